"""Pin the oracle's Coulomb-eigenbasis transformation of the products
(susceptibility.f:806-832) by what the mathematics fixes: the pair sum of
the transformed products equals T^H chi0 T of the untransformed ones (the
reason the reference may transform either M or chi0), with theta = 1 the
overlap step is the identity, and olapvec of a column equals the Fourier
coefficients of theta times the column's function (a real-space check)."""
import numpy as np

import oracle
from oracle import trafo


def _rand(g, *shape):
    return g.standard_normal(shape) + 1j * g.standard_normal(shape)


def test_trafo_chi0_is_congruence():
    g = np.random.default_rng(1)
    n, nb, dim, nw = 300, 24, 17, 3
    M = _rand(g, n, nb) / np.sqrt(nb)
    T = _rand(g, nb, nb)
    W = -g.uniform(0.1, 1.0, (n, nw))
    Mt = trafo.coulomb_trafo(M, T, dim)
    assert Mt.shape == (n, dim)
    chi = oracle.chi0_pairsum(np.asfortranarray(M), np.asfortranarray(W))
    chit = oracle.chi0_pairsum(np.asfortranarray(Mt), np.asfortranarray(W))
    for w in range(nw):
        full = oracle.unpack_herm(chi[:, w], nb)
        want = T[:, :dim].conj().T @ full @ T[:, :dim]
        got = oracle.unpack_herm(chit[:, w], dim)
        np.testing.assert_allclose(got, want, atol=1e-12 * np.abs(want).max())


def test_ctrafo_olap_theta_one_is_identity():
    g = np.random.default_rng(2)
    nbasp, ng, dim = 5, 9, 4
    gptm = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                     [0, 0, -1], [1, 1, 0], [1, 0, 1]])
    T = _rand(g, nbasp + ng, dim)
    cs = np.zeros((5, 5, 5), complex)
    cs[2, 2, 2] = 1.0
    np.testing.assert_array_equal(trafo.ctrafo_olap(T, nbasp, gptm, cs, (2, 2, 2)), T)


def test_ctrafo_olap_is_theta_times_function():
    """olapvec(c)(G_i) = Fourier coefficient at G_i of conj(theta) times the
    function sum_j c(G_j) e^{i G_j r}, evaluated on an alias-free grid."""
    g = np.random.default_rng(3)
    r = np.arange(-1, 2)
    G = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    c = _rand(g, len(G), 1)
    sb = 1
    cs = _rand(g, 3, 3, 3) * 0.1
    cs[1, 1, 1] = 0.9
    ce = 2
    csp = np.zeros((5, 5, 5), complex)
    csp[ce - sb:ce + sb + 1, ce - sb:ce + sb + 1, ce - sb:ce + sb + 1] = cs
    got = trafo.ctrafo_olap(c, 0, G, csp, (ce, ce, ce))[:, 0]
    N = 8
    x = np.arange(N) / N
    X = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1).reshape(-1, 3)
    f = np.exp(2j * np.pi * X @ G.T) @ c[:, 0]
    th = np.zeros(len(X), complex)
    for z in range(-sb, sb + 1):
        for y in range(-sb, sb + 1):
            for xx in range(-sb, sb + 1):
                th += cs[z + sb, y + sb, xx + sb] * np.exp(2j * np.pi * X @ np.array([xx, y, z]))
    prod = np.conj(th) * f
    want = np.array([np.mean(prod * np.exp(-2j * np.pi * X @ Gi)) for Gi in G])
    np.testing.assert_allclose(got, want, atol=1e-13)
