"""Pin the Drude-term oracle (oracle/drude.py):

* the Fermi-surface corner weights of one tetrahedron (tetrahedron.f:756-
  791) are the energy derivative of the exact integrals of the barycentric
  coordinates over the occupied part {e < E} (the clipped polytope of
  tests/test_oracle_tetra.py) in the two upper cross-section regimes; in
  the lowest (e1 < E < e2) the reference's area 3 f1 f2/(e3 - e1)
  (:779) is that derivative times (e4 - e1)/(e3 - e1) -- the exact one
  is 3 f1 f2/(e4 - e1) -- and the oracle keeps the reference's form;
* on a whole mesh, wintgr4 equals the same derivative assembled cube by
  cube from those exact integrals;
* degenerate bands get equal weights; the plasma frequency of a band with
  a constant momentum is the closed form 4 pi/(3 vol) * 2 |p|^2 * DOS; a
  degenerate pair with opposite momenta contributes nothing;
* drude, drudec and the head update are the closed forms.
"""
import math

import numpy as np
import pytest

from oracle import drude, tetra
from tests.test_oracle_tetra import clipped_corner_weights


def _dclip(et, E, h=1e-6):
    return (clipped_corner_weights(et, E + h) - clipped_corner_weights(et, E - h)) / (2 * h)


def _ref_factor(et, E):
    """The reference's lowest-regime area over the exact one (1 elsewhere)."""
    s = np.sort(et)
    return (s[3] - s[0]) / (s[2] - s[0]) if s[0] < E <= s[1] else 1.0


@pytest.mark.parametrize("seed", range(5))
def test_cross_section_is_derivative_of_clipped_integral(seed):
    g = np.random.default_rng(seed)
    et = g.uniform(-1, 1, 4)
    s = np.sort(et)
    for E in (0.3 * s[0] + 0.7 * s[1], 0.5 * (s[1] + s[2]), 0.2 * s[2] + 0.8 * s[3]):
        atria, wtria = drude._cross_section(et, E)
        got = (wtria * atria[None, :]).sum(1) / 3
        want = _dclip(et, E) / 4 * _ref_factor(et, E)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-8)


def _mesh_bands(n, seed):
    g = np.random.default_rng(seed)
    kx, ky, kz = np.meshgrid(*(np.arange(n) / n,) * 3, indexing="ij")
    k = np.stack([kx, ky, kz], -1).reshape(-1, 3)
    c = np.cos(2 * np.pi * k)
    e0 = -0.3 * c.sum(1) + 0.05 * g.standard_normal(len(k)) * 0.1
    e1 = 0.2 - 0.25 * (c[:, 0] * c[:, 1] + c[:, 2]) + 0.4
    return np.sort(np.stack([e0, e1, e1 + 2.0], 1), axis=1)


def test_mesh_weights_equal_assembled_derivative():
    n = 3
    rlat = np.eye(3)
    ene = _mesh_bands(n, 1)
    ef = 0.05
    W = drude.fermi_surface_weights((n, n, n), rlat, ene, ef)
    nk = n**3
    tet = tetra.get_tetra(rlat, (n, n, n))
    pk = tetra.mesh_index((n, n, n))
    want = np.zeros_like(W)
    for b in range(ene.shape[1]):
        for i1 in range(n):
            for i2 in range(n):
                for i3 in range(n):
                    kc = [pk(i1 + d1, i2 + d2, i3 + d3)
                          for d3 in (0, 1) for d2 in (0, 1) for d1 in (0, 1)]
                    for it in range(6):
                        et = ene[kc, b][tet[it]]
                        if et.min() >= ef or et.max() <= ef:
                            continue
                        d = _dclip(et, ef) / 4 / 6 / nk * _ref_factor(et, ef)
                        for i in range(4):
                            want[kc[tet[it][i]], b] += d[i]
    assert W.sum() > 0
    np.testing.assert_allclose(W, want, rtol=0, atol=1e-6 * W.max())


def test_degenerate_bands_share_weight():
    n = 2
    ene = _mesh_bands(n, 2)
    ene[:, 1] = ene[:, 0] + 1e-10  # exactly degenerate within 1e-8
    W = drude.fermi_surface_weights((n, n, n), np.eye(3), ene, float(np.median(ene[:, 0])))
    np.testing.assert_allclose(W[:, 0], W[:, 1], rtol=0, atol=1e-15)


def test_plasma_closed_forms():
    n = 3
    ene = _mesh_bands(n, 3)
    ef, vol = 0.05, 120.0
    W = drude.fermi_surface_weights((n, n, n), np.eye(3), ene, ef)
    nk, nb = W.shape
    p = np.array([0.3, -0.2, 0.1])
    mom = np.zeros((nk, nb, 3), complex)
    mom[:, 0, :] = p + 0.01j  # the imaginary part is dropped (real re array)
    deg = np.tile(np.arange(nb), (nk, 1))
    got = drude.plasma_frequency(W, mom, deg, vol)
    want = math.sqrt(4 * math.pi / 3 / vol * 2 * np.sum(p**2) * W[:, 0].sum())
    assert got == pytest.approx(want, rel=1e-13)
    # a degenerate pair with opposite momenta averages to zero
    mom2 = np.zeros((nk, nb, 3), complex)
    mom2[:, 0, :], mom2[:, 1, :] = p, -p
    deg2 = deg.copy()
    deg2[:, 0] = 1
    deg2[:, 1] = 1
    assert drude.plasma_frequency(W, mom2, deg2, vol) == 0.0


def test_plasma_spin_sum_closed_forms():
    """Two spins: plasma^2 is the sum of the spins' squares (constant momenta
    p_s on one band give 4 pi/(3 vol) sum_s |p_s|^2 DOS_s); two equal spins
    give the spin-unpolarised value (one spin times 2)."""
    n = 3
    ene = _mesh_bands(n, 3)
    ef, vol = 0.05, 120.0
    W = drude.fermi_surface_weights((n, n, n), np.eye(3), ene, ef)
    nk, nb = W.shape
    deg = np.tile(np.arange(nb), (nk, 1))
    p_up, p_dn = np.array([0.3, -0.2, 0.1]), np.array([0.05, 0.4, -0.25])
    mu = np.zeros((nk, nb, 3), complex)
    md = np.zeros((nk, nb, 3), complex)
    mu[:, 0, :], md[:, 0, :] = p_up, p_dn
    got = drude.plasma_frequency_spins([(W, mu, deg), (W, md, deg)], vol)
    want = math.sqrt(4 * math.pi / 3 / vol * (np.sum(p_up**2) + np.sum(p_dn**2)) * W[:, 0].sum())
    assert got == pytest.approx(want, rel=1e-13)
    same = drude.plasma_frequency_spins([(W, mu, deg), (W, mu, deg)], vol)
    assert same == pytest.approx(drude.plasma_frequency(W, mu, deg, vol), rel=1e-14)
    assert drude.plasma_frequency_spins([(W, mu, deg)], vol, 2.0) == \
        drude.plasma_frequency(W, mu, deg, vol)


def test_drude_functions_and_head():
    assert drude.drude(0.5, 0.0) == 0.0
    assert drude.drude(0.5, 2.0) == pytest.approx(-16.0)
    assert drude.drude(0.0, 2.0) == math.inf
    assert drude.drudec(0.5 + 0.5j, 2.0) == pytest.approx(4.0 / (0.5 + 0.5j) ** 2)
    head = np.arange(12, dtype=complex).reshape(6, 2)
    h = drude.head_minus_drude(head, [0.5, 2.0], 1.0)
    d = [-4.0, -0.25]
    for k in range(6):
        for w in range(2):
            want = head[k, w] - (d[w] if k in (0, 2, 5) else 0.0)
            assert h[k, w] == want
