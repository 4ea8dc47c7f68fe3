"""GPU parity of the Coulomb-eigenbasis transformation of the products
(susceptibility.f:806-832) against the oracle, in place and out of place,
with ctrafo1 formed on the device (overlap.f:298-315), and the chi0 of the
transformed products against T^H chi0 T."""
import numpy as np
import pytest

import oracle
from oracle import fftgrid, trafo
from project1_amd import spexhip, synth
from tests.conftest import normwise

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _rand(g, *shape):
    return g.standard_normal(shape) + 1j * g.standard_normal(shape)


def _dev_f(ctx, a):
    return ctx.to_dev(np.asfortranarray(a).ravel(order="F"))


@pytest.mark.parametrize("n,nbasm,dim,inplace", [(1000, 70, 70, True), (3000, 150, 97, False),
                                                 (20000, 150, 97, True)])
def test_coulomb_trafo_vs_oracle(ctx, n, nbasm, dim, inplace):
    g = np.random.default_rng(n + dim)
    M = _rand(g, n, nbasm)
    T = _rand(g, nbasm, nbasm) / np.sqrt(nbasm)
    ldm, ldt = nbasm + 2, nbasm + 1
    Mp = np.zeros((n, ldm), complex)
    Mp[:, :nbasm] = M
    dM = ctx.to_dev(Mp)
    Tp = np.zeros((ldt, nbasm), complex)
    Tp[:nbasm] = T
    dT = _dev_f(ctx, Tp)
    want = trafo.coulomb_trafo(M, T, dim)
    if inplace:
        ctx.coulomb_trafo_dev(n, nbasm, dM, ldm, dT, ldt, dim)
        got = dM.download(np.complex128, (n, ldm))
        np.testing.assert_array_equal(got[:, dim:], Mp[:, dim:])  # columns past dim untouched
        got = got[:, :dim]
    else:
        ldo = dim + 3
        dO = ctx.alloc(16 * n * ldo).zero()
        ctx.coulomb_trafo_dev(n, nbasm, dM, ldm, dT, ldt, dim, dO, ldo)
        got = dO.download(np.complex128, (n, ldo))[:, :dim]
        np.testing.assert_array_equal(dM.download(np.complex128, (n, ldm)), Mp)  # input kept
        dO.free()
    mx, fro = normwise(got, want)
    print(f"coulomb trafo n={n} nbasm={nbasm} dim={dim}: max-rel {mx:.2e} fro-rel {fro:.2e}")
    assert mx < TOL
    dM.free()
    dT.free()


def test_ctrafo_olap_vs_oracle(ctx):
    s = synth.product_system(nmesh=2, nband=2, gcut=3.1, ngptm=300, lcut=2, nrad=2, lcutm=2,
                             nradm=2, seed=80)
    nbasp, ngptm = s.nbasp, s.ngptm
    nbasm, dim = nbasp + ngptm, 211
    g = np.random.default_rng(81)
    T = _rand(g, nbasm, dim)
    ext = s.gptm.max(0) - s.gptm.min(0)
    cs = fftgrid.stepfunction_coeffs(np.eye(3) * 1.1, s.vol, [1.0, 0.8], s.cent, ext)
    dT = _dev_f(ctx, T)
    dT1 = ctx.alloc(16 * nbasm * dim)
    ctx.ctrafo_olap_dev(nbasp, s.gptm, dim, dT, nbasm, cs, dT1, nbasm)
    got = dT1.download(np.complex128, (dim, nbasm)).T
    want = trafo.ctrafo_olap(T, nbasp, s.gptm, cs, ext)
    np.testing.assert_array_equal(got[:nbasp], T[:nbasp])
    mx, fro = normwise(got, want)
    print(f"ctrafo_olap ngptm={ngptm} dim={dim}: max-rel {mx:.2e} fro-rel {fro:.2e}")
    assert mx < TOL
    with pytest.raises(spexhip.SpexError):  # cstep too small for the G differences
        ctx.ctrafo_olap_dev(nbasp, s.gptm, dim, dT, nbasm, np.ones((3, 3, 3), complex), dT1, nbasm)
    dT.free()
    dT1.free()


def test_chi0_in_coulomb_basis_is_congruence(ctx):
    """chi0 of the transformed products (device) equals T^H chi0 T of the
    untransformed ones (oracle), at the single-q basis size Nb = 1200 with
    dim = 1000 eigenvectors kept, over 4096 transitions."""
    g = np.random.default_rng(82)
    n, nbasm, dim, nw = 4096, 1200, 1000, 2
    M = _rand(g, n, nbasm) / np.sqrt(nbasm)
    Q, _ = np.linalg.qr(_rand(g, nbasm, nbasm))
    T = Q[:, :dim]
    W = -g.uniform(0.05, 1.0, (n, nw))
    dM = ctx.to_dev(np.ascontiguousarray(M))
    dT = _dev_f(ctx, T)
    ctx.coulomb_trafo_dev(n, nbasm, dM, nbasm, dT, nbasm, dim)
    dW = _dev_f(ctx, W)
    dX = ctx.alloc(16 * dim * dim * nw)
    ctx.chi0_dev(n, dim, dM, nbasm, False, nw, dW, spexhip.PREC_FP64, dX, dim)
    got = ctx.pack_upper(dX, dim, nw, dim)
    chi = oracle.chi0_pairsum(np.asfortranarray(M), np.asfortranarray(W))
    for w in range(nw):
        want = T.conj().T @ oracle.unpack_herm(chi[:, w], nbasm) @ T
        mx, fro = normwise(oracle.unpack_herm(got[w], dim), want)
        print(f"w{w}: chi0 in the Coulomb basis vs T^H chi0 T: max-rel {mx:.2e} fro-rel {fro:.2e}")
        assert mx < 1e-11
    for b in (dM, dT, dW, dX):
        b.free()
