"""GPU parity of the Gamma-point (q = 0) head and wings against the oracle,
through the C ABI: the momentum scaling (bit-exact), the head and wing sums
(imaginary and real frequencies, both layouts of M, separate head weights,
several frequency passes, accumulation), the dielectric-side head/wing
steps around eps^-1, and the host drop-ins."""
import math

import numpy as np
import pytest

import oracle
from oracle import gamma
from project1_amd import spexhip
from tests.conftest import normwise

pytestmark = pytest.mark.gpu


def _f(a, dt):
    return np.asfortranarray(a, dtype=dt)


def _dev_f(ctx, a, dt):
    """Column-major device copy of a 2-D (or n-D Fortran) array."""
    return ctx.to_dev(np.asfortranarray(a, dt).ravel(order="F"))


def _data(seed, n, nb):
    g = np.random.default_rng(seed)
    cp = (g.standard_normal((n, nb)) + 1j * g.standard_normal((n, nb))) / math.sqrt(nb)
    dp = g.standard_normal((n, 3)) + 1j * g.standard_normal((n, 3))
    return g, cp, dp


def _M_dev(ctx, cp, colmajor):
    """colmajor: the reference cprod(t + n*I); else transition-major t*nb + I."""
    return ctx.to_dev(np.asfortranarray(cp).ravel(order="F") if colmajor
                      else np.ascontiguousarray(cp))


def test_momentum_scale_bit_exact(ctx):
    g = np.random.default_rng(1)
    nk, nocc, nunocc = 5, 7, 11
    eo = -g.uniform(0.0, 1.0, (nk, nocc))
    eu = g.uniform(0.0, 1.0, (nk, nunocc))
    eu[1, 3] = eo[1, 2] + 5e-9  # below the threshold
    eu[2, 0] = eo[2, 6]         # degenerate
    n = nk * nocc * nunocc
    mom = g.standard_normal((n, 3)) + 1j * g.standard_normal((n, 3))
    vol = 271.3
    dP = _dev_f(ctx, mom, np.complex128)
    ctx.momentum_scale_dev(ctx.to_dev(eo), ctx.to_dev(eu), nk, nocc, nunocc, vol, dP, n)
    got = dP.download(np.complex128, (3, n)).T
    want = gamma.momentum_scale(mom, eo, eu, vol)
    assert np.array_equal(got, want)
    assert (got[(1 * nocc + 2) * nunocc + 3] == 0).all() and (got[(2 * nocc + 6) * nunocc] == 0).all()
    dP.free()


@pytest.mark.parametrize("n,nb,nw,colmajor,sep", [
    (3000, 40, 12, False, False),
    (3000, 40, 12, True, True),
    (2049, 254, 20, False, True),   # dcprod columns straddle two workgroups; two passes (16 + 4)
    (517, 300, 3, True, False),
    (5, 7, 1, False, False),        # fewer transitions than one stage
])
def test_headwings_vs_oracle(ctx, n, nb, nw, colmajor, sep):
    g, cp, dp = _data(2, n, nb)
    D = -g.uniform(0.05, 2.0, (n, nw))
    Dh = -g.uniform(0.05, 2.0, (n, nw)) if sep else None
    dM = _M_dev(ctx, cp, colmajor)
    dP = _dev_f(ctx, dp, np.complex128)
    dD = _dev_f(ctx, D, np.float64)
    dDh = _dev_f(ctx, Dh, np.float64) if sep else None
    dH = ctx.alloc(16 * 9 * nw)
    dW = ctx.alloc(16 * 3 * nb * nw)
    ctx.headwings_dev(n, nb, dM, n if colmajor else nb, colmajor, dP, n, nw, dD, n, dH, dW,
                      Dh=dDh, alpha=2.0)
    head = dH.download(np.complex128, (nw, 3, 3)).transpose(2, 1, 0)  # (k1, k2, w)
    wing = dW.download(np.complex128, (nw, nb, 3)).transpose(2, 1, 0)  # (k, i, w)
    hp, wp = gamma.headwings_imag(cp, dp, D, Dh)
    for w in range(nw):
        full = gamma.unpack_head(2 * hp[:, w])
        mx, _ = normwise(head[:, :, w], full)
        assert mx < 1e-12, ("head", w, mx)
        mx, _ = normwise(wing[:, :, w], 2 * wp[:, :, w])
        assert mx < 1e-12, ("wing", w, mx)
    # beta = 1 accumulates a second call
    ctx.headwings_dev(n, nb, dM, n if colmajor else nb, colmajor, dP, n, nw, dD, n, dH, dW,
                      Dh=dDh, alpha=1.0, beta=1.0)
    wing2 = dW.download(np.complex128, (nw, nb, 3)).transpose(2, 1, 0)
    mx, _ = normwise(wing2, 3 * wp)
    assert mx < 1e-12
    for b in (dM, dP, dD, dH, dW) + ((dDh,) if sep else ()):
        b.free()


@pytest.mark.parametrize("n,nb,nw,colmajor,sep", [
    (2500, 33, 7, False, True),     # passes of 6 + 1 frequencies
    (2500, 260, 4, True, False),
])
def test_headwings_cplx_vs_oracle(ctx, n, nb, nw, colmajor, sep):
    g, cp, dp = _data(3, n, nb)
    Dr, Di = g.standard_normal((n, nw)), g.standard_normal((n, nw))
    Dhr, Dhi = (g.standard_normal((n, nw)), g.standard_normal((n, nw))) if sep else (None, None)
    dM = _M_dev(ctx, cp, colmajor)
    dP = _dev_f(ctx, dp, np.complex128)
    bufs = [_dev_f(ctx, x, np.float64) for x in (Dr, Di)]
    hb = [_dev_f(ctx, x, np.float64) for x in (Dhr, Dhi)] if sep else [None, None]
    dH = ctx.alloc(16 * 9 * nw)
    dW = ctx.alloc(16 * 3 * nb * 2 * nw)
    ctx.headwings_cplx_dev(n, nb, dM, n if colmajor else nb, colmajor, dP, n, nw, bufs[0],
                           bufs[1], n, dH, dW, Dhr=hb[0], Dhi=hb[1])
    head = dH.download(np.complex128, (nw, 3, 3)).transpose(2, 1, 0)
    wing = dW.download(np.complex128, (nw, 2, nb, 3)).transpose(3, 2, 1, 0)  # (k, i, v, w)
    hw, ww = gamma.headwings_real(cp, dp, Dr, Di, Dhr, Dhi)
    for w in range(nw):
        mx, _ = normwise(head[:, :, w], hw[:, :, w])
        assert mx < 1e-12, ("head", w, mx)
        for v in range(2):
            mx, _ = normwise(wing[:, :, v, w], ww[:, :, v, w])
            assert mx < 1e-12, ("wing", v, w, mx)
    for b in [dM, dP, dH, dW] + bufs + [x for x in hb if x is not None]:
        b.free()


def test_headwings_are_the_augmented_chi0_at_size(ctx):
    """At a large packet (n = 200k): the head and wings equal the last
    columns of the FP64 chi0 of the augmented matrix [M | P] (a different
    kernel), a size-independent identity."""
    n, nb, nw = 200_000, 300, 12
    g = np.random.default_rng(4)
    dM = ctx.alloc(16 * n * (nb + 3))
    ctx.synth_normal(dM, 2 * n * (nb + 3), 5, 1.0 / math.sqrt(nb))
    # M (transition-major, ld nb + 3) is the first nb columns, P the last 3
    # (copied to its own column-major n x 3 array)
    Mall = dM.download(np.complex128, (n, nb + 3))
    dP = _dev_f(ctx, Mall[:, nb:], np.complex128)
    D = -g.uniform(0.05, 2.0, (n, nw))
    dD = _dev_f(ctx, D, np.float64)
    dH = ctx.alloc(16 * 9 * nw)
    dW = ctx.alloc(16 * 3 * nb * nw)
    ctx.headwings_dev(n, nb, dM, nb + 3, False, dP, n, nw, dD, n, dH, dW)
    dX = ctx.alloc(16 * (nb + 3) ** 2 * nw)
    ctx.chi0_dev(n, nb + 3, dM, nb + 3, False, nw, dD, spexhip.PREC_FP64, dX, nb + 3)
    X = dX.download(np.complex128, (nw, nb + 3, nb + 3)).transpose(2, 1, 0)  # (I, J, w)
    head = dH.download(np.complex128, (nw, 3, 3)).transpose(2, 1, 0)
    wing = dW.download(np.complex128, (nw, nb, 3)).transpose(2, 1, 0)
    for w in range(nw):
        up = np.triu(X[nb:, nb:, w])
        mx, _ = normwise(np.triu(head[:, :, w]), up)
        assert mx < 1e-12, ("head", w, mx)
        mx, _ = normwise(wing[:, :, w], X[:nb, nb:, w].T)
        assert mx < 1e-12, ("wing", w, mx)
    for b in (dM, dP, dD, dH, dW, dX):
        b.free()


def _gamma_case(seed, nb, nw):
    g = np.random.default_rng(seed)
    chis, heads, wings = [], [], []
    for _ in range(nw):
        A = (g.standard_normal((nb, nb + 5)) + 1j * g.standard_normal((nb, nb + 5))) / math.sqrt(nb)
        chis.append(oracle.pack_upper(-(A @ A.conj().T)))
        B = (g.standard_normal((3, 3)) + 1j * g.standard_normal((3, 3))) / 10
        heads.append(oracle.pack_upper(-(B @ B.conj().T)))
        wings.append((g.standard_normal((3, nb)) + 1j * g.standard_normal((3, nb))) / 30)
    sq = g.uniform(0.3, 1.5, nb)
    return np.stack(chis, 1), np.stack(heads, 1), np.stack(wings, 2), sq


def test_gamma_dielectric_chain_vs_oracle(ctx):
    nb, nw = 150, 4
    chi, head, wing, sq = _gamma_case(6, nb, nw)
    eps_o, h_o, w_o = gamma.gamma_prepare(chi, head, wing, sq)
    einv_o = np.stack([oracle.inverse_hpd_packed(eps_o[:, f]) for f in range(nw)], 1)
    h_o, w_o = gamma.gamma_inverse(einv_o, h_o, w_o)
    # device chain: unpack -> prepare -> dielectric -> inverse -> head/wings
    dX = ctx.alloc(16 * nb * nb * nw)
    ctx.unpack(np.ascontiguousarray(chi.T), nb, nw, dX, nb)
    hfull = np.stack([gamma.unpack_head(head[:, f]) for f in range(nw)], 2)
    dH = _dev_f(ctx, hfull, np.complex128)
    dW = _dev_f(ctx, wing, np.complex128)
    dS = ctx.to_dev(sq)
    ctx.gamma_prepare_dev(nb, nw, dX, nb, dH, dW, dS)
    ctx.dielectric_dev(dX, nb, nw, nb, 0, dS, dX, nb)
    eps = ctx.pack_upper(dX, nb, nw, nb).T
    mx, _ = normwise(eps, eps_o)
    assert mx < 1e-14, ("eps", mx)
    ctx.inverse_hpd_dev(dX, nb, nw, nb)
    ctx.gamma_inverse_dev(nb, nw, dX, nb, dH, dW)
    einv = ctx.pack_upper(dX, nb, nw, nb).T
    h = dH.download(np.complex128, (nw, 3, 3)).transpose(2, 1, 0)
    w = dW.download(np.complex128, (nw, nb, 3)).transpose(2, 1, 0)
    mx, _ = normwise(einv, einv_o)
    assert mx < 1e-12, ("einv", mx)
    for f in range(nw):
        mx, _ = normwise(h[:, :, f], gamma.unpack_head(h_o[:, f]))
        assert mx < 1e-12, ("head", f, mx)
        mx, _ = normwise(w[:, :, f], w_o[:, :, f])
        assert mx < 1e-12, ("wing", f, mx)
    for b in (dX, dH, dW, dS):
        b.free()


def test_dielectric_inverse_gamma_host_dropin(ctx):
    nb, nw = 97, 3
    chi, head, wing, sq = _gamma_case(7, nb, nw)
    eps_o, h_o, w_o = gamma.gamma_prepare(chi, head, wing, sq)
    einv_o = np.stack([oracle.inverse_hpd_packed(eps_o[:, f]) for f in range(nw)], 1)
    h_o, w_o = gamma.gamma_inverse(einv_o, h_o, w_o)
    h = _f(head, np.complex128)
    w = _f(wing, np.complex128)
    einv = ctx.dielectric_inverse_gamma(_f(chi, np.complex128), sq, h, w)
    assert normwise(einv, einv_o)[0] < 1e-12
    assert normwise(h, h_o)[0] < 1e-12
    assert normwise(w, w_o)[0] < 1e-12


def test_susceptibility_headwings_host_dropin_accumulates(ctx):
    n, nb, nfh, nfc = 900, 21, 3, 2
    g, cp, dp = _data(8, n, nb)
    wh = -g.uniform(0.1, 1.0, (n, nfh))
    whh = -g.uniform(0.1, 1.0, (n, nfh))
    wcr, wci = g.standard_normal((n, nfc)), g.standard_normal((n, nfc))
    headh = np.zeros((6, nfh), np.complex128, order="F")
    wingh = np.zeros((3, nb, nfh), np.complex128, order="F")
    headc = np.zeros((3, 3, nfc), np.complex128, order="F")
    wingc = np.zeros((3, nb, 2, nfc), np.complex128, order="F")
    for _ in range(2):
        ctx.susceptibility_headwings(cp, dp, wh, headh, wingh, wcr, wci, headc, wingc, whh=whh)
    hp, wp = gamma.headwings_imag(cp, dp, wh, whh)
    hc, wc = gamma.headwings_real(cp, dp, wcr, wci)
    assert normwise(headh, 2 * hp)[0] < 1e-12
    assert normwise(wingh, 2 * wp)[0] < 1e-12
    assert normwise(headc, 2 * hc)[0] < 1e-12
    assert normwise(wingc, 2 * wc)[0] < 1e-12


def test_headwings_rejects_bad_shapes(ctx):
    dM = ctx.alloc(16 * 64 * 8)
    dP = ctx.alloc(16 * 64 * 3)
    dD = ctx.alloc(8 * 64)
    dH = ctx.alloc(16 * 9)
    dW = ctx.alloc(16 * 3 * 8)
    with pytest.raises(spexhip.SpexError):  # wing buffer too small for nb = 9
        ctx.headwings_dev(64, 9, dM, 9, False, dP, 64, 1, dD, 64, dH, dW)
    with pytest.raises(spexhip.SpexError):  # ldp < n
        ctx.headwings_dev(64, 8, dM, 8, False, dP, 32, 1, dD, 64, dH, dW)
    for b in (dM, dP, dD, dH, dW):
        b.free()


def test_head_scalar_and_gamma_rpa(ctx):
    """eps(1,1) = trace(head)/3 or v^T head v (correlation.f:656-665, :683),
    then the RPA integrand of that eps against the oracle."""
    nb, nw = 40, 3
    chi, head, wing, sq = _gamma_case(9, nb, nw)
    eps_o, h_o, _ = gamma.gamma_prepare(chi, head, wing, sq)
    dX = ctx.alloc(16 * nb * nb * nw)
    ctx.unpack(np.ascontiguousarray(chi.T), nb, nw, dX, nb)
    hfull = np.stack([gamma.unpack_head(head[:, f]) for f in range(nw)], 2)
    dH = _dev_f(ctx, hfull, np.complex128)
    dW = _dev_f(ctx, wing, np.complex128)
    dS = ctx.to_dev(sq)
    ctx.gamma_prepare_dev(nb, nw, dX, nb, dH, dW, dS)
    ctx.dielectric_dev(dX, nb, nw, nb, 0, dS, dX, nb)
    vec = np.array([0.6, 0.0, 0.8])
    for v in (None, vec):
        ctx.gamma_head_scalar_dev(nb, nw, dX, nb, dH, v)
        got = ctx.pack_upper(dX, nb, nw, nb).T
        for f in range(nw):
            want = gamma.head_scalar(h_o[:, f], v)
            assert abs(got[0, f] - want) <= 1e-13 * abs(want), (v, f)
    # RPA at Gamma with the angular average in eps(1,1)
    ctx.gamma_head_scalar_dev(nb, nw, dX, nb, dH, None)
    got = ctx.rpa_integrand_dev(dX, nb, nw, nb, 0.25)
    for f in range(nw):
        e = eps_o[:, f].copy()
        e[0] = gamma.head_scalar(h_o[:, f]).real
        want = oracle.rpa_integrand(e, 0.25)
        assert abs(got[f] - want) <= 1e-10 * abs(want), (f, got[f], want)
    for b in (dX, dH, dW, dS):
        b.free()


def _absorptive_chi(g, nb, nw):
    """A real-frequency-like chi: Hermitian part of either sign, anti-
    Hermitian part -i C C^H (absorption), so eps has a non-negative
    imaginary part."""
    out = np.empty((nb, nb, nw), complex)
    for f in range(nw):
        A = (g.standard_normal((nb, nb)) + 1j * g.standard_normal((nb, nb))) / math.sqrt(nb)
        Cm = (g.standard_normal((nb, nb)) + 1j * g.standard_normal((nb, nb))) / math.sqrt(nb)
        out[:, :, f] = 0.3 * (A + A.conj().T) / 2 - 0.5j * (Cm @ Cm.conj().T)
    return out


@pytest.mark.parametrize("vkind", [0, 1])
def test_general_dielectric_and_inverse_vs_numpy(ctx, vkind):
    nb, nw = 130, 3
    g = np.random.default_rng(30 + vkind)
    chi = _absorptive_chi(g, nb, nw)
    if vkind == 0:
        vh = g.uniform(0.3, 1.2, nb)
        V = np.diag(vh).astype(complex)
        dV = ctx.to_dev(vh)
    else:
        q, _ = np.linalg.qr(g.standard_normal((nb, nb)))
        V = (q * g.uniform(0.3, 1.2, nb)) @ q.T
        V = V.astype(complex)
        dV = _dev_f(ctx, V, np.complex128)
    dX = _dev_f(ctx, chi, np.complex128)
    dE = ctx.alloc(16 * nb * nb * nw)
    ctx.dielectric_general_dev(dX, nb, nw, nb, vkind, dV, dE, nb)
    eps = dE.download(np.complex128, (nw, nb, nb)).transpose(2, 1, 0)
    want_eps = np.stack([np.eye(nb) - V @ chi[:, :, f] @ V for f in range(nw)], 2)
    assert normwise(eps, want_eps)[0] < 1e-13
    ctx.inverse_general_dev(dE, nb, nw, nb)
    inv = dE.download(np.complex128, (nw, nb, nb)).transpose(2, 1, 0)
    for f in range(nw):
        mx, _ = normwise(inv[:, :, f], np.linalg.inv(want_eps[:, :, f]))
        assert mx < 1e-12, (f, mx)
    for b in (dV, dX, dE):
        b.free()


def test_general_inverse_pivots_like_zgetrf(ctx):
    """Partial pivoting (wrapper.f:2357-2358 calls zgetrf/zgetri): an
    invertible matrix with a zero leading pivot, and well-conditioned
    matrices whose rows are scrambled across the 64-row panels so that the
    natural pivots are tiny (1e-12) or zero, invert to numpy's inverse."""
    nb = 70
    X = np.eye(nb, dtype=complex)
    X[0, 0] = 0.0
    X[0, 1] = X[1, 0] = 1.0  # invertible, but the first pivot is zero
    dX = _dev_f(ctx, X, np.complex128)
    ctx.inverse_general_dev(dX, nb, 1, nb)
    got = dX.download(np.complex128, (nb, nb)).T
    np.testing.assert_allclose(got, np.linalg.inv(X), rtol=0, atol=1e-15)
    dX.free()
    nb, nw = 200, 3
    g = np.random.default_rng(31)
    mats = []
    for f in range(nw):
        B = np.eye(nb) + 0.3 * (g.standard_normal((nb, nb)) + 1j * g.standard_normal((nb, nb))) / np.sqrt(nb)
        A = B[g.permutation(nb)]  # same condition number, pivots scattered over the panels
        A[0, 0] = 1e-12 if f == 0 else 0.0
        A[64, 64] = 1e-13  # the first pivot of the second panel, before elimination
        mats.append(A)
    A = np.stack(mats, 2)
    dX = _dev_f(ctx, A, np.complex128)
    ctx.inverse_general_dev(dX, nb, nw, nb)
    inv = dX.download(np.complex128, (nw, nb, nb)).transpose(2, 1, 0)
    for f in range(nw):
        want = np.linalg.inv(A[:, :, f])
        mx, fro = normwise(inv[:, :, f], want)
        print(f"pivoted inverse f={f}: cond {np.linalg.cond(A[:, :, f]):.1f} max-rel {mx:.2e}")
        assert mx < 1e-12, (f, mx)
    dX.free()


def test_general_inverse_rejects_a_singular_matrix(ctx):
    nb = 90
    g = np.random.default_rng(32)
    X = g.standard_normal((nb, nb)) + 1j * g.standard_normal((nb, nb))
    X[:, 70] = 0.0  # no pivot exists in column 70
    dX = _dev_f(ctx, X, np.complex128)
    with pytest.raises(spexhip.SpexError, match="pivot"):
        ctx.inverse_general_dev(dX, nb, 1, nb)
    dX.free()


def test_gamma_real_frequency_chain_vs_oracle(ctx):
    nb, nw = 96, 2
    g = np.random.default_rng(40)
    chi = _absorptive_chi(g, nb, nw)
    sq = g.uniform(0.3, 1.2, nb)
    head = (g.standard_normal((3, 3, nw)) + 1j * g.standard_normal((3, 3, nw))) / 20
    wing = (g.standard_normal((3, nb, 2, nw)) + 1j * g.standard_normal((3, nb, 2, nw))) / 30
    eps_o, h_o, w_o = gamma.gamma_prepare_cplx(chi, head, wing, sq)
    einv_o = np.stack([np.linalg.inv(eps_o[:, :, f]) for f in range(nw)], 2)
    h_o, w_o = gamma.gamma_inverse_cplx(einv_o, h_o, w_o)
    dX = _dev_f(ctx, chi, np.complex128)
    dH = _dev_f(ctx, head, np.complex128)
    dW = _dev_f(ctx, wing, np.complex128)
    dS = ctx.to_dev(sq)
    ctx.gamma_prepare_cplx_dev(nb, nw, dX, nb, dH, dW, dS)
    ctx.dielectric_general_dev(dX, nb, nw, nb, 0, dS, dX, nb)
    ctx.inverse_general_dev(dX, nb, nw, nb)
    ctx.gamma_inverse_cplx_dev(nb, nw, dX, nb, dH, dW)
    einv = dX.download(np.complex128, (nw, nb, nb)).transpose(2, 1, 0)
    h = dH.download(np.complex128, (nw, 3, 3)).transpose(2, 1, 0)
    w = dW.download(np.complex128, (nw, 2, nb, 3)).transpose(3, 2, 1, 0)
    assert normwise(einv, einv_o)[0] < 1e-12
    assert normwise(h, h_o)[0] < 1e-12
    assert normwise(w, w_o)[0] < 1e-12
    for b in (dX, dH, dW, dS):
        b.free()


def test_real_frequency_host_dropin_vs_numpy(ctx):
    nb, nw = 64, 2
    g = np.random.default_rng(50)
    chi = _absorptive_chi(g, nb, nw)
    sq = g.uniform(0.3, 1.2, nb)
    got = ctx.dielectric_inverse_general(chi, sq)
    for f in range(nw):
        want = np.linalg.inv(np.eye(nb) - sq[:, None] * chi[:, :, f] * sq[None, :])
        assert normwise(got[:, :, f], want)[0] < 1e-12
