"""Pinning of the Gamma-point head/wing oracle (oracle/gamma.py): the head
and wing sums are entries of the chi0 pair sum of the augmented matrix
[cprod | dcprod] (an independent, golden-pinned oracle); the reference's
head/wing update of the inverse is the Schur complement of the bordered
matrix; the scaling and dielectric steps have closed forms."""
import math

import numpy as np
import pytest

import oracle
from oracle import gamma, hilbert


def _rand(g, n, nb):
    cp = (g.standard_normal((n, nb)) + 1j * g.standard_normal((n, nb))) / math.sqrt(nb)
    dp = g.standard_normal((n, 3)) + 1j * g.standard_normal((n, 3))
    return cp, dp


def _packed_index(i, j):
    """zero-based packed-upper index of (i, j), i <= j."""
    return j * (j + 1) // 2 + i


def test_headwings_are_entries_of_the_augmented_pair_sum():
    g = np.random.default_rng(11)
    n, nb, nw = 300, 9, 3
    cp, dp = _rand(g, n, nb)
    D = -g.uniform(0.1, 2.0, (n, nw))
    head, wing = gamma.headwings_imag(cp, dp, D)
    aug = oracle.chi0_pairsum(np.concatenate([cp, dp], axis=1), D)  # (npack, nw)
    for k2 in range(3):
        for k1 in range(k2 + 1):
            want = aug[_packed_index(nb + k1, nb + k2)]
            np.testing.assert_allclose(head[_packed_index(k1, k2)], want, rtol=1e-12, atol=1e-12)
        for i in range(nb):
            np.testing.assert_allclose(wing[k2, i], aug[_packed_index(i, nb + k2)], rtol=1e-12,
                                       atol=1e-12)


def test_separate_head_weights_leave_the_wings_alone():
    g = np.random.default_rng(12)
    n, nb, nw = 120, 4, 2
    cp, dp = _rand(g, n, nb)
    D = g.standard_normal((n, nw))
    Dh = g.standard_normal((n, nw))
    h1, w1 = gamma.headwings_imag(cp, dp, D, Dh)
    h2, _ = gamma.headwings_imag(cp, dp, Dh)
    _, w2 = gamma.headwings_imag(cp, dp, D)
    np.testing.assert_array_equal(h1, h2)
    np.testing.assert_array_equal(w1, w2)


def test_real_frequency_headwings_are_the_complex_weighted_augmented_matrix():
    g = np.random.default_rng(13)
    n, nb, nw = 200, 6, 2
    cp, dp = _rand(g, n, nb)
    Dr, Di = g.standard_normal((n, nw)), g.standard_normal((n, nw))
    head, wing = gamma.headwings_real(cp, dp, Dr, Di)
    C = hilbert.chi0_complex_weights(np.concatenate([cp, dp], axis=1), Dr, Di)
    np.testing.assert_allclose(head, C[nb:, nb:], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(wing[:, :, 0], C[:nb, nb:].transpose(1, 0, 2), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(wing[:, :, 1], C[nb:, :nb], rtol=1e-12, atol=1e-12)


def test_imaginary_head_is_hermitian_and_real_weights_give_conjugate_wings():
    """With Di = 0 the second real-frequency wing is the conjugate of the
    first, and the imaginary-axis head equals the packed upper of the full
    real-frequency head."""
    g = np.random.default_rng(14)
    n, nb, nw = 150, 5, 2
    cp, dp = _rand(g, n, nb)
    D = g.standard_normal((n, nw))
    hc, wc = gamma.headwings_real(cp, dp, D, np.zeros_like(D))
    hh, wh = gamma.headwings_imag(cp, dp, D)
    np.testing.assert_allclose(wc[:, :, 1], np.conj(wc[:, :, 0]), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(wc[:, :, 0], wh, rtol=1e-13, atol=1e-13)
    for f in range(nw):
        np.testing.assert_allclose(gamma.unpack_head(hh[:, f]), hc[:, :, f], rtol=1e-13, atol=1e-13)


def test_momentum_scale_closed_form_and_zero_rule():
    nk, nocc, nunocc = 2, 2, 3
    eo = np.array([[-1.0, -0.5], [-0.25, -0.125]])
    eu = np.array([[0.5, 1.0, -0.5 + 1e-8], [2.0, -0.125 - 5e-9, 0.75]])
    n = nk * nocc * nunocc
    mom = (np.arange(n * 3) + 1.0).reshape(n, 3) * (1 - 2j)
    vol = 4.0
    got = gamma.momentum_scale(mom, eo, eu, vol)
    k, b1, b2 = gamma.transitions(nk, nocc, nunocc)
    assert list(zip(k[:4], b1[:4], b2[:4])) == [(0, 0, 0), (0, 0, 1), (0, 0, 2), (0, 1, 0)]
    for t in range(n):
        d = eu[k[t], b2[t]] - eo[k[t], b1[t]]
        if abs(d) > 1e-8:
            np.testing.assert_allclose(got[t], mom[t] / 2.0 / d, rtol=1e-15)
        else:
            assert (got[t] == 0).all(), t
    # the (0,1,2) and (1,1,1) differences are exactly at / below the threshold
    assert (got[(0 * nocc + 1) * nunocc + 2] == 0).all()
    assert (got[(1 * nocc + 1) * nunocc + 1] == 0).all()


def test_gamma_prepare_closed_form():
    g = np.random.default_rng(15)
    nb, nw = 6, 2
    A = g.standard_normal((nb, nb)) + 1j * g.standard_normal((nb, nb))
    chi = -(A @ A.conj().T)
    chi_p = np.stack([oracle.pack_upper(chi), oracle.pack_upper(2 * chi)], axis=1)
    sq = g.uniform(0.5, 2.0, nb)
    head = g.standard_normal((6, nw)) + 1j * g.standard_normal((6, nw))
    wing = g.standard_normal((3, nb, nw)) + 1j * g.standard_normal((3, nb, nw))
    eps, h, w = gamma.gamma_prepare(chi_p, head, wing, sq)
    for f, s in enumerate((1, 2)):
        c = s * chi.copy()
        c[0, :] = 0
        c[:, 0] = 0
        want = np.eye(nb) - sq[:, None] * c * sq[None, :]
        np.testing.assert_allclose(oracle.unpack_herm(eps[:, f], nb), want, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(h[[0, 2, 5]], 1 - 4 * math.pi * head[[0, 2, 5]], rtol=1e-15)
    np.testing.assert_allclose(h[[1, 3, 4]], -4 * math.pi * head[[1, 3, 4]], rtol=1e-15)
    assert (w[:, 0, :] == 0).all()
    np.testing.assert_allclose(w[:, 1:], -math.sqrt(4 * math.pi) * sq[None, 1:, None] * wing[:, 1:],
                               rtol=1e-15)


@pytest.mark.parametrize("seed", [16, 17])
def test_gamma_inverse_is_the_bordered_matrix_inverse(seed):
    """The reference keeps S = H - conj(w) B^-1 w^T as the head and
    -w conj(B^-1) as the wings: the Schur complement and the scaled
    off-diagonal blocks of the Hermitian bordered matrix
    E = [[H, conj(w)], [w^T, B]], whose inverse has S^-1 as its head block,
    w_new^T S^-1 below and S^-1 conj(w_new) beside it."""
    g = np.random.default_rng(seed)
    nb, nw = 8, 2
    heads, wings, einvs, Es = [], [], [], []
    for _ in range(nw):
        A = g.standard_normal((nb + 3, nb + 3)) + 1j * g.standard_normal((nb + 3, nb + 3))
        E = A @ A.conj().T + (nb + 3) * np.eye(nb + 3)
        H, X, B = E[:3, :3], E[:3, 3:], E[3:, 3:]
        heads.append(oracle.pack_upper(H))
        wings.append(np.conj(X))
        einvs.append(oracle.pack_upper(np.linalg.inv(B)))
        Es.append(E)
    h, w = gamma.gamma_inverse(np.stack(einvs, 1), np.stack(heads, 1), np.stack(wings, 2))
    for f in range(nw):
        Ei = np.linalg.inv(Es[f])
        S = gamma.unpack_head(h[:, f])
        Si = np.linalg.inv(S)
        np.testing.assert_allclose(Si, Ei[:3, :3], rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(w[:, :, f].T @ Si, Ei[3:, :3], rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(Si @ np.conj(w[:, :, f]), Ei[:3, 3:], rtol=1e-11, atol=1e-13)


def test_multdiff_spectral_function_is_a_fold_with_scaled_hilbert_weights():
    """spectral_function with mm = 1, 2 equals the per-transition fold with
    wghtS(s, :) / frqs(s)**mm (s with frqs(s) = 0 left undivided)."""
    g = np.random.default_rng(18)
    n, nfrqs, nw = 80, 12, 3
    frqs = np.concatenate([[0.0], np.sort(g.uniform(0.05, 3.0, nfrqs - 1))])
    wt = g.uniform(0, 1, (n, nfrqs))
    wt[g.uniform(0, 1, wt.shape) < 0.2] = 1e-9
    wS = g.standard_normal((nfrqs, nw))
    cp, dp = _rand(g, n, 2)
    c = cp[:, 0] * np.conj(dp[:, 1])
    for mm in (0, 1, 2):
        scale = np.where(frqs > 0, frqs, 1.0) ** mm
        D = hilbert.fold_weights(wt, wS / scale[:, None], 1e-6)
        want = c.real @ D + 1j * (c.imag @ D)
        got = gamma.spectral_function_mm(c, wt, frqs, wS, mm, 1e-6)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def _multdiff_on_mesh(g, nk, nocc, nunocc, nfrqs):
    """Energies whose every transition difference is a positive mesh point
    frqs(s_t), and tetrahedron weights nonzero only at that s_t."""
    frqs = np.concatenate([[0.0], np.cumsum(g.uniform(0.05, 0.4, nfrqs - 1))])
    c = g.uniform(-1.0, 0.0, nk)
    s_of = g.integers(1, nfrqs, (nk, nunocc))
    eo = np.repeat(c[:, None], nocc, axis=1)
    eu = c[:, None] + frqs[s_of]
    n = nk * nocc * nunocc
    k, b1, b2 = gamma.transitions(nk, nocc, nunocc)
    wt = np.zeros((n, nfrqs))
    wt[np.arange(n), s_of[k, b2]] = g.uniform(0.1, 1.0, n)
    return frqs, eo, eu, wt


def test_multdiff_cancels_when_weight_sits_at_the_transition_energy():
    """D_l = D_0 / Delta^l for every transition, so the scaled column times the
    divided weights gives back the plain pair sum, on both axes."""
    g = np.random.default_rng(21)
    nk, nocc, nunocc, nfrqs, nb, nw = 3, 2, 4, 9, 5, 4
    frqs, eo, eu, wt = _multdiff_on_mesh(g, nk, nocc, nunocc, nfrqs)
    n = nk * nocc * nunocc
    cp, _ = _rand(g, n, nb)
    wSr = g.standard_normal((nfrqs, nw))
    wSi = g.standard_normal((nfrqs, nw))
    plain = hilbert.chi0_hilbert_pairsum(cp, wt, wSr)
    md = gamma.chi0_multdiff(cp, eo, eu, wt, frqs, wSr)
    np.testing.assert_allclose(md, plain, rtol=1e-12, atol=1e-12)
    Dr = hilbert.fold_weights(wt, wSr)
    Di = hilbert.fold_weights(wt, wSi)
    plainc = hilbert.chi0_complex_weights(cp, Dr, Di)
    mdc = gamma.chi0_multdiff(cp, eo, eu, wt, frqs, wSr, wSi)
    np.testing.assert_allclose(mdc, plainc, rtol=1e-12, atol=1e-12)
    # the column is scaled by e_unocc(k+q) - e_occ(k) (not its inverse or
    # negative) and only the first column changes
    sc = gamma.multdiff_scale(cp, eo, eu)
    kk, bb1, bb2 = gamma.transitions(nk, nocc, nunocc)
    np.testing.assert_array_equal(sc[:, 1:], cp[:, 1:])
    np.testing.assert_allclose(sc[:, 0], cp[:, 0] * (eu[kk, bb2] - eo[kk, bb1]), rtol=1e-15)


def test_multdiff_with_weight_off_the_transition_energy_differs():
    """Moving the weight one mesh point away breaks the cancellation in row
    and column 0 only: the body (i, j >= 1) is untouched."""
    g = np.random.default_rng(22)
    nk, nocc, nunocc, nfrqs, nb, nw = 2, 2, 3, 8, 4, 3
    frqs, eo, eu, wt = _multdiff_on_mesh(g, nk, nocc, nunocc, nfrqs)
    wt = np.roll(wt, 1, axis=1)
    wt[:, 0] = 0.0
    n = nk * nocc * nunocc
    cp, _ = _rand(g, n, nb)
    wS = g.standard_normal((nfrqs, nw))
    plain = hilbert.chi0_hilbert_pairsum(cp, wt, wS)
    md = gamma.chi0_multdiff(cp, eo, eu, wt, frqs, wS)
    np.testing.assert_allclose(md[1:, 1:], plain[1:, 1:], rtol=1e-12, atol=1e-12)
    assert np.abs(md[0] - plain[0]).max() > 1e-3 * np.abs(plain[0]).max()


def test_multdiff_inton_pairs_and_weights():
    """INTON: the body (i, j >= 1) is the plain pair sum with the md = 0
    weights; row and column 0 stay empty except (0,0), which is the sum of
    the scaled column with the md = 1 plus md = 2 weights (linearity); on
    both axes."""
    g = np.random.default_rng(23)
    nk, nocc, nunocc, nfrqs, nb, nw = 2, 3, 4, 10, 5, 3
    frqs, eo, eu, wt = _multdiff_on_mesh(g, nk, nocc, nunocc, nfrqs)
    n = nk * nocc * nunocc
    cp, _ = _rand(g, n, nb)
    wS = [g.standard_normal((nfrqs, nw)) for _ in range(3)]
    wSi = [g.standard_normal((nfrqs, nw)) for _ in range(3)]
    got = gamma.chi0_multdiff_inton(cp, eo, eu, wt, wS)
    body = hilbert.chi0_hilbert_pairsum(cp, wt, wS[0])
    np.testing.assert_allclose(got[1:, 1:], body[1:, 1:], rtol=1e-12, atol=1e-12)
    assert not got[0, 1:].any() and not got[1:, 0].any()
    sc = gamma.multdiff_scale(cp, eo, eu)
    h = hilbert.chi0_hilbert_pairsum(sc[:, :1], wt, wS[1] + wS[2])
    np.testing.assert_allclose(got[0, 0], h[0, 0], rtol=1e-12)
    gotc = gamma.chi0_multdiff_inton(cp, eo, eu, wt, wS, wSi)
    Dr = [hilbert.fold_weights(wt, w) for w in wS]
    Di = [hilbert.fold_weights(wt, w) for w in wSi]
    bodyc = hilbert.chi0_complex_weights(cp, Dr[0], Di[0])
    np.testing.assert_allclose(gotc[1:, 1:], bodyc[1:, 1:], rtol=1e-12, atol=1e-12)
    assert not gotc[0, 1:].any() and not gotc[1:, 0].any()
    hc = hilbert.chi0_complex_weights(sc[:, :1], Dr[1] + Dr[2], Di[1] + Di[2])
    np.testing.assert_allclose(gotc[0, 0], hc[0, 0], rtol=1e-12)


def test_multdiff_gamma_head_and_wings_cancel_on_the_mesh():
    """At Gamma, MULTDIFF keeps dcprod / sqrt(vol) and divides the head and
    wing weights by frqs^2 and frqs; with every transition's weight at
    frqs(s) = Delta_t that equals the multdiff_ = 0 form (dcprod / Delta,
    plain weights)."""
    g = np.random.default_rng(24)
    nk, nocc, nunocc, nfrqs, nb, nw = 3, 2, 3, 8, 4, 3
    frqs, eo, eu, wt = _multdiff_on_mesh(g, nk, nocc, nunocc, nfrqs)
    n = nk * nocc * nunocc
    cp, dp = _rand(g, n, nb)
    vol = 123.4
    wS = g.standard_normal((nfrqs, nw))
    sc = np.where(frqs > 0, frqs, 1.0)
    D0, D1, D2 = (hilbert.fold_weights(wt, wS / sc[:, None] ** ll) for ll in range(3))
    h0, w0 = gamma.headwings_imag(cp, gamma.momentum_scale(dp, eo, eu, vol), D0)
    p_md = gamma.momentum_scale(dp, None, None, vol)
    np.testing.assert_allclose(p_md, dp / math.sqrt(vol), rtol=1e-15)
    h1, w1 = gamma.headwings_imag(cp, p_md, D1, D2)
    np.testing.assert_allclose(h1, h0, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(w1, w0, rtol=1e-12, atol=1e-14)


def test_head_scalar_closed_forms():
    """trace/3 is the angular average of v^T head v over unit vectors (the
    mean of v v^T over the sphere is I/3), and a projection on an axis picks
    the diagonal element."""
    g = np.random.default_rng(19)
    B = g.standard_normal((3, 3)) + 1j * g.standard_normal((3, 3))
    H = np.eye(3) + B @ B.conj().T
    hp = oracle.pack_upper(H)
    v = g.standard_normal((20000, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    avg = np.mean([gamma.head_scalar(hp, x) for x in v[:4000]])
    assert abs(avg - gamma.head_scalar(hp)) < 0.05 * abs(gamma.head_scalar(hp))
    assert np.isclose(gamma.head_scalar(hp), np.trace(H) / 3)
    assert np.isclose(gamma.head_scalar(hp, [0.0, 1.0, 0.0]), H[1, 1])


def test_real_frequency_prepare_closed_form():
    g = np.random.default_rng(20)
    nb, nw = 5, 2
    chi = g.standard_normal((nb, nb, nw)) + 1j * g.standard_normal((nb, nb, nw))
    sq = g.uniform(0.5, 2.0, nb)
    head = g.standard_normal((3, 3, nw)) + 1j * g.standard_normal((3, 3, nw))
    wing = g.standard_normal((3, nb, 2, nw)) + 1j * g.standard_normal((3, nb, 2, nw))
    eps, h, w = gamma.gamma_prepare_cplx(chi, head, wing, sq)
    for f in range(nw):
        c = chi[:, :, f].copy()
        c[0, :] = 0
        c[:, 0] = 0
        np.testing.assert_allclose(eps[:, :, f], np.eye(nb) - np.diag(sq) @ c @ np.diag(sq),
                                   rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(h, np.eye(3)[:, :, None] - 4 * math.pi * head, rtol=1e-15)
    assert (w[:, 0] == 0).all()
    np.testing.assert_allclose(w[:, 1:], -math.sqrt(4 * math.pi) * sq[None, 1:, None, None] *
                               wing[:, 1:], rtol=1e-15)


def test_real_frequency_inverse_is_the_bordered_matrix_inverse():
    """Non-Hermitian case: with E = [[H, w2], [w1^T, B]] the reference's
    update gives S = H - w2 B^-1 w1^T, w1_new^T S^-1 below the head block of
    inv(E) and S^-1 w2_new beside it."""
    g = np.random.default_rng(21)
    nb, nw = 7, 2
    heads, wings, einvs, Es = [], [], [], []
    for _ in range(nw):
        E = np.eye(nb + 3) * (nb + 3) + g.standard_normal((nb + 3, nb + 3)) + \
            1j * g.standard_normal((nb + 3, nb + 3))
        H, X, Y, B = E[:3, :3], E[:3, 3:], E[3:, :3], E[3:, 3:]
        heads.append(H)
        wings.append(np.stack([Y.T, X], axis=2))  # wingc(:,:,1) = Y^T, wingc(:,:,2) = X
        einvs.append(np.linalg.inv(B))
        Es.append(E)
    h, w = gamma.gamma_inverse_cplx(np.stack(einvs, 2), np.stack(heads, 2), np.stack(wings, 3))
    for f in range(nw):
        Ei = np.linalg.inv(Es[f])
        Si = np.linalg.inv(h[:, :, f])
        np.testing.assert_allclose(Si, Ei[:3, :3], rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(w[:, :, 0, f].T @ Si, Ei[3:, :3], rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(Si @ w[:, :, 1, f], Ei[:3, 3:], rtol=1e-11, atol=1e-13)
