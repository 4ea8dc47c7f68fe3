"""Oracle restatement of the HILBERT frequency weights (hilbert.f getfrqs,
getwghtS) and of the HILBERT pair sum (susceptibility.f:1066-1077,
1793-1849, 1231-1236), pinned by the mathematics: quadrature of the
defining integrals, the mesh recursion, the HILBERT NONE limit and the
reference's own two orders of summation."""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle
from oracle import hilbert


def _hat(frqs, a, s):
    """Piecewise-linear basis function of mesh point s as getwghtS
    integrates it: the first interval starts at max(0, frqs[0] - a*winc),
    where the interpolant is zero."""
    winc = frqs[1] - frqs[0]
    knots = np.concatenate([[max(0.0, frqs[0] - a * winc)], frqs])
    y = np.zeros(len(knots))
    y[s + 1] = 1.0
    return knots, y


def _quad_weight(frqs, s, kernel, md=0):
    """int hat_s(w') w'^-md kernel(w') dw' over the mesh, interval by interval."""
    winc = frqs[1] - frqs[0]
    a = winc / (frqs[2] - frqs[0]) if len(frqs) >= 3 else 0.5
    knots, y = _hat(frqs, a, s)
    tot = 0j
    for k in range(max(0, s), min(len(knots) - 1, s + 2)):
        x0, x1, y0, y1 = knots[k], knots[k + 1], y[k], y[k + 1]

        def f(x, part):
            v = (y0 + (y1 - y0) * (x - x0) / (x1 - x0)) * x ** (-md) * kernel(x)
            return v.real if part == 0 else v.imag

        tot += integrate.quad(f, x0, x1, args=(0,), epsabs=1e-14, epsrel=1e-12, limit=200)[0]
        tot += 1j * integrate.quad(f, x0, x1, args=(1,), epsabs=1e-14, epsrel=1e-12, limit=200)[0]
    return tot


def test_mesh_recursion_and_span():
    """frqs(2)-frqs(1) = winc, each step stretched by h, first point
    wmin + winc/(1+h), last point covers wmax (hilbert.f:100-104)."""
    winc, h, wmin, wmax = 0.01, 1.12, 0.05, 3.0
    f = hilbert.hilbert_mesh((winc, h), wmin, wmax)
    a = 1 / (1 + h)
    assert abs(f[0] - (wmin + winc * a)) < 1e-15
    d = np.diff(f)
    assert np.allclose(d, winc * h ** np.arange(len(d)), rtol=1e-12)
    assert f[-1] - d[-1] * h * a >= wmax - 1e-12 or f[-1] >= wmax
    assert f[-2] - d[-2] * h * a < wmax
    # the stretching-factor consistency the reference asserts (hilbert.f:165-170)
    assert np.all(np.abs(a - (f[1:-1] - f[:-2]) / (f[2:] - f[:-2])) < 1e-13)


def test_mesh_fspec_point_count():
    """Old FSPEC form: -n points from zero, n-1 increments summing to wmax."""
    f = hilbert.hilbert_mesh((-20, 1.1), 0.3, 2.0)
    assert f[0] == 0.0
    w = 2.0 * (1.1 - 1) / (1.1 ** 19 - 1)
    assert abs(f[1] - w) < 1e-14


@pytest.mark.parametrize("omega", [0.05, 0.7, 4.0])
def test_imaginary_axis_weights_equal_quadrature(omega):
    """Both branches on w = i omega: int hat_s(w') [1/(w-w') + 1/(-w-w')] dw'
    = int hat_s(w') (-2 w') / (omega^2 + w'^2) dw', real."""
    f = hilbert.hilbert_mesh((0.02, 1.15), 0.08, 2.5)
    wr, wi = hilbert.hilbert_weights([1j * omega], f, 0.0, 0, 0)
    kern = lambda x: -2 * x / (omega ** 2 + x ** 2) + 0j
    for s in (0, 1, 5, len(f) // 2, len(f) - 1):
        want = _quad_weight(f, s, kern)
        assert abs(wr[s, 0] - want.real) < 1e-11 * max(1.0, abs(want))
        assert abs(wi[s, 0]) < 1e-12


def test_wings_weights_equal_quadrature():
    """md = 1 (wings): S(w') = S'(w')/w' with S' piecewise linear; positive
    branch only, real frequency with disorder."""
    f = hilbert.hilbert_mesh((0.03, 1.2), 0.2, 2.0)
    w, dis = 0.913, 0.02
    wr, wi = hilbert.hilbert_weights([w], f, dis, 1, 1)
    kern = lambda x: 1.0 / (w - x + 1j * dis)
    for s in (0, 3, len(f) // 2, len(f) - 1):
        want = _quad_weight(f, s, kern, 1)
        got = wr[s, 0] + 1j * wi[s, 0]
        assert abs(got - want) < 1e-9 * max(1.0, abs(want)), (s, got, want)


def test_head_weights_differ_from_their_integral_only_by_the_reference_sign():
    """md = 2 (head, q = 0 only): the reference forms the 1/w'^2 integral as
    int1 = (int2 + rdum1)/z (hilbert.f:250), where the defining integral
    int w'^-2/(z - w') dw' is (int2 - rdum1)/z, rdum1 = 1/w2 - 1/w1.  The
    oracle keeps the reference's form.  Pinned here: oracle - quadrature is
    exactly the hat-weighted 2 rdum1/z of every interval, computed apart."""
    f = hilbert.hilbert_mesh((0.03, 1.2), 0.2, 2.0)
    w, dis = 0.913, 0.02
    z = w + 1j * dis
    wr, wi = hilbert.hilbert_weights([w], f, dis, 2, 1)
    winc = f[1] - f[0]
    a = winc / (f[2] - f[0])
    knots = np.concatenate([[max(0.0, f[0] - a * winc)], f])
    corr = np.zeros(len(f), complex)
    for k in range(len(knots) - 1):
        w1, w2 = knots[k], knots[k + 1]
        d = 2 * (1 / w2 - 1 / w1) / z  # int1(reference) - int1(integral)
        if k > 0:
            corr[k - 1] += w2 * d / (w2 - w1)
        corr[k] += -w1 * d / (w2 - w1)
    kern = lambda x: 1.0 / (w - x + 1j * dis)
    for s in (0, 3, len(f) // 2, len(f) - 1):
        want = _quad_weight(f, s, kern, 2) + corr[s]
        got = wr[s, 0] + 1j * wi[s, 0]
        assert abs(got - want) < 1e-9 * max(1.0, abs(want)), (s, got, want)


def test_real_frequency_weights_equal_quadrature():
    """md = 0, both branches, real frequency away from the mesh points."""
    f = hilbert.hilbert_mesh((0.025, 1.1), 0.1, 2.2)
    w, dis = 1.2345, 0.05
    wr, wi = hilbert.hilbert_weights([w], f, dis, 0, 0)
    kern = lambda x: 1.0 / (w - x + 1j * dis) + 1.0 / (-w - x + 1j * dis)
    for s in (0, 2, 7, len(f) - 2):
        want = _quad_weight(f, s, kern)
        got = wr[s, 0] + 1j * wi[s, 0]
        assert abs(got - want) < 1e-9 * max(1.0, abs(want)), (s, got, want)


def test_delta_spectral_function_gives_hilbert_none_weights():
    """A transition of energy d carries the spectral function delta(w'-d).
    As a spectral density on a uniform mesh of spacing dw it is (1-x)/dw and
    x/dw on its two neighbours (the interpolant integrates to 1); the
    Hilbert sum then tends to the HILBERT NONE weight -2d/(omega^2+d^2)
    (susceptibility.f:963-968) as the mesh is refined, with an O(dw^2)
    error."""
    d, omega = 0.73, 0.4
    want = -2 * d / (omega ** 2 + d ** 2)
    errs = []
    for winc in (0.04, 0.02, 0.01):
        f = hilbert.hilbert_mesh((winc, 1.0), 0.05, 3.0)
        wr, _ = hilbert.hilbert_weights([1j * omega], f, 0.0, 0, 0)
        s = np.searchsorted(f, d) - 1
        dw = f[s + 1] - f[s]
        x = (d - f[s]) / dw
        wt = np.zeros((1, len(f)))
        wt[0, s], wt[0, s + 1] = (1 - x) / dw, x / dw
        errs.append(abs(hilbert.fold_weights(wt, wr)[0, 0] - want))
    assert errs[0] < 2e-3 * abs(want)
    assert errs[1] < errs[0] / 3 and errs[2] < errs[1] / 3


def test_fold_then_rank_k_equals_reference_spectral_function_order():
    """The reference sums per pair: S(s) = sum_t re(t) wtetra(t,s), then
    sum_s S(s) wghtS(s,w).  Folding first, D = wtetra wghtS, then the
    rank-k pair sum gives the same matrix (associativity); the threshold
    drops the same terms in both."""
    g = np.random.default_rng(11)
    n, nb = 60, 9
    f = hilbert.hilbert_mesh((0.05, 1.2), 0.1, 2.5)
    wr, _ = hilbert.hilbert_weights([0.2j, 1.1j, 3.0j], f, 0.0, 0, 0)
    wt = g.uniform(0, 1, (n, len(f))) * (g.uniform(0, 1, (n, len(f))) < 0.15)
    wt[wt < 0.05] = 0.0
    wt[0, 0] = 0.01  # below the threshold
    cprod = (g.standard_normal((n, nb)) + 1j * g.standard_normal((n, nb))) / 3
    ref = hilbert.chi0_hilbert_pairsum(cprod, wt, wr, wghtthr=0.02)
    D = hilbert.fold_weights(wt, wr, wghtthr=0.02)
    packed = oracle.chi0_herk(cprod, D)
    for w in range(3):
        full = oracle.unpack_herm(packed[:, w], nb)
        assert np.max(np.abs(full - ref[:, :, w])) < 1e-13 * np.max(np.abs(ref[:, :, w]))


def test_complex_weights_pair_loop_equals_direct_sum():
    """suscepc accumulation (susceptibility.f:1231-1236) against the direct
    sum over t of (Dr + i Di) cprod(t,i) conj(cprod(t,j)); with Di = 0 it is
    the Hermitian matrix of the HILBERT NONE pair sum."""
    g = np.random.default_rng(5)
    n, nb, nw = 40, 7, 2
    cprod = g.standard_normal((n, nb)) + 1j * g.standard_normal((n, nb))
    Dr = g.standard_normal((n, nw))
    Di = g.standard_normal((n, nw))
    got = hilbert.chi0_complex_weights(cprod, Dr, Di)
    want = np.einsum("tw,ti,tj->ijw", Dr + 1j * Di, cprod, np.conj(cprod))
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    herm = hilbert.chi0_complex_weights(cprod, Dr, np.zeros_like(Di))
    packed = oracle.chi0_pairsum(cprod, Dr)
    for w in range(nw):
        assert np.max(np.abs(oracle.unpack_herm(packed[:, w], nb) - herm[:, :, w])) < 1e-12 * np.max(np.abs(herm))
