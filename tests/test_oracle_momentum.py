"""Pin the oracle's momentum matrix (dwavefproducts.f:89-284) by what the
mathematics fixes:

* muffin-tin part: for wave functions psi = sum c(l,m,n) u_ln(r)/r Y_lm in a
  sphere (complex Y_lm with the Condon-Shortley phase, the reference's
  harmonicsr, numerics.f:587-660), the oracle with the radial integrals
  qmat1/qmat2 formed as the reference forms them (:145-173, here by
  Gauss-Legendre quadrature of analytic radial functions) equals a direct 3D
  quadrature of conj(psi1) (-i d/dx_j psi2) over the sphere, the derivative
  by central differences of psi2 itself -- independent of the gradient
  formula, its F and G coefficients, the lm bookkeeping and the spherical
  to Cartesian transform;
* plane-wave part: with theta = 1 (cstep = delta_G0) the closed form
  sum_G conj(c1) c2 (k+G) of -i grad on plane waves; with a step function,
  the integral of theta conj(psi1) (-i grad psi2) over the cell on an
  alias-free real-space grid (gradients exact in Fourier space).
"""
import math

import numpy as np
import pytest
from scipy import special

from oracle import momentum


def _ylm(l, m, theta, phi):
    if hasattr(special, "sph_harm_y"):
        return special.sph_harm_y(l, m, theta, phi)
    return special.sph_harm(m, l, phi, theta)


def test_ylm_convention_matches_reference_low_orders():
    """harmonicsr: Y_00 = 0.2820947917738781, Y_10 = 0.48860251190292 cos(theta),
    Y_11 = -r (x + i y) / r with the Condon-Shortley sign (numerics.f:603-631)."""
    th, ph = 0.7, 1.1
    assert _ylm(0, 0, th, ph) == pytest.approx(0.282094791773878)
    assert _ylm(1, 0, th, ph) == pytest.approx(0.48860251190292 * math.cos(th))
    r11 = 0.282094791773878 * math.sqrt(3) / math.sqrt(2)
    assert _ylm(1, 1, th, ph) == pytest.approx(-r11 * math.sin(th) * np.exp(1j * ph))
    assert _ylm(1, -1, th, ph) == pytest.approx(r11 * math.sin(th) * np.exp(-1j * ph))


# radial functions u_ln(r) = r^(l+1) (1 + a r) exp(-b r) (u = r R, R regular)
def _u(l, n, r):
    a, b = 0.3 + 0.2 * n, 0.8 + 0.15 * l + 0.1 * n
    return r ** (l + 1) * (1 + a * r) * np.exp(-b * r)


def _du(l, n, r):
    a, b = 0.3 + 0.2 * n, 0.8 + 0.15 * l + 0.1 * n
    return ((l + 1) * r**l * (1 + a * r) + r ** (l + 1) * a - b * r ** (l + 1) * (1 + a * r)) * np.exp(-b * r)


def _qmats(lcut, nindx, R):
    """qmat1/qmat2 as dwavefproducts.f:145-173 forms them (bas1 = u, bas2 = 0),
    the radial integrals by 60-point Gauss-Legendre on [0, R]."""
    x, w = np.polynomial.legendre.leggauss(60)
    r = 0.5 * R * (x + 1)
    w = 0.5 * R * w
    mx = max(nindx)
    q1 = np.zeros((mx, mx, lcut + 1))
    q2 = np.zeros((mx, mx, lcut + 1))
    for l in range(lcut + 1):
        for n2 in range(nindx[l]):
            db = _du(l, n2, r) - _u(l, n2, r) / r
            if l > 0:
                for n1 in range(nindx[l - 1]):
                    q1[n1, n2, l] = (np.sum(w * db * _u(l - 1, n1, r))
                                     + (l + 1) * np.sum(w * _u(l, n2, r) * _u(l - 1, n1, r) / r))
            if l < lcut:
                for n1 in range(nindx[l + 1]):
                    q2[n1, n2, l] = (np.sum(w * db * _u(l + 1, n1, r))
                                     - l * np.sum(w * _u(l, n2, r) * _u(l + 1, n1, r) / r))
    return q1, q2


def _psi(c, lcut, nindx, pts):
    """psi at Cartesian points (npts, 3) from coefficients c[lm] (reference order)."""
    r = np.linalg.norm(pts, axis=1)
    th = np.arccos(np.clip(pts[:, 2] / r, -1, 1))
    ph = np.arctan2(pts[:, 1], pts[:, 0])
    out = np.zeros(len(pts), complex)
    lm = 0
    for l in range(lcut + 1):
        for m in range(-l, l + 1):
            y = _ylm(l, m, th, ph)
            for n in range(nindx[l]):
                out += c[lm] * _u(l, n, r) / r * y
                lm += 1
    return out


def test_momentum_mt_equals_direct_quadrature():
    lcut, nindx, R = 3, [2, 1, 2, 1], 2.2
    nn = sum((2 * l + 1) * nindx[l] for l in range(lcut + 1))
    g = np.random.default_rng(3)
    c = g.standard_normal((nn, 1, 2, 1)) + 1j * g.standard_normal((nn, 1, 2, 1))
    q1, q2 = _qmats(lcut, nindx, R)
    T = dict(lcut=lcut, nindx=nindx, neq=1, qmat1=q1, qmat2=q2)
    got = momentum.momentum_mt([T], c, [0], 0, 1, 1, 1)[0]  # <psi_0| -i grad |psi_1>
    # direct quadrature: GL in r and cos(theta), trapezoid in phi
    xr, wr = np.polynomial.legendre.leggauss(48)
    r = 0.5 * R * (xr + 1)
    wr = 0.5 * R * wr
    xc, wc = np.polynomial.legendre.leggauss(12)
    nph = 24
    ph = 2 * np.pi * np.arange(nph) / nph
    rr, cc, pp = np.meshgrid(r, xc, ph, indexing="ij")
    ww = (wr[:, None, None] * wc[None, :, None] * (2 * np.pi / nph)) * rr**2
    st = np.sqrt(1 - cc**2)
    pts = np.stack([rr * st * np.cos(pp), rr * st * np.sin(pp), rr * cc], -1).reshape(-1, 3)
    p1 = _psi(c[:, 0, 0, 0], lcut, nindx, pts)
    h = 1e-5
    want = []
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        d2 = (_psi(c[:, 0, 1, 0], lcut, nindx, pts + e) - _psi(c[:, 0, 1, 0], lcut, nindx, pts - e)) / (2 * h)
        want.append(np.sum(ww.reshape(-1) * np.conj(p1) * (-1j) * d2))
    want = np.array(want)
    assert np.abs(want).max() > 1e-2
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-8 * np.abs(want).max())


def test_momentum_mt_non_hermitian_part_is_the_surface_term():
    """Inside a sphere -i grad is Hermitian only up to a surface term:
    <1|p|2> - conj(<2|p|1>) = -i oint conj(psi1) psi2 n dS over the sphere,
    here with the swapped pair computed by the same formula."""
    lcut, nindx, R = 2, [1, 2, 1], 1.9
    nn = sum((2 * l + 1) * nindx[l] for l in range(lcut + 1))
    g = np.random.default_rng(4)
    c = g.standard_normal((nn, 1, 2, 1)) + 1j * g.standard_normal((nn, 1, 2, 1))
    q1, q2 = _qmats(lcut, nindx, R)
    T = dict(lcut=lcut, nindx=nindx, neq=1, qmat1=q1, qmat2=q2)
    a = momentum.momentum_mt([T], c, [0], 0, 1, 1, 1)[0]
    b = momentum.momentum_mt([T], c, [0], 1, 1, 0, 1)[0]
    xc, wc = np.polynomial.legendre.leggauss(10)
    ph = 2 * np.pi * np.arange(20) / 20
    cc, pp = np.meshgrid(xc, ph, indexing="ij")
    st = np.sqrt(1 - cc**2)
    nrm = np.stack([st * np.cos(pp), st * np.sin(pp), cc], -1).reshape(-1, 3)
    wS = (wc[:, None] * np.ones_like(pp) * (2 * np.pi / 20)).reshape(-1) * R**2
    p1 = _psi(c[:, 0, 0, 0], lcut, nindx, R * nrm)
    p2 = _psi(c[:, 0, 1, 0], lcut, nindx, R * nrm)
    surf = np.array([np.sum(wS * np.conj(p1) * p2 * nrm[:, j]) for j in range(3)])
    np.testing.assert_allclose(a - np.conj(b), -1j * surf, atol=1e-9 * np.abs(a).max())


def test_momentum_pw_theta_one_closed_form():
    g = np.random.default_rng(5)
    rlat = np.array([[1.1, 0.2, 0.0], [0.0, 0.9, 0.1], [0.3, 0.0, 1.2]])
    G = np.array([[0, 0, 0], [1, 0, 0], [0, -1, 1], [2, 1, -1], [-1, -1, 0]])
    kpt = np.array([[0.25, -0.1, 0.3]])
    cpw = (g.standard_normal((5, 3, 1)) + 1j * g.standard_normal((5, 3, 1)))
    cext = (3, 2, 2)
    cs = np.zeros((5, 5, 7), complex)
    cs[2, 2, 3] = 1.0  # theta = 1
    got = momentum.momentum_pw(rlat, kpt, G[None], [5], cpw, cs, cext, [0], 0, 1, 1, 2)
    kg = (kpt[0] + G) @ rlat.T
    want = np.array([[np.sum(np.conj(cpw[:, 0, 0]) * cpw[:, b, 0] * kg[:, j]) for j in range(3)]
                     for b in (1, 2)])
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-14)


def test_momentum_pw_with_step_equals_real_space_integral():
    """sum over the cell of theta(r) conj(psi1(r)) (-i grad psi2)(r) on a
    grid fine enough to hold the integrand's Fourier components exactly
    (|G| <= 2 in the states, |G| <= 2 in theta), gradients taken in Fourier
    space; equals the oracle's olapvec form up to the cell volume factor of
    the discrete sum (the grid mean)."""
    g = np.random.default_rng(6)
    rlat = 2 * np.pi * np.eye(3)
    r = np.arange(-2, 3)
    G = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    G = G[(G * G).sum(1) <= 4]
    ng = len(G)
    kpt = np.array([[0.1, 0.2, -0.15]])
    cpw = g.standard_normal((ng, 2, 1)) + 1j * g.standard_normal((ng, 2, 1))
    sb = 2
    cs = (g.standard_normal((5, 5, 5)) + 1j * g.standard_normal((5, 5, 5))) * 0.1
    cs = 0.5 * (cs + np.conj(cs[::-1, ::-1, ::-1]))
    cs[sb, sb, sb] = 0.8
    ce = 4  # cstep must cover every G_j - G_i (the reference sizes it 2 max|G| + ...)
    csp = np.zeros((2 * ce + 1,) * 3, complex)
    csp[ce - sb:ce + sb + 1, ce - sb:ce + sb + 1, ce - sb:ce + sb + 1] = cs
    got = momentum.momentum_pw(rlat, kpt, G[None], [ng], cpw, csp, (ce, ce, ce), [0], 0, 1, 1, 1)[0]
    N = 16
    x = np.arange(N) / N
    X = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1).reshape(-1, 3)  # fractional coords
    ph = np.exp(2j * np.pi * X @ G.T)  # e^{i G.r} (the k phase cancels in conj(psi1) psi2)
    th = np.zeros(len(X), complex)
    for z in range(-sb, sb + 1):
        for y in range(-sb, sb + 1):
            for xx in range(-sb, sb + 1):
                th += cs[z + sb, y + sb, xx + sb] * np.exp(2j * np.pi * X @ np.array([xx, y, z]))
    psi1 = ph @ cpw[:, 0, 0]
    kg = (kpt[0] + G) @ rlat.T
    want = np.array([np.mean(th * np.conj(psi1) * (ph @ (cpw[:, 1, 0] * kg[:, j])))
                     for j in range(3)])
    assert np.abs(th.imag).max() < 1e-12
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
