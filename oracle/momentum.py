"""Momentum matrix of the k.p expansion at q = 0 -- TEST INFRASTRUCTURE, part
of the oracle; only tests/ may import it.

Restates momentum_matrix (dwavefproducts.f:89-284) for one spin, no SOC, k
points stored in full (storeibz false), in plain numpy loops:

  momentum(b2, b1, k, :) = <phi_b1,k | -i nabla | phi_b2,k>

* muffin-tin part (:129-254): the coefficients F and G of the gradient
  formula for spherical harmonics (:131-142), the radial integrals qmat1
  (l -> l-1) and qmat2 (l -> l+1) given by the caller, the left vectors
  cvec1/2/3 (spherical components m+1, m, m-1) built from cmt of band 1,
  their dot products with cmt of band 2 divided by i, and the transform
  (c1 - c3)/sqrt2, -i (c1 + c3)/sqrt2, c2 to Cartesian components;
* plane-wave part (:256-281): vec1 = olapvec(cpw of band 1) (overlap.f:
  298-315, out(i) = sum_j conj(cstep(G_j - G_i)) vec(j)), then
  sum_G conj(vec1(G)) cpw(G, band 2) (k + G)_cart.

Conventions: cmt[lm, ic, band, k] with lm running over l, then m = -l..l,
then the radial index (the reference cmt(maxlmindx, ncent, nband, nkpt));
qmat1/qmat2[n1, n2, l] as the reference qmat(n1, n2, l, itype) (0-based);
cstep as a [z][y][x] array of extents 2*cext+1; rlat[:, j] basis vector j.
What pins each part: tests/test_oracle_momentum.py (a 3D quadrature of
conj(psi1) (-i grad psi2) inside a sphere for the muffin-tin part, the
closed form for theta = 1 and a real-space grid integral with theta for the
plane-wave part).
"""
from __future__ import annotations

import math

import numpy as np


def fg_coeffs(lmax):
    """fcoeff, gcoeff [(lmax+1)^2, 3] with columns -1, 0, 1 (dwavefproducts.f:131-142)."""
    n = (lmax + 1) ** 2
    f = np.zeros((n, 3))
    g = np.zeros((n, 3))
    lm = 0
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            f[lm, 0] = -math.sqrt(1.0 * (l + m + 1) * (l + m + 2) / (2 * (2 * l + 1) * (2 * l + 3)))
            f[lm, 1] = math.sqrt(1.0 * (l - m + 1) * (l + m + 1) / ((2 * l + 1) * (2 * l + 3)))
            f[lm, 2] = -math.sqrt(1.0 * (l - m + 1) * (l - m + 2) / (2 * (2 * l + 1) * (2 * l + 3)))
            if l > 0:
                g[lm, 0] = math.sqrt(1.0 * (l - m) * (l - m - 1) / (2 * (2 * l - 1) * (2 * l + 1)))
                g[lm, 1] = math.sqrt(1.0 * (l - m) * (l + m) / ((2 * l - 1) * (2 * l + 1)))
                g[lm, 2] = math.sqrt(1.0 * (l + m) * (l + m - 1) / (2 * (2 * l - 1) * (2 * l + 1)))
            lm += 1
    return f, g


def _cvecs(cmt1, lcut, nindx, qmat1, qmat2, f, g):
    """Left vectors cvec1/2/3 of one band (dwavefproducts.f:198-240)."""
    nn = sum((2 * l + 1) * nindx[l] for l in range(lcut + 1))
    cv = np.zeros((3, nn), complex)
    lm_0, lm_1, lm = 0, nindx[0], 0
    for l in range(lcut):  # l -> l+1 functions of band 1 (qmat2)
        n0, n1 = nindx[l], nindx[l + 1]
        q = qmat2[:n1, :n0, l]
        for m in range(-l, l + 1):
            lm0 = lm_0 + (m + l) * n0
            for x, dm in enumerate((1, 0, -1)):
                lm1 = lm_1 + (m + dm + l + 1) * n1
                cv[x, lm0:lm0 + n0] = f[lm, x] * (cmt1[lm1:lm1 + n1] @ q)
            lm += 1
        lm_0 += (2 * l + 1) * n0
        lm_1 += (2 * l + 3) * n1
    lm_0, lm_1, lm = nindx[0], 0, 1
    for l in range(1, lcut + 1):  # l -> l-1 functions of band 1 (qmat1)
        n0, n1 = nindx[l], nindx[l - 1]
        q = qmat1[:n1, :n0, l]
        for m in range(-l, l + 1):
            lm0 = lm_0 + (m + l) * n0
            for x, dm in enumerate((1, 0, -1)):
                if abs(m + dm) <= l - 1:
                    lm1 = lm_1 + (m + dm + l - 1) * n1
                    cv[x, lm0:lm0 + n0] += g[lm, x] * (cmt1[lm1:lm1 + n1] @ q)
            lm += 1
        lm_0 += (2 * l + 1) * n0
        lm_1 += (2 * l - 1) * n1
    return cv, nn


def momentum_mt(types, cmt, k1, b1lo, nocc, b2lo, nunocc):
    """Muffin-tin part, (n, 3) complex in the transition order t = (k*nocc +
    b1)*nunocc + b2.  types: dicts lcut, nindx, neq, qmat1, qmat2."""
    lmax = max(T["lcut"] for T in types)
    f, g = fg_coeffs(lmax)
    out = np.zeros((len(k1) * nocc * nunocc, 3), complex)
    s2 = math.sqrt(2.0)
    for ik, k in enumerate(k1):
        ic = 0
        for T in types:
            for _ in range(T["neq"]):
                for i1 in range(nocc):
                    cv, nn = _cvecs(cmt[:, ic, b1lo + i1, k], T["lcut"], T["nindx"], T["qmat1"],
                                    T["qmat2"], f, g)
                    for i2 in range(nunocc):
                        c2 = cmt[:nn, ic, b2lo + i2, k]
                        c = [np.sum(np.conj(cv[x]) * c2) / 1j for x in range(3)]
                        t = (ik * nocc + i1) * nunocc + i2
                        out[t, 0] += (c[0] - c[2]) / s2
                        out[t, 1] -= (c[0] + c[2]) / s2 * 1j
                        out[t, 2] += c[1]
                ic += 1
    return out


def olapvec(vec, G, cstep, cext):
    """overlap.f:298-315: out(i) = sum_j conj(cstep(G_j - G_i)) vec(j)."""
    c = np.asarray(cstep)
    d = G[None, :, :] - G[:, None, :]  # [i, j] = G_j - G_i
    row = c[d[..., 2] + cext[2], d[..., 1] + cext[1], d[..., 0] + cext[0]]
    return (np.conj(row) * vec[None, :]).sum(1)


def momentum_pw(rlat, kpt, gpt, ngpt, cpw, cstep, cext, k1, b1lo, nocc, b2lo, nunocc):
    """Plane-wave part, (n, 3) complex; gpt[k, i] the G vectors of k (ngpt[k]
    of them), cpw[G, band, k]."""
    out = np.zeros((len(k1) * nocc * nunocc, 3), complex)
    rlat = np.asarray(rlat, float)
    for ik, k in enumerate(k1):
        ng = int(ngpt[k])
        G = np.asarray(gpt[k, :ng], np.int64)
        qg = (np.asarray(kpt[k], float)[None, :] + G) @ rlat.T  # rlat . (k + G)
        for i1 in range(nocc):
            v1 = olapvec(cpw[:ng, b1lo + i1, k], G, cstep, cext)
            for i2 in range(nunocc):
                v2 = cpw[:ng, b2lo + i2, k]
                t = (ik * nocc + i1) * nunocc + i2
                for x in range(3):
                    out[t, x] += np.sum(np.conj(v1) * (v2 * qg[:, x]))
    return out
