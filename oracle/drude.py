"""Drude term of the head at q = 0 for metals -- TEST INFRASTRUCTURE, part of
the oracle; only tests/ may import it.

Restates, in plain numpy/Python loops:
- tetrahedron4_init (tetrahedron.f:708-838) for a full cubic k mesh without
  symmetry reduction: the Fermi-surface weights wintgr4(n, k) -- per band
  and tetrahedron the cross-section with e = efermi, a triangle or a
  quadrangle of two triangles with areas atria and corner weights wtria
  (:756-791), scattered to the cube corners in the reference's loop order,
  divided by 3*6*nkpt -- then averaged over degenerate states (|e(n+1) -
  e(n)| < 1e-8, :826-838);
- the plasma frequency of susceptibility.f:1390-1441 and :1693-1700 for one
  spin: integrand(n, k) = sum_j (real part of the mean of momentum(i, i, k,
  j) over the degenerate group of n)^2, plasma^2 = 4 pi/(3 vol) sum
  wintgr4 * integrand, times 2 for a spin-unpolarised system, then the
  square root, set to zero below 1e-6 eV; for two spins the squares are
  summed before the root (the spin loop of :1395-1437);
- drude / drudec (correlation.f:1144-1169) and the head update of
  correlation.f:847-862: head(j, j, w) -= drude(freq(w)).

k points are numbered as in oracle.tetra (i3 fastest).  What pins each
function is in tests/test_oracle_drude.py.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import tetra

HARTREE = 27.2113831243217  # eV (susceptibility.f:1695)


def _rorderp(e):
    """rorderp (util.f:37-56): the stable ascending order of e."""
    return np.argsort(e, kind="stable")


def _cross_section(et, ef):
    """(atria[ntria], wtria[4, ntria]) of one tetrahedron with corner energies
    et cut by ef (tetrahedron.f:753-784), wtria in the original corner order."""
    pnt = _rorderp(et)
    e = et[pnt]
    wtria = np.zeros((4, 2))
    if ef > e[2]:
        f = (ef - e[:3]) / (e[3] - e[:3])
        atria = [3 * (ef - e[3]) ** 2 / ((e[3] - e[0]) * (e[3] - e[1]) * (e[3] - e[2]))]
        wtria[:3, 0] = 1 - f
        wtria[3, 0] = f[0] + f[1] + f[2]
    elif ef > e[1]:
        f = np.zeros(4)
        f[0:2] = (ef - e[0]) / (e[2:4] - e[0])
        f[2:4] = (ef - e[1]) / (e[2:4] - e[1])
        rdum = (e[3] - e[1]) * (e[2] - e[0])
        atria = [3 * f[2] * (e[2] - ef) / rdum, 3 * f[1] * (e[3] - ef) / rdum]
        wtria[0] = [1 - f[0], 2 - f[0] - f[1]]
        wtria[1] = [2 - f[2] - f[3], 1 - f[3]]
        wtria[2] = [f[0] + f[2], f[0]]
        wtria[3] = [f[3], f[1] + f[3]]
    else:
        f = (ef - e[0]) / (e[1:4] - e[0])
        atria = [3 * f[0] * f[1] / (e[2] - e[0])]
        wtria[0, 0] = 3 - f[0] - f[1] - f[2]
        wtria[1:, 0] = f
    out = np.zeros((4, 2))
    out[pnt] = wtria  # wtria(pnt,:) = wtria
    return np.array(atria), out[:, :len(atria)]


def fermi_surface_weights(nkpt3, rlat, ene, efermi):
    """wintgr4 (nk, nband) of tetrahedron4_init; ene (nk, nband) sorted per k."""
    ene = np.asarray(ene, float)
    nk, nband = ene.shape
    n1m, n2m, n3m = nkpt3
    assert n1m * n2m * n3m == nk
    tet = tetra.get_tetra(rlat, nkpt3)
    pk = tetra.mesh_index(nkpt3)
    W = np.zeros((nk, nband))
    for n in range(nband):
        e1 = ene[:, n]
        if e1.min() >= efermi or e1.max() <= efermi:
            continue
        for k3 in range(n3m):
            for k2 in range(n2m):
                for k1 in range(n1m):
                    kc = [pk(k1 + d1, k2 + d2, k3 + d3)
                          for d3 in (0, 1) for d2 in (0, 1) for d1 in (0, 1)]
                    ec = e1[kc]
                    for it in range(6):
                        et = ec[tet[it]]
                        if np.all(et >= efermi) or np.all(et <= efermi):
                            continue
                        atria, wtria = _cross_section(et, efermi)
                        for itr in range(len(atria)):
                            for i in range(4):
                                W[kc[tet[it][i]], n] += atria[itr] * wtria[i, itr]
    W /= 3 * 6 * nk
    # degenerate states (tetrahedron.f:826-838)
    for k in range(nk):
        n = 0
        while n < nband - 1:
            n1 = n
            while abs(ene[k, n1 + 1] - ene[k, n]) < 1e-8:
                n1 += 1
                if n1 == nband - 1:
                    break
            W[k, n:n1 + 1] = W[k, n:n1 + 1].sum() / (n1 - n + 1)
            n = n1 + 1
    return W


def plasma_frequency(wintgr4, mom_diag, deg, vol, spin_factor=2.0):
    """susceptibility.f:1404-1441, :1693-1700 for one spin of a spin-
    unpolarised system (spin_factor 2).  mom_diag (nk, nband, 3) complex:
    momentum(i, i, k, :); deg (nk, nband) zero-based last band of each band's
    degenerate group (SPEX's deg, 0-based)."""
    W = np.asarray(wintgr4, float)
    nk, nband = W.shape
    integrand = np.zeros((nk, nband))
    for k in range(nk):
        i1 = 0
        while i1 < nband:
            i2 = int(deg[k][i1])
            if i2 >= nband:
                break
            re = np.real(mom_diag[k, i1:i2 + 1, :].sum(0)) / (i2 - i1 + 1)
            integrand[k, i1:i2 + 1] = np.sum(np.abs(re) ** 2)
            i1 = i2 + 1
    p2 = 4 * math.pi / 3 / vol * np.sum(W * integrand) * spin_factor
    if p2 < -1e-10:
        raise ValueError("plasma frequency undefined")
    p = math.sqrt(max(0.0, p2))
    if p * HARTREE < 1e-6:
        p = 0.0
    return p


def plasma_frequency_spins(spins, vol, spin_factor=1.0):
    """The loop over spins of susceptibility.f:1395-1441: plasma += 4 pi/(3
    vol) sum(wdrude * integrand) for each (wintgr4, mom_diag, deg) of
    `spins`, times spin_factor (2 for one spin of an unpolarised system
    without SOC, :1439), then the checks and the root of :1440-1441,
    1693-1700."""
    p2 = 0.0
    for W, mom_diag, deg in spins:
        W = np.asarray(W, float)
        nk, nband = W.shape
        integrand = np.zeros((nk, nband))
        for k in range(nk):
            i1 = 0
            while i1 < nband:
                i2 = int(deg[k][i1])
                if i2 >= nband:
                    break
                re = np.real(mom_diag[k, i1:i2 + 1, :].sum(0)) / (i2 - i1 + 1)
                integrand[k, i1:i2 + 1] = np.sum(np.abs(re) ** 2)
                i1 = i2 + 1
        p2 = p2 + 4 * math.pi / 3 / vol * np.sum(W * integrand)
    p2 = p2 * spin_factor
    if p2 < -1e-10:
        raise ValueError("plasma frequency undefined")
    p = math.sqrt(max(0.0, p2))
    if p * HARTREE < 1e-6:
        p = 0.0
    return p


def drude(frq, plasma):
    """correlation.f:1147-1157 (imaginary frequencies)."""
    if plasma <= 0:
        return 0.0
    return math.inf if frq == 0 else -plasma**2 / frq**2


def drudec(frq, plasma):
    """correlation.f:1159-1169 (complex frequencies)."""
    if plasma <= 0:
        return 0j
    return complex(math.inf) if frq == 0 else plasma**2 / complex(frq) ** 2


def head_minus_drude(head, freq, plasma, headc=None, freqc=None):
    """correlation.f:847-862: head (6, nw) packed, headc (3, 3, nwc)."""
    h = np.array(head, complex)
    for i, f in enumerate(freq):
        for j, k in enumerate((0, 2, 5)):  # packed diagonal: k = j(j+1)/2 + j
            h[k, i] -= drude(f, plasma)
    if headc is None:
        return h
    hc = np.array(headc, complex)
    for i, f in enumerate(freqc):
        for j in range(3):
            hc[j, j, i] -= drudec(f, plasma)
    return h, hc
