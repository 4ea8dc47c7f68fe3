"""Hilbert-mesh (HILBERT) frequency weights -- TEST INFRASTRUCTURE, part of
the oracle; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import it.

Restates, in plain numpy/Python:
- getfrqs   hilbert.f:106-174  the Hilbert mesh frqs(1:nfrqs) on [wmin, wmax];
- getwghtS  hilbert.f:193-300  the weights wghtS(s, j) of the Hilbert
  transform  chi(w_j) = sum_s (wghtSr + i wghtSi)(s, j) S(frqs(s));
- the HILBERT branch of the pair sum: the threshold on the tetrahedron
  weights (susceptibility.f:1066-1077, wtetra > wghtthr), the spectral
  function S(s) = sum_t re(t) wtetra(t, s) and the Hilbert sums
  (spectral_function, susceptibility.f:1793-1849), and the accumulation of
  the real-frequency matrix suscepc (:1231-1236).

What pins each function is in tests/test_oracle_hilbert.py: quadrature of
the defining integrals (scipy), the closed-form mesh recursion, the HILBERT
NONE limit of a delta spectral function, and the reference's own two orders
of summation.
"""
from __future__ import annotations

import cmath
import math

import numpy as np


def hilbert_mesh(fspec, wmin, wmax):
    """frqs of getfrqs (hilbert.f:106-174) for fspec = (fspec1, fspec2).

    fspec1 > 0: HILBERT (first increment, stretching factor); fspec1 < 0:
    the old FSPEC form (mesh from zero; a negative integer is a point count).
    fspec1 == 0: no mesh (HILBERT NONE)."""
    f1, h = float(fspec[0]), float(fspec[1])
    if f1 == 0:
        return np.zeros(0)
    wmin1 = 0.0 if f1 < 0 else wmin
    wdif = wmax - wmin1
    winc = abs(f1)
    if f1 < 0:
        n = int(round(-f1))
        if abs(n + f1) < 1e-12:
            winc = wdif / (n - 1) if h == 1 else wdif * (h - 1) / (h ** (n - 1) - 1)
    a = 1 / (1 + h)
    frqs = []
    rdum = 0.0 if f1 < 0 else wmin1 + winc * a
    winc1 = winc
    while rdum - winc1 * a < wmax:
        frqs.append(rdum)
        rdum = rdum + winc1
        winc1 = winc1 * h
    return np.array(frqs)


def _ln(z):
    """Principal complex log with log(-x + 0i) = log(x) + i pi for either
    sign of the zero imaginary part (hilbert.f:285-297)."""
    a, b = z.real, z.imag
    if a < 0 and b == 0:
        return complex(math.log(-a), math.pi)
    return cmath.log(z)


def hilbert_weights(frq, frqs, disorder=0.0, md=0, iterm=0):
    """(wghtSr, wghtSi), each (nfrqs, nfrq), of getwghtS (hilbert.f:193-300)."""
    frq = np.asarray(frq, complex)
    frqs = np.asarray(frqs, float)
    nfrq, nfrqs = len(frq), len(frqs)
    wr = np.zeros((nfrqs, nfrq))
    wi = np.zeros((nfrqs, nfrq))
    if nfrq == 0:
        return wr, wi
    winc = frqs[1] - frqs[0]
    a = 0.5
    if nfrqs >= 3:
        a = winc / (frqs[2] - frqs[0])
    img = 1j
    signs = (1, -1) if iterm == 0 else ((1,) if iterm == 1 else (-1,))
    for isign in signs:
        for j in range(nfrq):
            w = complex(frq[j]) * isign
            i2 = 1 if frqs[0] == 0 else 0  # 1-based interval ends, as in the reference
            while i2 < nfrqs:
                i1 = i2
                rdum = abs(w - frqs[i1])  # frqs(i1+1), 0-based frqs[i1]
                if rdum > min(winc / 3, 1e-6):
                    i2 = i1 + 1
                else:
                    i2 = i1 + 2
                    if i2 > nfrqs:
                        break
                w1 = max(0.0, frqs[0] - a * winc) if i1 == 0 else frqs[i1 - 1]
                w2 = frqs[i2 - 1]
                int1 = 0j
                int2 = 0j
                cdum = 0j
                r0 = r1 = r2 = 0.0
                if w + w1 - img * disorder != 0:
                    cdum = _ln(w - w1 + img * disorder) - _ln(w - w2 + img * disorder)
                if w1 != 0 and md > 0:
                    r0 = math.log(w2) - math.log(w1)
                    r1 = w2 ** -1 - w1 ** -1
                    r2 = (w2 ** -2 - w1 ** -2) / 2
                if md == 0:
                    int1 = cdum
                    int2 = cdum * (w + img * disorder) - (w2 - w1)
                elif md == 1:
                    int2 = cdum
                    int1 = (cdum + r0) / (w + img * disorder) if w - img * disorder != 0 else r1
                elif md == 2:
                    if w - img * disorder != 0:
                        int2 = (cdum + r0) / (w + img * disorder)
                        int1 = (int2 + r1) / (w + img * disorder)
                    else:
                        int2 = r1
                        int1 = r2
                if i2 - i1 == 1:
                    c = (w2 * int1 - int2) / (w2 - w1)
                    if i1 > 0:
                        wr[i1 - 1, j] += c.real
                        wi[i1 - 1, j] += c.imag
                    c = (int2 - w1 * int1) / (w2 - w1)
                    wr[i2 - 1, j] += c.real
                    wi[i2 - 1, j] += c.imag
                else:  # i2 - i1 == 2: a mesh point too close to the pole
                    if i1 > 0:
                        wr[i1 - 1, j] += isign
                    i1 = i1 + 1
                    wr[i1 - 1, j] += int1.real
                    wi[i1 - 1, j] += int1.imag
                    wr[i2 - 1, j] = wr[i1 - 1, j] - isign
    return wr, wi


def fold_weights(wtetra, wghtS, wghtthr=0.0):
    """D(t, j) = sum over s with wtetra(t, s) > wghtthr of wtetra(t, s) wghtS(s, j)
    -- the per-transition form of the HILBERT pair sum, with the
    reference's threshold (susceptibility.f:1066-1077).  wtetra (n, nfrqs),
    wghtS (nfrqs, nfrq); returns (n, nfrq)."""
    wt = np.where(np.asarray(wtetra) > wghtthr, wtetra, 0.0)
    return wt @ np.asarray(wghtS)


def chi0_hilbert_pairsum(cprod, wtetra, wghtSr, wghtthr=0.0):
    """Literal reference order for the imaginary-axis matrix: per pair (i<=j)
    re/im(t) = cprod(t,i) conj(cprod(t,j)) (susceptibility.f:1195-1201), the
    spectral function S(s) = sum over indexed t of re(t) wtetra (:1793-1815)
    and sus(w) = sum_s S(s) wghtSr(s, w) (:1838-1841).  Returns the full
    Hermitian (nb, nb, nfrq) matrix (small sizes only)."""
    cp = np.asarray(cprod, complex)
    n, nb = cp.shape
    wt = np.asarray(wtetra, float)
    nfrqs, nfrq = np.asarray(wghtSr).shape
    idx = [np.nonzero(wt[:, s] > wghtthr)[0] for s in range(nfrqs)]
    out = np.zeros((nb, nb, nfrq), complex)
    for j in range(nb):
        for i in range(j + 1):
            c = cp[:, i] * np.conj(cp[:, j])
            reS = np.array([np.dot(c.real[ix], wt[ix, s]) for s, ix in enumerate(idx)])
            imS = np.array([np.dot(c.imag[ix], wt[ix, s]) for s, ix in enumerate(idx)])
            v = reS @ wghtSr + 1j * (imS @ wghtSr)
            out[i, j] = v
            out[j, i] = np.conj(v)
    return out


def chi0_complex_weights(cprod, Dr, Di):
    """Real-frequency matrix suscepc(i, j, w) = sum_t (Dr + i Di)(t, w)
    cprod(t, i) conj(cprod(t, j)) for all i, j -- the accumulation of
    susceptibility.f:1231-1236 (sus1 + i susc1 + (-susc2 + i sus2) above the
    diagonal, its counterpart below).  Literal pair loop over i <= j."""
    cp = np.asarray(cprod, complex)
    n, nb = cp.shape
    Dr = np.asarray(Dr, float)
    Di = np.asarray(Di, float)
    nw = Dr.shape[1]
    out = np.zeros((nb, nb, nw), complex)
    for j in range(nb):
        for i in range(j + 1):
            c = cp[:, i] * np.conj(cp[:, j])
            re, im = c.real, c.imag
            for w in range(nw):
                sus1, susc1 = re @ Dr[:, w], re @ Di[:, w]
                sus2, susc2 = im @ Dr[:, w], im @ Di[:, w]
                cdum = complex(sus1, susc1)
                cdum1 = complex(-susc2, sus2)
                out[i, j, w] += cdum + cdum1
                if i != j:
                    out[j, i, w] += cdum - cdum1
    return out
