"""Gamma-point (q = 0) head and wings -- TEST INFRASTRUCTURE, part of the
oracle; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import it.

Restates, in plain numpy/Python:
- the momentum-matrix scaling of the k.p expansion, susceptibility.f:858-873
  (dcprod / sqrt(vol), then / (e_unocc - e_occ), zero where that difference
  is at most 1e-8 in size; multdiff_ = 0);
- the head and wing sums of the pair loop, susceptibility.f:1092-1157, with
  the HILBERT NONE spectral function (:1780-1788) or folded Hilbert weights:
  head(k1,k2) = sum_t D(t) dcprod(t,k1) conj(dcprod(t,k2)) over k1 <= k2
  (packed, ij = k1 + k2(k2+1)/2 zero-based), wing(k2,i) = sum_t D(t)
  cprod(t,i) conj(dcprod(t,k2)), and their real-frequency forms with the
  complex weights (sus1 + i sus2 + i susc1 - susc2 and its partner);
- the MULTDIFF form of spectral_function (:1817-1824): the spectral function
  divided by frqs(s)**mm where frqs(s) > 0;
- MULTDIFF ON at any q (multdiff_ = 1, :588-592): cprod(:,1) times the
  band-energy difference (:912-925) and the pair loop with mm = the number
  of pair indices that are the first basis function (:1193); MULTDIFF
  INTON (multdiff_ = 2): the three passes ll = 0, 1, 2 over the pairs they
  visit (:1088-1090, 1168-1170), each with its own Hilbert weights;
- the Gamma-point dielectric step, correlation.f:570-646 (body row and
  column of the first basis function set to zero, wing(:,1,:) = 0,
  head = 1 - 4 pi head, wing = -sqrt(4 pi) sqcoul wing), the head's scalar
  in eps(1,1) (:651-665, :683), and the head and wings of the inverse,
  correlation.f:749-759.

What pins each function is in tests/test_oracle_gamma.py: the chi0 pair sum
of the augmented matrix [cprod | dcprod] (whose head and wing entries these
sums are), closed forms of the scaling and of the dielectric step, the
inverse of the bordered matrix [[head, conj(wing)], [wing^T, body]] (whose
Schur complement and off-diagonal blocks the reference's update forms), the
per-transition fold of the MULTDIFF spectral function, and MULTDIFF's own
purpose: when every transition's spectral weight sits at the mesh point
frqs(s) equal to its energy difference, the scaled column and the 1/frqs**l
weights cancel and the MULTDIFF matrix is the plain one (any wrong power,
sign or direction of the scaling breaks that).
"""
from __future__ import annotations

import math

import numpy as np

import oracle

FOURPI = 4 * math.pi


def transitions(nk, nocc, nunocc):
    """(k, b1, b2) of transition t = (k*nocc + b1)*nunocc + b2, the order of
    the reference loops k -> occupied -> unoccupied (susceptibility.f:861-866)."""
    t = np.arange(nk * nocc * nunocc)
    return t // (nocc * nunocc), (t // nunocc) % nocc, t % nunocc


def momentum_scale(mom, e_occ, e_unocc, vol):
    """dcprod of susceptibility.f:858-873 for multdiff_ = 0: mom (n, 3)
    complex, e_occ (nk, nocc), e_unocc (nk, nunocc) at k (q = 0).  The
    division of a complex by a real is taken componentwise (what the complex
    division gives for a zero imaginary divisor).  e_occ None: multdiff_ > 0,
    dcprod / sqrt(vol) only (:857)."""
    if e_occ is None:
        p = np.array(mom, complex)
        sv = math.sqrt(vol)
        return (p.real / sv) + 1j * (p.imag / sv)
    e_occ = np.asarray(e_occ, float)
    e_unocc = np.asarray(e_unocc, float)
    nk, nocc = e_occ.shape
    nunocc = e_unocc.shape[1]
    k, b1, b2 = transitions(nk, nocc, nunocc)
    sv = math.sqrt(vol)
    p = np.array(mom, complex)
    re = p.real / sv
    im = p.imag / sv
    d = e_unocc[k, b2] - e_occ[k, b1]
    ok = np.abs(d) > 1e-8
    out = np.zeros_like(p)
    out.real[ok] = re[ok] / d[ok, None]
    out.imag[ok] = im[ok] / d[ok, None]
    return out


def _sums(c, Dr, Di=None):
    """sus1, sus2 (and susc1, susc2) of spectral_function's HILBERT NONE
    branch: ddot(re, w), ddot(im, w) per frequency."""
    s1 = c.real @ Dr
    s2 = c.imag @ Dr
    if Di is None:
        return s1, s2
    return s1, s2, c.real @ Di, c.imag @ Di


def headwings_imag(cprod, dcprod, D, Dh=None):
    """Imaginary frequencies, susceptibility.f:1096-1127 (headh, wingh).
    cprod (n, nb), dcprod (n, 3), D (n, nw) weights of the wings, Dh those of
    the head (D if None; they differ under MULTDIFF, :1097 and :1130).
    Returns head (6, nw) packed upper and wing (3, nb, nw)."""
    cp = np.asarray(cprod, complex)
    dp = np.asarray(dcprod, complex)
    D = np.asarray(D, float)
    Dh = D if Dh is None else np.asarray(Dh, float)
    nb = cp.shape[1]
    nw = D.shape[1]
    head = np.zeros((6, nw), complex)
    for k2 in range(3):
        for k1 in range(k2 + 1):
            s1, s2 = _sums(dp[:, k1] * np.conj(dp[:, k2]), Dh)
            head[k2 * (k2 + 1) // 2 + k1] += s1 + 1j * s2
    wing = np.zeros((3, nb, nw), complex)
    for k2 in range(3):
        for i in range(nb):
            s1, s2 = _sums(cp[:, i] * np.conj(dp[:, k2]), D)
            wing[k2, i] += s1 + 1j * s2
    return head, wing


def headwings_real(cprod, dcprod, Dr, Di, Dhr=None, Dhi=None):
    """Real frequencies, susceptibility.f:1112-1125 and :1144-1151 (headc,
    wingc): complex weights Dr + i Di (wghtSr and wghtSi folded).  Returns
    head (3, 3, nw) full and wing (3, nb, 2, nw)."""
    cp = np.asarray(cprod, complex)
    dp = np.asarray(dcprod, complex)
    Dr = np.asarray(Dr, float)
    Di = np.asarray(Di, float)
    Dhr = Dr if Dhr is None else np.asarray(Dhr, float)
    Dhi = Di if Dhi is None else np.asarray(Dhi, float)
    nb = cp.shape[1]
    nw = Dr.shape[1]
    head = np.zeros((3, 3, nw), complex)
    for k2 in range(3):
        for k1 in range(k2 + 1):
            s1, s2, c1, c2 = _sums(dp[:, k1] * np.conj(dp[:, k2]), Dhr, Dhi)
            head[k1, k2] += s1 + 1j * s2 + 1j * c1 - c2
            if k1 != k2:
                head[k2, k1] += s1 - 1j * s2 + 1j * c1 + c2
    wing = np.zeros((3, nb, 2, nw), complex)
    for k2 in range(3):
        for i in range(nb):
            s1, s2, c1, c2 = _sums(cp[:, i] * np.conj(dp[:, k2]), Dr, Di)
            wing[k2, i, 0] += s1 + 1j * s2 + 1j * c1 - c2
            wing[k2, i, 1] += s1 - 1j * s2 + 1j * c1 + c2
    return head, wing


def spectral_function_mm(c, wtetra, frqs, wghtS, mm, wghtthr=0.0):
    """spectral_function with mm /= 0 (susceptibility.f:1793-1841) for one
    pair: reS(s) = sum over indexed t of re(t) wtetra(t, s), divided by
    frqs(s)**mm where frqs(s) > 0, then sum_s reS(s) wghtS(s, w) (and the
    same for im).  Returns the complex sus(w) = sus1 + i sus2."""
    wt = np.asarray(wtetra, float)
    frqs = np.asarray(frqs, float)
    nfrqs = wt.shape[1]
    reS = np.zeros(nfrqs)
    imS = np.zeros(nfrqs)
    for s in range(nfrqs):
        ix = np.nonzero(wt[:, s] > wghtthr)[0]
        reS[s] = np.dot(c.real[ix], wt[ix, s])
        imS[s] = np.dot(c.imag[ix], wt[ix, s])
        if mm != 0 and frqs[s] > 0:
            r = frqs[s] ** mm
            reS[s] /= r
            imS[s] /= r
    return reS @ wghtS + 1j * (imS @ wghtS)


def multdiff_scale(cprod, e_occ, e_unocc):
    """susceptibility.f:912-925: cprod(:,1) times the band-energy difference
    ene(j, k+q) - ene(i, k) of its transition (e_occ (nk, nocc) at k, e_unocc
    (nk, nunocc) at k+q); the other columns unchanged.  Returns a copy."""
    e_occ = np.asarray(e_occ, float)
    e_unocc = np.asarray(e_unocc, float)
    nk, nocc = e_occ.shape
    nunocc = e_unocc.shape[1]
    k, b1, b2 = transitions(nk, nocc, nunocc)
    cp = np.array(cprod, complex)
    for t in range(cp.shape[0]):
        cp[t, 0] = cp[t, 0] * (e_unocc[k[t], b2[t]] - e_occ[k[t], b1[t]])
    return cp


def chi0_multdiff(cprod, e_occ, e_unocc, wtetra, frqs, wghtSr, wghtSi=None, wghtthr=0.0):
    """The pair loop with MULTDIFF ON (multdiff_ = 1, susceptibility.f:588-592,
    912-925, 1159-1236): column 0 of cprod scaled by multdiff_scale, each pair
    i <= j with l = count([i, j] == first) passed to spectral_function_mm
    with mm = l.  wghtSi None: the imaginary-axis matrix (susceph, :1228-1230)
    as the full Hermitian (nb, nb, nw) array; else the real-frequency suscepc
    (:1231-1236) from sus1 + i sus2 (wghtSr) and susc1 + i susc2 (wghtSi).
    Literal loops: small sizes only."""
    cp = multdiff_scale(cprod, e_occ, e_unocc)
    nb = cp.shape[1]
    nw = np.asarray(wghtSr).shape[1]
    out = np.zeros((nb, nb, nw), complex)
    for j in range(nb):
        for i in range(j + 1):
            c = cp[:, i] * np.conj(cp[:, j])
            mm = int(i == 0) + int(j == 0)
            s = spectral_function_mm(c, wtetra, frqs, wghtSr, mm, wghtthr)
            if wghtSi is None:
                out[i, j] = s
                out[j, i] = np.conj(s)
                continue
            sc = spectral_function_mm(c, wtetra, frqs, wghtSi, mm, wghtthr)
            sus1, sus2, susc1, susc2 = s.real, s.imag, sc.real, sc.imag
            cdum = sus1 + 1j * susc1
            cdum1 = -susc2 + 1j * sus2
            out[i, j] += cdum + cdum1
            if i != j:
                out[j, i] += cdum - cdum1
    return out


def chi0_multdiff_inton(cprod, e_occ, e_unocc, wtetra, wghtS, wghtSi=None, wghtthr=0.0):
    """The pair loop with MULTDIFF INTON (multdiff_ = 2, with trafo;
    susceptibility.f:591, 1088-1090, 1159-1250): column 0 scaled by
    multdiff_scale; for ll = 0, 1, 2 the Hilbert weights of md = ll
    (wghtS[ll], (nfrqs, nw) each; wghtSi likewise for real frequencies);
    ll = 0 takes the pairs i <= j with neither index on the first basis
    function, ll > 0 the pairs with j = first (so only (first, first)), all
    with mm = 0.  Entries not visited stay 0.  Shapes and axes as
    chi0_multdiff."""
    cp = multdiff_scale(cprod, e_occ, e_unocc)
    nb = cp.shape[1]
    nfrqs = np.asarray(wtetra).shape[1]
    nw = np.asarray(wghtS[0]).shape[1]
    out = np.zeros((nb, nb, nw), complex)
    for ll in range(3):
        for j in range(nb):
            for i in range(j + 1):
                if ll == 0 and (i == 0 or j == 0):
                    continue
                if ll > 0 and j > 0:
                    continue
                c = cp[:, i] * np.conj(cp[:, j])
                s = spectral_function_mm(c, wtetra, np.zeros(nfrqs), wghtS[ll], 0, wghtthr)
                if wghtSi is None:
                    out[i, j] += s
                    if i != j:
                        out[j, i] += np.conj(s)
                    continue
                sc = spectral_function_mm(c, wtetra, np.zeros(nfrqs), wghtSi[ll], 0, wghtthr)
                cdum = s.real + 1j * sc.real
                cdum1 = -sc.imag + 1j * s.imag
                out[i, j] += cdum + cdum1
                if i != j:
                    out[j, i] += cdum - cdum1
    return out


def gamma_prepare(chi_packed, head, wing, sqcoul):
    """correlation.f:570-629 for ikpt = 1 on the imaginary axis.  chi_packed
    (npack, nw) = suscep, head (6, nw) packed, wing (3, nb, nw), sqcoul (nb).
    Returns (eps_packed, head, wing): the first basis function's row of the
    packed matrix and wing(:,1,:) set to zero (:597-607), eps = 1 - v^1/2
    chi v^1/2 (:612-621), head = identityp(3) - 4 pi head (:624-626) and
    wing(:,i,:) = -sqrt(4 pi) sqcoul(i) wing(:,i,:) (:627-629)."""
    sqcoul = np.asarray(sqcoul, float)
    nb = sqcoul.size
    chi = np.array(chi_packed, complex)
    j = 0
    for i in range(1, nb + 1):  # j = 1, 2, 4, 7, ... (1-based): suscep(1, i)
        chi[j] = 0
        j += i
    w = np.array(wing, complex)
    w[:, 0, :] = 0
    nw = chi.shape[1]
    eps = np.stack([oracle.dielectric_diag(chi[:, f], sqcoul) for f in range(nw)], axis=1)
    idp = np.zeros(6)
    idp[[0, 2, 5]] = 1.0  # identityp(3): packed diagonal
    h = idp[:, None] - FOURPI * np.asarray(head, complex)
    w = -math.sqrt(FOURPI) * sqcoul[None, :, None] * w
    return eps, h, w


def gamma_inverse(einv_packed, head, wing):
    """Head and wings of the inverse, correlation.f:749-759 (imaginary
    axis): wing2 = conj(wing); wing = -wing conj(einv); head = head +
    packmat(wing2 wing^T).  einv_packed (npack, nw), head (6, nw), wing
    (3, nb, nw).  Returns (head, wing)."""
    einv_packed = np.asarray(einv_packed, complex)
    w = np.array(wing, complex)
    h = np.array(head, complex)
    nb = w.shape[1]
    nw = w.shape[2]
    iu = [(k1, k2) for k2 in range(3) for k1 in range(k2 + 1)]
    for f in range(nw):
        e = oracle.unpack_herm(einv_packed[:, f], nb)
        w2 = np.conj(w[:, :, f])
        w[:, :, f] = -w[:, :, f] @ np.conj(e)
        m = w2 @ w[:, :, f].T
        h[:, f] += np.array([m[k1, k2] for k1, k2 in iu])
    return h, w


def gamma_prepare_cplx(chi, head, wing, sqcoul):
    """Real frequencies, correlation.f:597-607 and :632-646: chi (nb, nb,
    nw) = suscepc with row and column 1 set to zero, eps = 1 - sqcoul chi
    sqcoul (p_diagmatdiag(-sqcoul, ., sqcoul), p_plusdiag), headc =
    identity(3) - 4 pi headc (3, 3, nw), wingc(:,i,:,:) = -sqrt(4 pi)
    sqcoul(i) wingc (3, nb, 2, nw) after wingc(:,1,:,:) = 0."""
    sq = np.asarray(sqcoul, float)
    c = np.array(chi, complex)
    c[0, :, :] = 0
    c[:, 0, :] = 0
    eps = np.eye(sq.size)[:, :, None] + (-sq)[:, None, None] * c * sq[None, :, None]
    h = np.eye(3)[:, :, None] - FOURPI * np.asarray(head, complex)
    w = np.array(wing, complex)
    w[:, 0] = 0
    w = -math.sqrt(FOURPI) * sq[None, :, None, None] * w
    return eps, h, w


def gamma_inverse_cplx(einv, head, wing):
    """correlation.f:760-763 per real frequency: wingc(:,:,1) =
    -matmat(wingc(:,:,1), transpose(dielecc)); headc += matmat(wingc(:,:,2),
    transpose(wingc(:,:,1))) (the new first wing); wingc(:,:,2) =
    -matmat(wingc(:,:,2), dielecc).  einv (nb, nb, nw)."""
    h = np.array(head, complex)
    w = np.array(wing, complex)
    for f in range(w.shape[3]):
        e = np.asarray(einv[:, :, f], complex)
        w[:, :, 0, f] = -w[:, :, 0, f] @ e.T
        h[:, :, f] = h[:, :, f] + w[:, :, 1, f] @ w[:, :, 0, f].T
        w[:, :, 1, f] = -w[:, :, 1, f] @ e
    return h, w


def head_scalar(head_packed, vec=None):
    """The head's scalar for eps(1,1) (correlation.f:656-665, :683):
    trace(head)/3, or dotprod(vec, matvec(head, vec)) for a unit vec.
    head_packed (6,) packed upper."""
    h = unpack_head(head_packed)
    if vec is None:
        return np.trace(h) / 3
    vec = np.asarray(vec, float)
    return vec @ h @ vec


def unpack_head(hp):
    """(6,) packed upper -> (3, 3) Hermitian."""
    h = np.zeros((3, 3), complex)
    ij = 0
    for k2 in range(3):
        for k1 in range(k2 + 1):
            h[k1, k2] = hp[ij]
            h[k2, k1] = np.conj(hp[ij])
            ij += 1
    return h
