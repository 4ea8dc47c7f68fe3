"""Tetrahedron weights of the Hilbert mesh -- TEST INFRASTRUCTURE, part of
the oracle; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import it.

Restates, in plain numpy/Python:
- get_tetra (tetrahedron.f:219-271, mode 1): the six tetrahedra of a k-mesh
  cube around its shortest main diagonal;
- cuttetraN (tetrahedron.f:347-412): the corner weights of the part of a
  tetrahedron where a linear function lies below a threshold;
- tetrahedron5_init (tetrahedron.f:903-1332) for a full cubic k mesh
  without symmetry reduction (kpt1p the identity, kflag = 1) and bands that
  the Fermi energy does not cut (insulators: case (c0) of the reference):
  the weight of each transition t = (k, b1, b2) in the frequency bin
  [w(i) - dw(i), w(i) + dw(i)] of the Hilbert mesh, divided by 2 dw(i).
  Degenerate-state averaging (:1233-1273) is not part of it;
- the same with bands cut by the Fermi energy (metals, case (c1)): each
  tetrahedron restricted to its occupied (e_occ < efermi) and unoccupied
  (e_unocc > efermi) part through cuttetraN's sub-tetrahedra (:1045-1129),
  then cut by d < frq1(i) (:1176-1202);
- the averaging of the weights over degenerate states (:1233-1273), for
  groups inside the band windows;
- the weights of a symmetry-reduced k set (kpt1p): the sum over the
  full-mesh points that map to each parent;
- the final averaging over equivalent k points of kpt1 under the symmetries
  that keep q (:1282-1318);
- the TETRAF rows of a reduced set (:1092-1096, 1212-1216);
- band windows that cut a degenerate group (:984-1012, 1191-1194,
  1237-1273): widened windows, clamped rows and the whole-group average;
- gauss4_init (gauss.f:120-186), the Gaussian weights of the GAUSS key.

k points are numbered k = (i1*N2 + i2)*N3 + i3 for the mesh point
(i1, i2, i3)/(N1, N2, N3); cube corners 1..8 are (0,0,0), (1,0,0),
(0,1,0), (1,1,0), (0,0,1), (1,0,1), (0,1,1), (1,1,1) offsets, as kcube in
tetrahedron.f:1014-1021.

What pins each function is in tests/test_oracle_tetra.py: exact
integration of the barycentric coordinates over the cut tetrahedron
(independent subdivision quadrature), the volume and sum rules, flat bands
(all weight in one bin), and the total weight of the Brillouin zone; the
Gaussian weights by the quadrature of the Gaussian over each bin, the
telescoping sum over the tiled bins of the exponential mesh (the Gaussian
mass inside the mesh times the occupation factors) and star sums.
"""
from __future__ import annotations

import math

import numpy as np

ETA = 1e-14  # cuttetraN's tolerance (tetrahedron.f:355)

# the four decompositions of get_tetra, 1-based corner numbers, one row of
# four corners per tetrahedron (tetrahedron.f:253-268)
_DECOMP = [
    [[1, 2, 3, 6], [5, 7, 3, 6], [1, 5, 3, 6], [2, 4, 3, 6], [4, 8, 3, 6], [7, 8, 3, 6]],  # 3-6
    [[5, 6, 2, 7], [1, 5, 2, 7], [1, 3, 2, 7], [8, 6, 2, 7], [4, 3, 2, 7], [8, 4, 2, 7]],  # 2-7
    [[2, 6, 1, 8], [2, 4, 1, 8], [3, 4, 1, 8], [3, 7, 1, 8], [5, 7, 1, 8], [5, 6, 1, 8]],  # 1-8
    [[2, 6, 4, 5], [1, 2, 4, 5], [1, 3, 4, 5], [3, 7, 4, 5], [7, 8, 4, 5], [6, 8, 4, 5]],  # 4-5
]


def tetra_choice(rlat, nkpt3):
    """Index (0..3) of the decomposition get_tetra(mode 1) picks: the
    shortest of the diagonals 3-6, 2-7, 1-8, 4-5 of the mesh cell spanned by
    rlat(:,i)/nkpt3(i) (first on ties, like minloc)."""
    rl = np.asarray(rlat, float) / np.asarray(nkpt3, float)[None, :]
    r1, r2, r3 = rl[:, 0], rl[:, 1], rl[:, 2]
    diag = [np.dot(v, v) for v in (r1 + r3 - r2, r2 + r3 - r1, r1 + r2 + r3, r1 + r2 - r3)]
    return int(np.argmin(diag))


def get_tetra(rlat, nkpt3):
    """(6, 4) zero-based corner indices of the six tetrahedra."""
    return np.array(_DECOMP[tetra_choice(rlat, nkpt3)]) - 1


def cuttetraN(etetra_in, ef):
    """(wtetra(4), vtetra(3), ntetra) of cuttetraN (tetrahedron.f:347-412):
    the corner weights (barycentric coordinates integrated over the part of
    the tetrahedron where the linear function with corner values etetra_in
    is below ef, scaled so that the whole tetrahedron gives 1 per corner)."""
    e_in = np.asarray(etetra_in, float)
    # pnt: the reference's size ordering (minloc picks the first minimum)
    pnt = np.full(4, 4)
    pnt[int(np.argmin(e_in))] = 1
    m = np.where(pnt > 1, e_in, np.inf)
    pnt[int(np.argmin(m))] = 2
    m = np.where(pnt > 2, e_in, np.inf)
    pnt[int(np.argmin(m))] = 3
    # pnt(pnt) = [1,2,3,4]: invert the rank map
    order = np.empty(4, int)
    order[pnt - 1] = np.arange(4)
    e = e_in[order]
    p = order  # p[j] = position (0-based) of the (j+1)-th smallest value
    w = np.zeros(4)
    v = np.zeros(3)
    if ef >= e[3] - ETA:
        return np.ones(4), np.array([1.0, 0.0, 0.0]), 1
    if ef >= e[2]:
        f = (ef - e[:3]) / (e[3] - e[:3])
        v[0] = f[1]
        v[1] = f[0] * (1 - f[1]) * (1 - f[2])
        v[2] = f[2] * (1 - f[1])
        w[p[0]] = v[0] + v[1] * (2 - f[0]) + v[2]
        w[p[1]] = v[0] + (v[0] + v[1] + v[2]) * (1 - f[1])
        w[p[2]] = v[0] + v[1] * (1 - f[2]) + v[2] * (2 - f[2])
        w[p[3]] = v[0] * f[1] + v[1] * f[0] + (v[1] + v[2]) * (f[1] + f[2])
        return w, v, 3
    if ef > e[1]:
        f = np.empty(4)
        f[:2] = (ef - e[0]) / (e[2:4] - e[0])
        f[2:] = (ef - e[1]) / (e[2:4] - e[1])
        v[0] = f[0] * f[1]
        v[1] = f[0] * (1 - f[1]) * f[3]
        v[2] = f[2] * (1 - f[0]) * f[3]
        w[p[0]] = v[0] * (3 - f[0] - f[1]) + v[1] * (2 - f[0] - f[1]) + v[2] * (1 - f[0])
        w[p[1]] = v[0] + v[1] * (2 - f[3]) + v[2] * (3 - f[2] - f[3])
        w[p[2]] = (v[0] + v[1] + v[2]) * f[0] + v[2] * f[2]
        w[p[3]] = (v[0] + v[1]) * f[1] + (v[1] + v[2]) * f[3]
        return w, v, 3
    if ef >= e[0] + ETA:
        f = (ef - e[0]) / (e[1:4] - e[0])
        v[0] = f[0] * f[1] * f[2]
        w[p[0]] = v[0] * (4 - f[0] - f[1] - f[2])
        w[p[1]] = v[0] * f[0]
        w[p[2]] = v[0] * f[1]
        w[p[3]] = v[0] * f[2]
        return w, v, 1
    return w, v, 0


def cuttetraN_sub(etetra_in, ef):
    """cuttetraN with its optional tetra output (tetrahedron.f:347-412):
    (wtetra(4), vtetra(3), ntetra, T(ntetra, 4, 4)) where T[m][i, j] is the
    barycentric coordinate, on input corner i, of corner j of sub-tetrahedron
    m of the part below ef, and vtetra(m) its volume fraction."""
    e_in = np.asarray(etetra_in, float)
    w, v, nt = cuttetraN(e_in, ef)
    pnt = np.full(4, 4)
    pnt[int(np.argmin(e_in))] = 1
    m = np.where(pnt > 1, e_in, np.inf)
    pnt[int(np.argmin(m))] = 2
    m = np.where(pnt > 2, e_in, np.inf)
    pnt[int(np.argmin(m))] = 3
    order = np.empty(4, int)
    order[pnt - 1] = np.arange(4)
    e = e_in[order]
    if nt == 0:
        return w, v, 0, np.zeros((0, 4, 4))
    if ef >= e[3] - ETA:
        cols = [[[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]]
    elif ef >= e[2]:
        f1, f2, f3 = (ef - e[:3]) / (e[3] - e[:3])
        cols = [[[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 1 - f2, 0, f2]],
                [[1, 0, 0, 0], [1 - f1, 0, 0, f1], [0, 1 - f2, 0, f2], [0, 0, 1 - f3, f3]],
                [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1 - f2, 0, f2], [0, 0, 1 - f3, f3]]]
    elif ef > e[1]:
        f1, f2 = (ef - e[0]) / (e[2:4] - e[0])
        f3, f4 = (ef - e[1]) / (e[2:4] - e[1])
        cols = [[[1, 0, 0, 0], [0, 1, 0, 0], [1 - f1, 0, f1, 0], [1 - f2, 0, 0, f2]],
                [[0, 1, 0, 0], [1 - f1, 0, f1, 0], [1 - f2, 0, 0, f2], [0, 1 - f4, 0, f4]],
                [[0, 1, 0, 0], [1 - f1, 0, f1, 0], [0, 1 - f4, 0, f4], [0, 1 - f3, f3, 0]]]
    else:
        f1, f2, f3 = (ef - e[0]) / (e[1:4] - e[0])
        cols = [[[1, 0, 0, 0], [1 - f1, f1, 0, 0], [1 - f2, 0, f2, 0], [1 - f3, 0, 0, f3]]]
    T = np.zeros((nt, 4, 4))
    for mm in range(nt):
        X = np.array(cols[mm], float).T  # X[i_sorted, j]: column j lists sorted-corner coords
        T[mm][order, :] = X  # tetra(pnt, :, m) = X
    return w, v, nt, T


def _restricted(e1t, e2t, ef, cut1, cut2):
    """The part of one tetrahedron with occupied e1 < ef and unoccupied
    e2 > ef as (wtetra_(4), [(v, T)]) -- the tetra_/vtetra_/wtetra_ of
    tetrahedron5_init (:1100-1129).  cut1/cut2: the cube's bands are cut."""
    if cut1 and not cut2:
        w, v, nt, T = cuttetraN_sub(e1t, ef)
        return w, [(v[m], T[m]) for m in range(nt)]
    if cut2 and not cut1:
        w, v, nt, T = cuttetraN_sub(-np.asarray(e2t), -ef)
        return w, [(v[m], T[m]) for m in range(nt)]
    _, v2, nt2, T2 = cuttetraN_sub(-np.asarray(e2t), -ef)
    w_ = np.zeros(4)
    subs = []
    for m2 in range(nt2):
        e1s = np.asarray(e1t) @ T2[m2]
        w, v, nt, T1 = cuttetraN_sub(e1s, ef)
        w_ += v2[m2] * (T2[m2] @ w)
        for m1 in range(nt):
            subs.append((v2[m2] * v[m1], T2[m2] @ T1[m1]))
    return w_, subs


def corner_weights_at(dt, part, f):
    """One tetrahedron's contribution to its 4 corners at the bin edge f
    (tetrahedron.f:1160-1202): dt the band-energy differences at its
    corners, part the (wtetra_, sub-tetrahedra) of its occupied-unoccupied
    part when the cube's bands are cut (None: not cut, the whole
    tetrahedron, case (c0))."""
    dt = np.asarray(dt, float)
    if part is None:
        if np.all(dt >= f):
            return np.zeros(4)
        if np.all(dt < f):
            return np.ones(4)
        return cuttetraN(dt, f)[0]
    w_, subs = part
    if not subs or np.all(dt >= f):
        return np.zeros(4)
    if np.all(dt < f):
        return w_
    out = np.zeros(4)
    for v, T in subs:
        ds_ = dt @ T
        if np.all(ds_ >= f):
            continue
        if np.all(ds_ < f):
            out += v * T.sum(1)
        else:
            out += v * (T @ cuttetraN(ds_, f)[0])
    return out


def tetrahedron5_weights_metal(nkpt3, rlat, e_occ, e_unocc, efermi, frq):
    """tetrahedron5_init with bands cut by the Fermi energy (cases (c0) and
    (c1), tetrahedron.f:1045-1213), full cubic mesh, kflag = 1, no symmetry;
    same output layout as tetrahedron5_weights.  Literal loop order."""
    e_occ = np.asarray(e_occ, float)
    e_unocc = np.asarray(e_unocc, float)
    nk, nocc = e_occ.shape
    nunocc = e_unocc.shape[1]
    n1m, n2m, n3m = nkpt3
    tetra = get_tetra(rlat, nkpt3)
    frq1, _ = hilbert_bins(frq)
    nfrq = frq1.size
    pk = mesh_index(nkpt3)
    W = np.zeros((nk, nocc, nunocc, nfrq))
    cnt = np.array([np.count_nonzero(tetra == i) for i in range(8)], float)
    for i1 in range(n1m):
        for i2 in range(n2m):
            for i3 in range(n3m):
                kcube = [pk(i1 + d1, i2 + d2, i3 + d3)
                         for d3 in (0, 1) for d2 in (0, 1) for d1 in (0, 1)]
                for b2 in range(nunocc):
                    ecube2 = e_unocc[kcube, b2]
                    if np.all(ecube2 < efermi):
                        continue
                    cut2 = ecube2.min() < efermi
                    for b1 in range(nocc):
                        ecube1 = e_occ[kcube, b1]
                        if np.all(ecube1 > efermi):
                            continue
                        cut1 = ecube1.max() > efermi
                        dcube = ecube2 - ecube1
                        dmin, dmax = dcube.min(), dcube.max()
                        if cut1 or cut2:
                            parts = [_restricted(ecube1[tetra[it]], ecube2[tetra[it]], efermi,
                                                 cut1, cut2) for it in range(6)]
                            wadd0 = np.zeros(8)
                            for it in range(6):
                                wadd0[tetra[it]] += parts[it][0]
                        else:
                            wadd0 = cnt
                        for ifr in range(nfrq):
                            f = frq1[ifr]
                            if f < dmin:
                                continue
                            if f > dmax:
                                wadd = wadd0
                            else:
                                wadd = np.zeros(8)
                                for it in range(6):
                                    wadd[tetra[it]] += corner_weights_at(
                                        dcube[tetra[it]], parts[it] if (cut1 or cut2) else None, f)
                            for c in range(8):
                                W[kcube[c], b1, b2, ifr] += wadd[c]
    rdum = 4.0 * 6 * nk
    out = np.empty_like(W)
    for ifr in range(nfrq - 1, 0, -1):
        out[..., ifr] = (W[..., ifr] - W[..., ifr - 1]) / (2 * (frq1[ifr] - frq[ifr]) * rdum)
    out[..., 0] = W[..., 0] / (2 * (frq1[0] - frq[0]) * rdum)
    if frq[0] == 0:
        out[..., 0] = 0
    return out.reshape(nk * nocc * nunocc, nfrq)


def reduce_to_parents(w, nk, nocc, nunocc, kpt1p, nkpt1):
    """The k points of a symmetry-reduced set: tetrahedron5_init scatters a
    cube corner k to its parent kpt1p(k) and skips corners without one
    (tetrahedron.f:1022, :1092-1095, :1207-1210), so the row of a parent is
    the sum of the full-mesh rows of the points that map to it.  w (n_full,
    nfrqs); kpt1p (nk,) 0-based parent or -1.  Returns (nkpt1*nocc*nunocc,
    nfrqs), in the loop order of the reference's scatter (k ascending)."""
    W = np.asarray(w, float).reshape(nk, nocc, nunocc, -1)
    out = np.zeros((nkpt1, nocc, nunocc, W.shape[-1]))
    for k in range(nk):
        if kpt1p[k] >= 0:
            out[kpt1p[k]] += W[k]
    return out.reshape(nkpt1 * nocc * nunocc, -1)


def reduce_tetraf(w, nk, nocc, nunocc, kpt1, kpt1p, nkpt1):
    """The rows of a reduced k set under TETRAF (getinput.f:1605): a cube
    corner k contributes only if k is itself listed in kpt1 (tetrahedron.f:
    1092-1096), to its parent kpt1p(k), and the row of parent p is then
    multiplied by count(kpt1p == p) (:1212-1216).  w (n_full, nfrqs);
    kpt1 (nkpt1,) full-mesh indices; kpt1p (nk,) 0-based parent or -1."""
    W = np.asarray(w, float).reshape(nk, nocc, nunocc, -1)
    listed = set(int(k) for k in kpt1)
    out = np.zeros((nkpt1, nocc, nunocc, W.shape[-1]))
    for k in range(nk):
        if kpt1p[k] >= 0 and k in listed:
            out[kpt1p[k]] += W[k]
    for p in range(nkpt1):
        out[p] *= np.count_nonzero(np.asarray(kpt1p) == p)
    return out.reshape(nkpt1 * nocc * nunocc, -1)


def _group(deg_row, b):
    """(first, last) of band b's degenerate group from SPEX's deg convention:
    deg(b) is the last band of the group when b is its first band, else the
    first band (global.f deg; tetrahedron.f:1237-1240)."""
    d = int(deg_row[b])
    if d >= b:
        return b, d
    return d, int(deg_row[d])


def average_degenerate(w, nk, nocc, nunocc, deg_occ, deg_unocc):
    """Degenerate-state averaging of tetrahedron5_init (tetrahedron.f:
    1233-1273) for groups inside the band windows: first over each group of
    occupied states at k (all unoccupied b2, all bins), then over each group
    of unoccupied states at k+q.  w (n, nfrqs) with t = (k*nocc + b1)*nunocc
    + b2; deg_occ (nk, nocc), deg_unocc (nk, nunocc) window-relative."""
    W = np.array(w, float).reshape(nk, nocc, nunocc, -1)
    for k in range(nk):
        n = _group(deg_occ[k], 0)[0]
        while n < nocc:
            first, last = _group(deg_occ[k], n)
            W[k, first:last + 1] = W[k, first:last + 1].sum(0) / (last - first + 1)
            n = last + 1
    for k in range(nk):
        n = _group(deg_unocc[k], 0)[0]
        while n < nunocc:
            first, last = _group(deg_unocc[k], n)
            W[k, :, first:last + 1] = (W[k, :, first:last + 1].sum(1) / (last - first + 1))[:, None]
            n = last + 1
    return W.reshape(nk * nocc * nunocc, -1)


def tetrahedron5_weights_bands(nkpt3, rlat, ene1, ene2, deg1, deg2, b1low, b1up, b2low, b2up,
                               efermi, frq, kpt1=None, kpt1p=None, nkpt1=None, tetraf=False,
                               alignbd=False):
    """tetrahedron5_init with band windows that may cut a degenerate group
    (tetrahedron.f:984-1012, 1080-1093, 1191-1194, 1237-1273), kflag 1.
    ene1 (nk, nb1) at k and ene2 (nk, nb2) at k+q hold all bands, deg1 and
    deg2 the SPEX deg tables of the same bands (0-based, SPEX's convention)
    at k and k+q; windows b1low..b1up, b2low..b2up 0-based inclusive.

    Per mesh point k the windows are widened to the whole groups of their
    edge bands (d1low(k) .. d1up(k), d2low(k) .. d2up(k)); a transition (n1,
    n2) of the widened windows adds its weight to the row of the window band
    nearest to it (n clamped into the window) of its corner k; the rows of a
    reduced set are formed from those (reduce_to_parents, or reduce_tetraf
    under TETRAF); then the reference's averaging over the groups at kpt1(p)
    divides each group's in-window sum by the whole group's size (skipped
    under alignbd: tetrahedron5_init returns first, :1231).  Returns
    (nkr*nocc*nunocc, nfrqs), nkr = nkpt1 or nk."""
    ene1, ene2 = np.asarray(ene1, float), np.asarray(ene2, float)
    nk = ene1.shape[0]
    nocc, nunocc = b1up - b1low + 1, b2up - b2low + 1
    lo1 = np.array([_group(deg1[k], b1low)[0] for k in range(nk)])
    up1 = np.array([_group(deg1[k], b1up)[1] for k in range(nk)])
    lo2 = np.array([_group(deg2[k], b2low)[0] for k in range(nk)])
    up2 = np.array([_group(deg2[k], b2up)[1] for k in range(nk)])
    a1, z1, a2, z2 = lo1.min(), up1.max(), lo2.min(), up2.max()
    nx1, nx2 = z1 - a1 + 1, z2 - a2 + 1
    Wx = tetrahedron5_weights_metal(nkpt3, rlat, ene1[:, a1:z1 + 1], ene2[:, a2:z2 + 1], efermi,
                                    frq).reshape(nk, nx1, nx2, -1)
    acc = np.zeros((nk, nocc, nunocc, Wx.shape[-1]))
    for k in range(nk):
        for n1 in range(lo1[k], up1[k] + 1):
            r1 = min(b1up, max(b1low, n1)) - b1low
            for n2 in range(lo2[k], up2[k] + 1):
                r2 = min(b2up, max(b2low, n2)) - b2low
                acc[k, r1, r2] += Wx[k, n1 - a1, n2 - a2]
    acc = acc.reshape(nk * nocc * nunocc, -1)
    if kpt1p is None:
        rows, kd = acc, np.arange(nk)
    else:
        kd = np.asarray(kpt1)
        if tetraf:
            rows = reduce_tetraf(acc, nk, nocc, nunocc, kpt1, kpt1p, nkpt1)
        else:
            rows = reduce_to_parents(acc, nk, nocc, nunocc, kpt1p, nkpt1)
    if alignbd:
        return rows
    W = rows.reshape(len(kd), nocc, nunocc, -1).copy()
    for r, k in enumerate(kd):  # occupied groups (:1237-1255)
        n = _group(deg1[k], b1low)[0]
        while n <= b1up:
            nn = _group(deg1[k], n)[1]
            lo, hi = max(b1low, n), min(b1up, nn)
            if lo <= hi:
                W[r, lo - b1low:hi - b1low + 1] = (W[r, lo - b1low:hi - b1low + 1].sum(0)
                                                   / (nn - n + 1))
            n = nn + 1
    for r, k in enumerate(kd):  # unoccupied groups (:1257-1273)
        n = _group(deg2[k], b2low)[0]
        while n <= b2up:
            nn = _group(deg2[k], n)[1]
            lo, hi = max(b2low, n), min(b2up, nn)
            if lo <= hi:
                W[r, :, lo - b2low:hi - b2low + 1] = (W[r, :, lo - b2low:hi - b2low + 1].sum(1)
                                                      / (nn - n + 1))[:, None]
            n = nn + 1
    return W.reshape(len(kd) * nocc * nunocc, -1)


def gauss4_weights(e_occ, e_unocc, efermi, sigma, frq, kpt1p=None, nkpt1=None):
    """gauss4_init (gauss.f:120-186), kflag 1, in the reference's loop order:
    for each mesh point k with a parent row p (kpt1p[k] >= 0; all k with
    kpt1p None), each transition (b1, b2) and each bin i with frq(i) /= 0,
      w(p,b1,b2,i) += r1 r2 g2fac factor (erf(gfac (f2 - d)) - erf(gfac (f1 -
      d))) / (2 df),
    r1 = (1 + erf((efermi - e1) / (sqrt2 sigma1))) / 2, r2 = (1 + erf((e2 -
    efermi) / (sqrt2 sigma1))) / 2, d = e2 - e1, gfac = 1/(sqrt2 sigma2),
    g2fac = 1/(sqrt(2 pi) sigma2), factor = sqrt(pi/2) sigma2, the bin [f1,
    f2] = frq(i) -/+ df with df = a (frq(i+1) - frq(i)) (last bin: (1 - a)
    (frq(i) - frq(i-1))) and a the mesh factor of :137-147; then / nk.
    e_occ (nk, nocc) at k, e_unocc (nk, nunocc) at k+q.  Returns
    (nrow*nocc*nunocc, nfrqs)."""
    eo, eu = np.asarray(e_occ, float), np.asarray(e_unocc, float)
    frq = np.asarray(frq, float)
    nk, nocc = eo.shape
    nunocc = eu.shape[1]
    nf = len(frq)
    rdum, a = 0.0, 0.0
    for i in range(1, nf):
        if frq[i] <= frq[i - 1]:
            raise ValueError("frequency mesh not ascending")
        if i != nf - 1:
            rdum += (frq[i] - frq[i - 1]) / (frq[i + 1] - frq[i - 1])
            r = rdum / i
            if i > 1 and abs(r - a) > 1e-10:
                raise ValueError("mesh not exponential")
            a = r
    s1, s2 = sigma
    sqrt2 = math.sqrt(2.0)
    gfac = 1.0 / (sqrt2 * s2)
    g2fac = 1.0 / (math.sqrt(2 * math.pi) * s2)
    factor = math.sqrt(0.5 * math.pi) * s2
    nrow = nk if kpt1p is None else nkpt1
    w = np.zeros((nrow, nocc, nunocc, nf))
    for k in range(nk):
        p = k if kpt1p is None else int(kpt1p[k])
        if p < 0:
            continue
        for b1 in range(nocc):
            r1 = (1 + math.erf((efermi - eo[k, b1]) / (sqrt2 * s1))) / 2
            for b2 in range(nunocc):
                r2 = (1 + math.erf((eu[k, b2] - efermi) / (sqrt2 * s1))) / 2
                d = eu[k, b2] - eo[k, b1]
                for i in range(nf):
                    if frq[i] == 0:
                        continue
                    df = a * (frq[i + 1] - frq[i]) if i != nf - 1 else (1 - a) * (frq[i] - frq[i - 1])
                    f1, f2 = frq[i] - df, frq[i] + df
                    w[p, b1, b2, i] += (r1 * r2 * g2fac * factor
                                        * (math.erf(gfac * (f2 - d)) - math.erf(gfac * (f1 - d)))
                                        / (2 * df))
    return (w / nk).reshape(nrow * nocc * nunocc, nf)


def equivalent_groups(kpt1, kpt1p, kptsym, ikpt):
    """The groups of tetrahedron5_init's final averaging (tetrahedron.f:
    1282-1318), in the reference's loop order: for each k of kpt1 not yet
    done, the symmetries that keep q (kptsym(ikpt, isym) == ikpt) map
    kpt1(k) to points whose kpt1p indices k1 are counted; if one k1 differs
    from k, the k1 with a count form a group (ascending) that is averaged
    with the counts as weights and marked done.  0-based: kpt1 (nkpt1,)
    full-mesh indices, kpt1p (nk,) index into kpt1 or -1, kptsym (nk, nsym)
    image of each point under each symmetry.  Returns [(members, counts)]."""
    kptsym = np.asarray(kptsym)
    nkpt1 = len(kpt1)
    little = [s for s in range(kptsym.shape[1]) if kptsym[ikpt, s] == ikpt]
    done = np.zeros(nkpt1, bool)
    groups = []
    for k in range(nkpt1):
        if done[k]:
            continue
        cnt = np.zeros(nkpt1, int)
        moved = False
        for s in little:
            k1 = kpt1p[kptsym[kpt1[k], s]]
            if k1 >= 0:
                cnt[k1] += 1
                moved = moved or k1 != k
        if moved:
            mem = np.flatnonzero(cnt > 0)
            done[mem] = True
            groups.append((mem, cnt[mem]))
    return groups


def average_equivalent(w, nocc, nunocc, kpt1, kpt1p, kptsym, ikpt):
    """tetrahedron.f:1282-1318: the weights of each group of equivalent k
    points replaced by their count-weighted mean, summed member by member
    in ascending order (rdum = rdum + d1up*w, unfused), then divided by the
    sum of the counts.  Groups are applied one after another, as the
    reference's loop does.  w (nkpt1*nocc*nunocc, nfrqs), t = (k*nocc +
    b1)*nunocc + b2."""
    W = np.array(w, float).reshape(len(kpt1), nocc, nunocc, -1)
    for mem, cnt in equivalent_groups(kpt1, kpt1p, kptsym, ikpt):
        acc = np.zeros(W.shape[1:])
        for m, c in zip(mem, cnt):
            acc = acc + float(c) * W[m]
        W[mem] = acc / float(cnt.sum())
    return W.reshape(-1, W.shape[-1])


def deg_from_energies(e, tol):
    """SPEX-style deg array (window-relative) of sorted energies e (nk, nb):
    bands within tol of their neighbour form a group."""
    e = np.asarray(e, float)
    nk, nb = e.shape
    deg = np.zeros((nk, nb), int)
    for k in range(nk):
        b = 0
        while b < nb:
            c = b
            while c + 1 < nb and abs(e[k, c + 1] - e[k, c]) <= tol:
                c += 1
            deg[k, b] = c
            for x in range(b + 1, c + 1):
                deg[k, x] = b
            b = c + 1
    return deg


def hilbert_bins(frq):
    """(frq1, dw) of tetrahedron5_init (tetrahedron.f:951-975): the upper
    bin edges frq1(i) = w(i) + dw(i) and the half widths dw(i) of an
    exponential mesh (a checked constant, dw(i) = a (w(i+1) - w(i)))."""
    frq = np.asarray(frq, float)
    nfrq = frq.size
    if nfrq < 2:
        raise ValueError("Number of Hilbert frequencies must be larger than one.")
    if np.any(np.diff(frq) <= 0):
        raise ValueError("input frequency mesh not properly ordered.")
    rdum = 0.0
    for i in range(1, nfrq - 1):
        a = (frq[i] - frq[i - 1]) / (frq[i + 1] - frq[i - 1])
        rdum += a
        if abs(rdum / i - a) > 1e-10:
            raise ValueError("Hilbert mesh not exponential.")
    a = rdum / (nfrq - 2) if nfrq > 2 else 0.5
    frq1 = np.empty(nfrq)
    frq1[:-1] = frq[:-1] + a * (frq[1:] - frq[:-1])
    frq1[-1] = frq[-1] + (1 - a) * (frq[-1] - frq[-2])
    return frq1, frq1 - frq


def mesh_index(nkpt3):
    n1, n2, n3 = nkpt3

    def pk(i1, i2, i3):
        return ((i1 % n1) * n2 + (i2 % n2)) * n3 + (i3 % n3)

    return pk


def tetrahedron5_weights(nkpt3, rlat, e_occ, e_unocc, efermi, frq):
    """wintgr5 of tetrahedron5_init for the whole packet, as (n, nfrqs) with
    t = (k*nocc + b1)*nunocc + b2 fastest -- the wtetra(n, nfrqs) of
    susceptibility.f:936-948.  e_occ (nk, nocc) at k, e_unocc (nk, nunocc)
    at k + q, indexed by k.  Literal loop order of the reference (cubes k,
    then n2, n1, frequencies; tetrahedra in order within a cube)."""
    e_occ = np.asarray(e_occ, float)
    e_unocc = np.asarray(e_unocc, float)
    nk, nocc = e_occ.shape
    nunocc = e_unocc.shape[1]
    n1m, n2m, n3m = nkpt3
    assert n1m * n2m * n3m == nk
    if e_occ.max() > efermi or e_unocc.min() < efermi:
        raise ValueError("a band is cut by the Fermi energy (metal): not covered")
    tetra = get_tetra(rlat, nkpt3)
    frq1, dw = hilbert_bins(frq)
    nfrq = frq1.size
    pk = mesh_index(nkpt3)
    W = np.zeros((nk, nocc, nunocc, nfrq))
    wadd0 = np.array([np.count_nonzero(tetra == i) for i in range(8)], float)
    for i1 in range(n1m):
        for i2 in range(n2m):
            for i3 in range(n3m):
                kcube = [pk(i1 + d1, i2 + d2, i3 + d3)
                         for d3 in (0, 1) for d2 in (0, 1) for d1 in (0, 1)]
                for b2 in range(nunocc):
                    ecube2 = e_unocc[kcube, b2]
                    for b1 in range(nocc):
                        ecube1 = e_occ[kcube, b1]
                        dcube = ecube2 - ecube1
                        dmin, dmax = dcube.min(), dcube.max()
                        for ifr in range(nfrq):
                            f = frq1[ifr]
                            if f < dmin:
                                continue
                            if f > dmax:
                                wadd = wadd0
                            else:
                                wadd = np.zeros(8)
                                for it in range(6):
                                    dt = dcube[tetra[it]]
                                    if np.all(dt >= f):
                                        continue
                                    if np.all(dt < f):
                                        wadd[tetra[it]] += 1
                                    else:
                                        wadd[tetra[it]] += cuttetraN(dt, f)[0]
                            for c in range(8):
                                W[kcube[c], b1, b2, ifr] += wadd[c]
    rdum = 4.0 * 6 * nk
    out = np.empty_like(W)
    for ifr in range(nfrq - 1, 0, -1):
        out[..., ifr] = (W[..., ifr] - W[..., ifr - 1]) / (2 * (frq1[ifr] - frq[ifr]) * rdum)
    out[..., 0] = W[..., 0] / (2 * (frq1[0] - frq[0]) * rdum)
    if frq[0] == 0:
        out[..., 0] = 0
    return out.reshape(nk * nocc * nunocc, nfrq)
