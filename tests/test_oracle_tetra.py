"""Pinning of the tetrahedron-weight oracle (oracle/tetra.py): cuttetraN
against exact integration over the clipped tetrahedron (a convex polytope
triangulated by scipy), the volume rule, the decomposition choice, and for
the whole mesh: flat bands (all weight in one bin), the Brillouin-zone sum
rule, and an independent assembly from the clipped-polytope integrals."""
import itertools
import math
from pathlib import Path

import numpy as np
import pytest
from scipy.spatial import ConvexHull, Delaunay

from oracle import tetra

# unit tetrahedron: corner j at these Cartesian points; barycentric
# coordinates of a point x are lambda = B^-1 (x - v0) completed to sum 1
V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def _bary(x):
    lam = x - V[0]  # for this tetrahedron lambda_1..3 are the coordinates
    return np.column_stack([1 - lam.sum(1), lam])


def clipped_corner_weights(d, f):
    """Exact integral of each barycentric coordinate over {x in T: d(x) < f}
    for the linear d with corner values d, scaled so that the whole
    tetrahedron gives 1 per corner (4/vol * integral)."""
    d = np.asarray(d, float)
    pts = [V[j] for j in range(4) if d[j] < f]
    for a in range(4):
        for b in range(a + 1, 4):
            if (d[a] < f) != (d[b] < f):
                s = (f - d[a]) / (d[b] - d[a])
                pts.append(V[a] + s * (V[b] - V[a]))
    pts = np.array(pts)
    if len(pts) < 4:
        return np.zeros(4)
    try:
        if ConvexHull(pts).volume < 1e-15:
            return np.zeros(4)
    except Exception:
        return np.zeros(4)
    tri = Delaunay(pts)
    vol_t = 1.0 / 6.0
    w = np.zeros(4)
    for s in tri.simplices:
        p = pts[s]
        vol = abs(np.linalg.det(p[1:] - p[0])) / 6.0
        w += vol * _bary(p).mean(0)
    return 4.0 * w / vol_t


@pytest.mark.parametrize("seed", range(6))
def test_cuttetraN_is_the_exact_clipped_integral(seed):
    g = np.random.default_rng(seed)
    d = g.uniform(-1, 1, 4)
    ds = np.sort(d)
    # thresholds in each regime: below, (d1,d2), (d2,d3), (d3,d4), above
    fs = [ds[0] - 0.1, *(0.5 * (ds[:-1] + ds[1:])), 0.3 * ds[1] + 0.7 * ds[2], ds[3] + 0.1]
    for f in fs:
        w, v, nt = tetra.cuttetraN(d, f)
        want = clipped_corner_weights(d, f)
        np.testing.assert_allclose(w, want, rtol=1e-9, atol=1e-12)
        # volume rule: the pieces' volumes add up to the mean corner weight
        np.testing.assert_allclose(v[:nt].sum() if nt else 0.0, w.sum() / 4, rtol=1e-12, atol=1e-15)


def test_cuttetraN_order_of_equal_values_and_limits():
    w, _, nt = tetra.cuttetraN([0.5, 0.5, 0.5, 0.5], 0.5 + 1e-15)
    assert nt == 1 and (w == 1).all()  # within eta of the top: whole tetrahedron
    w, _, nt = tetra.cuttetraN([0.2, 0.4, 0.6, 0.8], 0.2)
    assert nt == 0 and (w == 0).all()  # at the bottom: nothing (strict, eta)


def test_tetra_choice_shortest_diagonal():
    cubic = np.eye(3)
    assert tetra.tetra_choice(cubic, (4, 4, 4)) == 0  # all equal: the first (3-6)
    # fcc reciprocal lattice (bcc vectors): a1 + a2 + a3 is the short diagonal 1-8? check by lengths
    rl = np.array([[-1.0, 1.0, 1.0], [1.0, -1.0, 1.0], [1.0, 1.0, -1.0]]).T
    r1, r2, r3 = rl.T
    lens = [np.dot(v, v) for v in (r1 + r3 - r2, r2 + r3 - r1, r1 + r2 + r3, r1 + r2 - r3)]
    assert tetra.tetra_choice(rl, (2, 2, 2)) == int(np.argmin(lens))
    t = tetra.get_tetra(rl, (2, 2, 2))
    assert t.shape == (6, 4) and set(np.unique(t)) == set(range(8))
    # every tetrahedron holds both ends of the chosen diagonal
    ends = [(2, 5), (1, 6), (0, 7), (3, 4)][tetra.tetra_choice(rl, (2, 2, 2))]
    assert all(ends[0] in row and ends[1] in row for row in t)


def _mesh_case(g, nkpt3, nocc, nunocc):
    nk = int(np.prod(nkpt3))
    eo = -g.uniform(0.2, 1.0, (nk, nocc))
    eu = g.uniform(0.1, 1.2, (nk, nunocc))
    return eo, eu


def test_flat_bands_put_all_weight_in_one_bin():
    nkpt3 = (3, 2, 2)
    nk = 12
    frq = 0.05 * (1.3 ** np.arange(12) - 1)  # exponential mesh starting at 0
    frq1, dw = tetra.hilbert_bins(frq)
    d0 = 0.5 * (frq1[5] + frq1[6])  # inside bin 6
    eo = np.full((nk, 1), -d0 / 2)
    eu = np.full((nk, 2), d0 / 2)
    w = tetra.tetrahedron5_weights(nkpt3, np.eye(3), eo, eu, 0.0, frq)
    want = np.zeros(len(frq))
    want[6] = 1.0 / (nk * 2 * dw[6])
    for t in range(w.shape[0]):
        np.testing.assert_allclose(w[t], want, rtol=1e-13, atol=1e-15)


def test_brillouin_zone_sum_rule():
    g = np.random.default_rng(3)
    nkpt3 = (2, 3, 2)
    eo, eu = _mesh_case(g, nkpt3, 2, 3)
    frq = 0.02 * (1.25 ** np.arange(24) - 1) + 0.01
    frq1, dw = tetra.hilbert_bins(frq)
    assert frq1[-1] > (eu.max() - eo.min())
    w = tetra.tetrahedron5_weights(nkpt3, np.eye(3), eo, eu, 0.0, frq)
    nk, nocc, nunocc = eo.shape[0], 2, 3
    tot = (w * (2 * dw)[None, :]).reshape(nk, nocc, nunocc, -1).sum(axis=(0, 3))
    np.testing.assert_allclose(tot, 1.0, rtol=1e-12)


def test_mesh_weights_from_clipped_integrals():
    """Independent assembly: per cube and tetrahedron the exact clipped
    integrals at every bin edge, scattered to the corner k points, then the
    bin differences -- against the oracle's cuttetraN path."""
    g = np.random.default_rng(5)
    nkpt3 = (2, 2, 2)
    rl = np.array([[1.0, 0.1, 0.0], [0.0, 1.2, 0.0], [0.2, 0.0, 0.9]]).T
    eo, eu = _mesh_case(g, nkpt3, 1, 2)
    frq = 0.1 * (1.4 ** np.arange(10) - 1) + 0.2
    frq1, dw = tetra.hilbert_bins(frq)
    got = tetra.tetrahedron5_weights(nkpt3, rl, eo, eu, 0.0, frq)
    tt = tetra.get_tetra(rl, nkpt3)
    pk = tetra.mesh_index(nkpt3)
    nk = 8
    cum = np.zeros((nk, 1, 2, len(frq)))
    for i1 in range(2):
        for i2 in range(2):
            for i3 in range(2):
                kc = [pk(i1 + a, i2 + b, i3 + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                for b2 in range(2):
                    dc = eu[kc, b2] - eo[kc, 0]
                    for j, f in enumerate(frq1):
                        for row in tt:
                            w = clipped_corner_weights(dc[row], f)
                            for c, wc in zip(row, w):
                                cum[kc[c], 0, b2, j] += wc
    want = np.empty_like(cum)
    want[..., 1:] = np.diff(cum, axis=-1) / (2 * dw[1:] * 24 * nk)
    want[..., 0] = cum[..., 0] / (2 * dw[0] * 24 * nk)
    np.testing.assert_allclose(got, want.reshape(got.shape), rtol=1e-8, atol=1e-10)


def test_metals_and_bad_meshes_are_rejected():
    nkpt3 = (2, 2, 2)
    eo = np.full((8, 1), -0.1)
    eu = np.full((8, 1), 0.2)
    eu[3, 0] = -0.05  # an unoccupied band below the Fermi energy at one k
    frq = 0.1 * (1.3 ** np.arange(6) - 1)
    with pytest.raises(ValueError):
        tetra.tetrahedron5_weights(nkpt3, np.eye(3), eo, eu, 0.0, frq)
    with pytest.raises(ValueError):
        tetra.hilbert_bins(np.array([0.0, 0.1, 0.15, 0.4]))  # not exponential


def clipped_weights_multi(halfspaces):
    """Exact 4/vol * integral of each barycentric coordinate over the part of
    the unit tetrahedron where every (values, c, sign) half-space holds:
    sign * (linear function with corner values `values`) < sign * c."""
    pts = [V[j] for j in range(4)]
    for vals, c, sgn in halfspaces:
        vals = np.asarray(vals, float)
        lam = _bary(np.array(pts))
        g = sgn * (lam @ vals - c)  # < 0 inside
        keep = [p for p, gi in zip(pts, g) if gi < 0]
        for a in range(len(pts)):
            for b in range(a + 1, len(pts)):
                if (g[a] < 0) != (g[b] < 0):
                    s = g[a] / (g[a] - g[b])
                    keep.append(pts[a] + s * (pts[b] - pts[a]))
        pts = keep
        if len(pts) < 4:
            return np.zeros(4)
    pts = np.unique(np.round(np.array(pts), 15), axis=0)
    if len(pts) < 4:
        return np.zeros(4)
    try:
        if ConvexHull(pts).volume < 1e-14:
            return np.zeros(4)
    except Exception:
        return np.zeros(4)
    tri = Delaunay(pts)
    w = np.zeros(4)
    for s in tri.simplices:
        p = pts[s]
        vol = abs(np.linalg.det(p[1:] - p[0])) / 6.0
        w += vol * _bary(p).mean(0)
    return 4.0 * w * 6.0


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("mode", ["occ", "unocc", "both"])
def test_metal_corner_weights_are_exact_clipped_integrals(seed, mode):
    """The (c1) corner weights (occupied/unoccupied part through cuttetraN's
    sub-tetrahedra, then the cut d < f) against the exact integral over
    {e_occ < ef, e_unocc > ef, d < f}."""
    g = np.random.default_rng(100 + seed)
    ef = 0.0
    e1 = g.uniform(-0.6, 0.3, 4) if mode != "unocc" else g.uniform(-0.6, -0.05, 4)
    e2 = g.uniform(-0.3, 0.8, 4) if mode != "occ" else g.uniform(0.05, 0.8, 4)
    cut1, cut2 = e1.max() > ef, e2.min() < ef
    part = tetra._restricted(e1, e2, ef, cut1, cut2)
    d = e2 - e1
    for f in np.linspace(d.min() - 0.05, d.max() + 0.05, 9):
        got = tetra.corner_weights_at(d, part, f)
        want = clipped_weights_multi([(e1, ef, 1), (e2, ef, -1), (d, f, 1)])
        np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-11)


def test_metal_path_without_cuts_is_the_insulator_path():
    g = np.random.default_rng(21)
    nkpt3 = (2, 3, 2)
    eo, eu = _mesh_case(g, nkpt3, 2, 3)
    frq = 0.02 * (1.25 ** np.arange(20) - 1) + 0.01
    a = tetra.tetrahedron5_weights(nkpt3, RL_T, eo, eu, 0.0, frq)
    b = tetra.tetrahedron5_weights_metal(nkpt3, RL_T, eo, eu, 0.0, frq)
    assert np.array_equal(a, b)


RL_T = np.array([[1.0, 0.1, 0.0], [0.0, 1.2, 0.0], [0.2, 0.0, 0.9]]).T


def test_metal_zone_sum_is_the_occupied_unoccupied_volume():
    """Summed over k and bins, each (b1, b2) carries the volume fraction of
    {e_occ(k) < ef < e_unocc(k+q)} under linear interpolation in every
    tetrahedron -- computed here by exact clipping, independently."""
    g = np.random.default_rng(22)
    nkpt3 = (2, 2, 3)
    nk = 12
    kx = np.arange(nk)
    band = 0.4 * np.cos(2 * np.pi * (kx % 3) / 3) + 0.2 * ((kx // 6) - 0.5)
    eo = np.stack([band - 0.05, band - 0.6], 1) + 0.02 * g.standard_normal((nk, 2))
    eu = np.stack([band + 0.1, band + 0.7], 1) + 0.02 * g.standard_normal((nk, 2))
    ef = 0.05
    frq = 0.02 * (1.3 ** np.arange(20) - 1) + 0.001
    frq1, dw = tetra.hilbert_bins(frq)
    assert frq1[-1] > (eu.max() - eo.min())
    w = tetra.tetrahedron5_weights_metal(nkpt3, RL_T, eo, eu, ef, frq)
    tot = (w * (2 * dw)[None, :]).reshape(nk, 2, 2, -1).sum(axis=(0, 3))
    tt = tetra.get_tetra(RL_T, nkpt3)
    pk = tetra.mesh_index(nkpt3)
    want = np.zeros((2, 2))
    for i1 in range(2):
        for i2 in range(2):
            for i3 in range(3):
                kc = [pk(i1 + a, i2 + b, i3 + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                for b1 in range(2):
                    for b2 in range(2):
                        for row in tt:
                            e1 = eo[np.array(kc)[row], b1]
                            e2 = eu[np.array(kc)[row], b2]
                            cw = clipped_weights_multi([(e1, ef, 1), (e2, ef, -1)])
                            want[b1, b2] += cw.sum() / 4 / 6 / nk
    assert ((want > 0.01) & (want < 0.99)).any()  # some pairs are cut by the Fermi energy
    np.testing.assert_allclose(tot, want, rtol=1e-8, atol=1e-12)


def test_degenerate_averaging_rules():
    """Averaging over degenerate groups: members end up equal, group sums
    are kept, singleton groups change nothing, and on exactly degenerate
    bands (two copies of one band) the raw weights already agree."""
    g = np.random.default_rng(31)
    nkpt3 = (2, 2, 2)
    nk = 8
    eo, eu = _mesh_case(g, nkpt3, 3, 4)
    eo.sort(axis=1)
    eu.sort(axis=1)
    eo[:, 1] = eo[:, 2]  # occupied bands 1, 2 degenerate everywhere (still sorted)
    eu[:, 0] = eu[:, 1]  # unoccupied 0, 1 degenerate everywhere
    frq = 0.02 * (1.25 ** np.arange(20) - 1) + 0.01
    w = tetra.tetrahedron5_weights(nkpt3, np.eye(3), eo, eu, 0.0, frq)
    do, du = tetra.deg_from_energies(eo, 1e-12), tetra.deg_from_energies(eu, 1e-12)
    assert (do[:, 1] == 2).all() and (do[:, 2] == 1).all() and (du[:, 0] == 1).all()
    a = tetra.average_degenerate(w, nk, 3, 4, do, du)
    W, A = w.reshape(nk, 3, 4, -1), a.reshape(nk, 3, 4, -1)
    np.testing.assert_allclose(A, W, rtol=1e-13, atol=1e-15)  # identical bands: no change
    # a perturbation inside the groups is averaged out, the group sums kept
    P = W + g.uniform(0, 1e-3, W.shape)
    B = tetra.average_degenerate(P.reshape(w.shape), nk, 3, 4, do, du).reshape(P.shape)
    np.testing.assert_allclose(B[:, 1], B[:, 2], rtol=1e-14)
    np.testing.assert_allclose(B[:, :, 0], B[:, :, 1], rtol=1e-14)
    np.testing.assert_allclose(B.sum(axis=(1, 2)), P.sum(axis=(1, 2)), rtol=1e-13)
    np.testing.assert_array_equal(B[:, 0, 2:], P[:, 0, 2:])  # singleton groups untouched


def _literal_scatter(nkpt3, eo, eu, frq, kpt1p, nkpt1, kpt1=None):
    """tetrahedron5_init's cube loop for one occupied band, written out: the
    corner weights of each cube go to kcubep = kpt1p(kcube), corners without
    a parent skipped and, under TETRAF (kpt1 given), corners not listed in
    kpt1 skipped too and each parent's sum scaled by count(kpt1p == p);
    then the bin differences."""
    nk = int(np.prod(nkpt3))
    nunocc = eu.shape[1]
    tt = tetra.get_tetra(np.eye(3), nkpt3)
    pk = tetra.mesh_index(nkpt3)
    frq1, dw = tetra.hilbert_bins(frq)
    cum = np.zeros((nkpt1, 1, nunocc, len(frq)))
    listed = set(range(nk)) if kpt1 is None else set(kpt1)
    for i1 in range(nkpt3[0]):
        for i2 in range(nkpt3[1]):
            for i3 in range(nkpt3[2]):
                kc = [pk(i1 + a, i2 + b, i3 + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                for b2 in range(nunocc):
                    dc = eu[kc, b2] - eo[kc, 0]
                    for j, f in enumerate(frq1):
                        wadd = np.zeros(8)
                        for row in tt:
                            wadd[row] += tetra.corner_weights_at(dc[row], None, f)
                        for c in range(8):
                            if kpt1p[kc[c]] >= 0 and kc[c] in listed:
                                cum[kpt1p[kc[c]], 0, b2, j] += wadd[c]
    if kpt1 is not None:
        for p in range(nkpt1):
            cum[p] *= np.count_nonzero(np.asarray(kpt1p) == p)
    want = np.empty_like(cum)
    want[..., 1:] = np.diff(cum, axis=-1) / (2 * dw[1:] * 24 * nk)
    want[..., 0] = cum[..., 0] / (2 * dw[0] * 24 * nk)
    return want.reshape(nkpt1 * nunocc, len(frq))


def test_parents_reduction_is_the_reference_scatter():
    """A literal scatter with kcubep = kpt1p(kcube) (corners without a parent
    skipped) equals the star sums of the full-mesh weights."""
    g = np.random.default_rng(41)
    nkpt3 = (2, 2, 2)
    nk = 8
    eo, eu = _mesh_case(g, nkpt3, 1, 2)
    frq = 0.05 * (1.3 ** np.arange(12) - 1) + 0.02
    kpt1p = np.array([0, 1, 1, 2, 0, -1, 2, 3])  # some stars, one point without a parent
    nkpt1 = 4
    full = tetra.tetrahedron5_weights(nkpt3, np.eye(3), eo, eu, 0.0, frq)
    got = tetra.reduce_to_parents(full, nk, 1, 2, kpt1p, nkpt1)
    want = _literal_scatter(nkpt3, eo, eu, frq, kpt1p, nkpt1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_tetraf_reduction_is_the_reference_scatter():
    """The same literal scatter under TETRAF (only corners listed in kpt1,
    rows scaled by the star sizes) equals reduce_tetraf, including a parent
    whose listed point is not the first of its star and a listed point that
    has no parent."""
    g = np.random.default_rng(43)
    nkpt3 = (2, 2, 2)
    nk = 8
    eo, eu = _mesh_case(g, nkpt3, 1, 2)
    frq = 0.05 * (1.3 ** np.arange(12) - 1) + 0.02
    kpt1p = np.array([0, 1, 1, 2, 0, -1, 2, 3])
    kpt1 = [0, 2, 6, 7]  # parent 1 listed by its second member, parent 2 by 6
    full = tetra.tetrahedron5_weights(nkpt3, np.eye(3), eo, eu, 0.0, frq)
    got = tetra.reduce_tetraf(full, nk, 1, 2, kpt1, kpt1p, 4)
    want = _literal_scatter(nkpt3, eo, eu, frq, kpt1p, 4, kpt1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    # the rows are count * the one listed row, not the star sums
    F = full.reshape(nk, 2, -1)
    np.testing.assert_allclose(got.reshape(4, 2, -1)[1], 2 * F[2], rtol=1e-14)
    assert not np.allclose(got, tetra.reduce_to_parents(full, nk, 1, 2, kpt1p, 4))


def test_tetraf_equals_star_sums_under_exact_symmetry():
    """Bands invariant under the 12 operations (axis permutations, with and
    without inversion) that also keep the Kuhn decomposition along the
    (1,1,1) diagonal: equivalent k points then carry equal full-mesh weights,
    so count(star) * w(representative) is the star sum."""
    n = 4
    rl = np.array([[1.0, -0.4, -0.4], [-0.4, 1.0, -0.4], [-0.4, -0.4, 1.0]])
    assert tetra.tetra_choice(rl, (n, n, n)) == 2  # the diagonal 1-8
    pk = tetra.mesh_index((n, n, n))
    pts = np.array([(i1, i2, i3) for i1 in range(n) for i2 in range(n) for i3 in range(n)])
    ops = []
    for perm in itertools.permutations(range(3)):
        P = np.eye(3, dtype=int)[list(perm)]
        ops += [P, -P]
    x = 2 * np.pi * pts / n
    sym = np.cos(x).sum(1) + 0.7 * np.cos(x.sum(1)) + 0.3 * np.cos(2 * x).sum(1) * np.cos(x).sum(1)
    eo = np.stack([-2.0 + 0.2 * sym], 1)
    eu = np.stack([1.7 - 0.2 * sym, 1.8 + 0.05 * sym ** 2], 1)
    frq = 0.05 * (1.25 ** np.arange(30) - 1) + 0.02
    full = tetra.tetrahedron5_weights((n, n, n), rl, eo, eu, 0.0, frq)
    F = full.reshape(n ** 3, 2, -1)
    for R in ops:
        img = [pk(*(R @ p)) for p in pts]
        np.testing.assert_allclose(F[img], F, rtol=1e-11, atol=1e-14 * np.abs(F).max())
    kpt1, kpt1p, rep = [], np.full(n ** 3, -1), {}
    for k, p in enumerate(pts):
        key = min(pk(*(R @ p)) for R in ops)
        if key not in rep:
            rep[key] = len(kpt1)
            kpt1.append(k)
        kpt1p[k] = rep[key]
    sizes = np.bincount(kpt1p)
    assert sizes.max() > 1 and len(kpt1) < n ** 3
    got = tetra.reduce_tetraf(full, n ** 3, 1, 2, kpt1, kpt1p, len(kpt1))
    want = tetra.reduce_to_parents(full, n ** 3, 1, 2, kpt1p, len(kpt1))
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-14 * np.abs(want).max())


def _cubic_ops():
    """The 48 signed permutation matrices (the cubic point group)."""
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3), int)
            for i in range(3):
                R[i, perm[i]] = signs[i]
            ops.append(R)
    return ops


def _symmetric_mesh(n, q):
    """A cubic n^3 mesh with bands invariant under the cubic group, the
    image table kptsym(k, isym) and the index of q."""
    pk = tetra.mesh_index((n, n, n))
    pts = np.array([(i1, i2, i3) for i1 in range(n) for i2 in range(n) for i3 in range(n)])
    ops = _cubic_ops()
    kptsym = np.array([[pk(*(R @ p)) for R in ops] for p in pts])

    def band(p, shift):
        x = 2 * np.pi * (p + shift) / n
        return np.cos(x).sum(-1)
    eo = np.stack([-3.2 + 0.3 * band(pts, 0)], 1)
    eu = np.stack([1.5 - 0.4 * band(pts, np.array(q)), 2.0 - 0.2 * band(pts, np.array(q))], 1)
    return pts, ops, kptsym, pk(*q), eo, eu


def test_equivalent_average_is_the_orbit_mean_under_the_little_group():
    """All k of a 4^3 mesh in kpt1 (no reduction): each k's weights become
    the mean over its orbit under the symmetries that keep q = (1,0,0)/4
    (an orbit built as a set here), equal on every member, and the
    decomposition's symmetry breaking is really there to remove."""
    n, q = 4, (1, 0, 0)
    pts, ops, kptsym, iq, eo, eu = _symmetric_mesh(n, q)
    nk = n ** 3
    frq = 0.05 * (1.25 ** np.arange(40) - 1) + 0.01
    w = tetra.tetrahedron5_weights((n, n, n), np.eye(3), eo, eu, 0.0, frq)
    kpt1, kpt1p = np.arange(nk), np.arange(nk)
    got = tetra.average_equivalent(w, 1, 2, kpt1, kpt1p, kptsym, iq).reshape(nk, 2, -1)
    W = w.reshape(nk, 2, -1)
    little = [R for R in ops if tuple(R @ np.array(q) % n) == q]
    assert len(little) == 8
    pk = tetra.mesh_index((n, n, n))
    broken = 0.0
    for k in range(nk):
        orbit = sorted({pk(*(R @ pts[k])) for R in little})
        np.testing.assert_allclose(got[k], W[orbit].mean(0), rtol=1e-13, atol=1e-15)
        for m in orbit:
            np.testing.assert_array_equal(got[m], got[k])
        broken = max(broken, np.abs(W[orbit] - W[orbit].mean(0)).max())
    assert broken > 1e-6 * np.abs(W).max()
    # the whole group (not the little group) would mix other orbits in
    full = [sorted({pk(*(R @ p)) for R in ops}) for p in pts]
    assert any(len(o) > len({pk(*(R @ pts[k])) for R in little}) for k, o in enumerate(full))


def test_equivalent_average_weights_by_counts_and_skips_reduced_sets():
    """Hand-worked cases: two symmetries (identity and one that sends point
    0 to point 2, whose parent is 1), kpt1 = (0, 1): point 0's images have
    parents 0 and 1, each counted once -> the plain mean; three symmetries
    whose images of point 0 have parents (0, 0, 2) -> counts (2, 1) and the
    mean (2 w0 + w2)/3, point 1 untouched; a symmetry that moves q is not
    used; a k set that is irreducible under the little group is left as it
    is."""
    w = np.array([[1.0, 10.0], [4.0, 40.0]])  # (nkpt1 * 1 * 1, nfrqs)
    kptsym = np.array([[0, 2], [1, 1], [2, 0]])  # (nk, nsym): symmetry 1 swaps 0 and 2
    got = tetra.average_equivalent(w, 1, 1, [0, 1], [0, 1, 1], kptsym, 1)
    np.testing.assert_array_equal(got, [[2.5, 25.0], [2.5, 25.0]])
    # three points, counts (2, 1): identity twice on 0's image table
    kptsym3 = np.array([[0, 0, 3], [1, 1, 1], [2, 2, 2], [3, 3, 0]])
    w3 = np.array([[3.0], [5.0], [7.0]])
    got3 = tetra.average_equivalent(w3, 1, 1, [0, 1, 2], [0, 1, 2, 2], kptsym3, 1)
    np.testing.assert_array_equal(got3[[0, 2]], [[(2 * 3.0 + 7.0) / 3]] * 2)
    assert got3[1, 0] == 5.0
    # q outside the symmetry's little group: symmetry 1 does not keep q = 0
    kq = np.array([[0, 2], [1, 0], [2, 1]])
    np.testing.assert_array_equal(tetra.average_equivalent(w, 1, 1, [0, 1], [0, 1, 1], kq, 0), w)
    # irreducible set: every image has the point's own parent
    n, q = 4, (1, 0, 0)
    pts, ops, kptsym, iq, _, _ = _symmetric_mesh(n, q)
    pk = tetra.mesh_index((n, n, n))
    little = [s for s, R in enumerate(ops) if kptsym[iq, s] == iq]
    rep = {}
    kpt1, kpt1p = [], np.full(n ** 3, -1)
    for k in range(n ** 3):
        key = min(kptsym[k, s] for s in little)
        if key not in rep:
            rep[key] = len(kpt1)
            kpt1.append(k)
        kpt1p[k] = rep[key]
    wr = np.random.default_rng(5).uniform(0, 1, (len(kpt1) * 2, 3))
    assert tetra.equivalent_groups(kpt1, kpt1p, kptsym, iq) == []
    np.testing.assert_array_equal(tetra.average_equivalent(wr, 1, 2, kpt1, kpt1p, kptsym, iq), wr)


def _degenerate_bands(g, nk, nb, sign, copies, local=()):
    """Sorted bands (nk, nb) of one sign, bands copies[i] exactly equal to
    the band below them on every k, and at the k points in `local` one more
    degenerate pair (bands 3 and 4)."""
    e = np.sort(sign * g.uniform(0.1, 1.2, (nk, nb)), axis=1)
    for b in copies:
        e[:, b] = e[:, b - 1]
    for k in local:
        e[k, 4] = e[k, 3]
    return e


@pytest.mark.parametrize("w1,w2", [((2, 3), (1, 3)), ((1, 4), (2, 2)), ((3, 3), (0, 2))])
def test_cut_windows_give_the_rows_of_whole_groups(w1, w2):
    """Windows that cut degenerate groups give, in every window row, what a
    window holding the whole groups gives after the ordinary averaging: the
    purpose of the reference's widening.  Full mesh, degenerate pairs on
    every k (bands 1-2 occupied, 2-3 unoccupied) and on a few k only."""
    g = np.random.default_rng(61)
    nkpt3 = (2, 3, 2)
    nk = 12
    e1 = _degenerate_bands(g, nk, 6, -1.0, copies=(2,), local=(0, 5, 7))
    e2 = _degenerate_bands(g, nk, 6, 1.0, copies=(3,), local=(1, 5))
    d1, d2 = tetra.deg_from_energies(e1, 1e-12), tetra.deg_from_energies(e2, 1e-12)
    frq = 0.04 * (1.25 ** np.arange(20) - 1)
    got = tetra.tetrahedron5_weights_bands(nkpt3, RL_T, e1, e2, d1, d2, *w1, *w2, 0.0, frq)
    full = tetra.tetrahedron5_weights_metal(nkpt3, RL_T, e1, e2, 0.0, frq)
    wide = tetra.average_degenerate(full, nk, 6, 6, d1, d2).reshape(nk, 6, 6, -1)
    want = wide[:, w1[0]:w1[1] + 1, w2[0]:w2[1] + 1].reshape(got.shape)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15 * np.abs(want).max())
    # the windows really cut groups (at some k) for these cases
    cut = any(tetra._group(d1[k], w1[0])[0] < w1[0] or tetra._group(d1[k], w1[1])[1] > w1[1]
              or tetra._group(d2[k], w2[0])[0] < w2[0] or tetra._group(d2[k], w2[1])[1] > w2[1]
              for k in range(nk))
    assert cut


def test_cut_windows_on_reduced_sets():
    """The same on the rows of a reduced k set, star sums and TETRAF, with
    the degeneracies shared by every member of a star."""
    g = np.random.default_rng(62)
    nkpt3 = (2, 2, 2)
    nk = 8
    e1 = _degenerate_bands(g, nk, 5, -1.0, copies=(2,))
    e2 = _degenerate_bands(g, nk, 5, 1.0, copies=(1, 4))
    d1, d2 = tetra.deg_from_energies(e1, 1e-12), tetra.deg_from_energies(e2, 1e-12)
    frq = 0.04 * (1.25 ** np.arange(16) - 1)
    kpt1p = np.array([0, 1, 1, 2, 0, -1, 2, 3])
    kpt1 = np.array([0, 1, 3, 7])
    full = tetra.tetrahedron5_weights_metal(nkpt3, np.eye(3), e1, e2, 0.0, frq)
    for tf in (False, True):
        got = tetra.tetrahedron5_weights_bands(nkpt3, np.eye(3), e1, e2, d1, d2, 2, 3, 1, 3, 0.0,
                                               frq, kpt1, kpt1p, 4, tetraf=tf)
        red = (tetra.reduce_tetraf(full, nk, 5, 5, kpt1, kpt1p, 4) if tf
               else tetra.reduce_to_parents(full, nk, 5, 5, kpt1p, 4))
        wide = tetra.average_degenerate(red, 4, 5, 5, d1[kpt1], d2[kpt1]).reshape(4, 5, 5, -1)
        want = wide[:, 2:4, 1:4].reshape(got.shape)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15 * np.abs(want).max())


def test_uncut_windows_equal_the_plain_chain():
    g = np.random.default_rng(63)
    nkpt3 = (2, 2, 3)
    nk = 12
    e1 = _degenerate_bands(g, nk, 5, -1.0, copies=(2,))
    e2 = _degenerate_bands(g, nk, 5, 1.0, copies=(2,))
    d1, d2 = tetra.deg_from_energies(e1, 1e-12), tetra.deg_from_energies(e2, 1e-12)
    frq = 0.04 * (1.25 ** np.arange(16) - 1)
    got = tetra.tetrahedron5_weights_bands(nkpt3, np.eye(3), e1, e2, d1, d2, 1, 2, 1, 3, 0.0, frq)
    w = tetra.tetrahedron5_weights_metal(nkpt3, np.eye(3), e1[:, 1:3], e2[:, 1:4], 0.0, frq)
    want = tetra.average_degenerate(w, nk, 2, 3, d1[:, 1:3] - 1, d2[:, 1:4] - 1)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("w1,w2", [((2, 3), (1, 3)), ((3, 3), (0, 2))])
def test_alignbd_rows_are_clamped_sums_that_the_averaging_divides(w1, w2):
    """ALIGNBD stops before the averaging: every window row holds the
    all-band weights of the bands clamped into it at its k (its own band,
    and at a window edge the members of the edge group outside the window),
    and averaging those rows over the window part of each group and dividing
    by the whole group's size gives the rows without ALIGNBD."""
    g = np.random.default_rng(64)
    nkpt3 = (2, 3, 2)
    nk = 12
    e1 = _degenerate_bands(g, nk, 6, -1.0, copies=(2,), local=(0, 5, 7))
    e2 = _degenerate_bands(g, nk, 6, 1.0, copies=(3,), local=(1, 5))
    d1, d2 = tetra.deg_from_energies(e1, 1e-12), tetra.deg_from_energies(e2, 1e-12)
    frq = 0.04 * (1.25 ** np.arange(20) - 1)
    raw = tetra.tetrahedron5_weights_bands(nkpt3, RL_T, e1, e2, d1, d2, *w1, *w2, 0.0, frq,
                                           alignbd=True)
    avg = tetra.tetrahedron5_weights_bands(nkpt3, RL_T, e1, e2, d1, d2, *w1, *w2, 0.0, frq)
    full = tetra.tetrahedron5_weights_metal(nkpt3, RL_T, e1, e2, 0.0, frq).reshape(nk, 6, 6, -1)
    no, nu = w1[1] - w1[0] + 1, w2[1] - w2[0] + 1
    R = raw.reshape(nk, no, nu, -1)
    A = avg.reshape(nk, no, nu, -1)

    def members(deg, k, lo, hi, b):
        a = tetra._group(deg[k], lo)[0] if b == lo else b
        z = tetra._group(deg[k], hi)[1] if b == hi else b
        return range(a, z + 1)

    for k in range(nk):
        for r1 in range(no):
            for r2 in range(nu):
                m1 = members(d1, k, *w1, w1[0] + r1)
                m2 = members(d2, k, *w2, w2[0] + r2)
                want = full[k][np.ix_(list(m1), list(m2))].sum(axis=(0, 1))
                np.testing.assert_allclose(R[k, r1, r2], want, rtol=1e-13,
                                           atol=1e-15 * np.abs(full).max())
                f1, l1 = tetra._group(d1[k], w1[0] + r1)
                f2, l2 = tetra._group(d2[k], w2[0] + r2)
                p1 = slice(max(f1, w1[0]) - w1[0], min(l1, w1[1]) - w1[0] + 1)
                p2 = slice(max(f2, w2[0]) - w2[0], min(l2, w2[1]) - w2[0] + 1)
                mean = R[k, p1, p2].sum(axis=(0, 1)) / ((l1 - f1 + 1) * (l2 - f2 + 1))
                np.testing.assert_allclose(A[k, r1, r2], mean, rtol=1e-13,
                                           atol=1e-15 * np.abs(full).max())
    assert np.abs(raw - avg).max() > 1e-6 * np.abs(avg).max()


def _exp_mesh(n, w0, r, zero=False):
    """An exponential Hilbert mesh: steps growing by r, optionally from 0."""
    steps = w0 * r ** np.arange(n - 1)
    start = 0.0 if zero else w0
    return start + np.concatenate([[0.0], np.cumsum(steps)])


def test_gauss_weights_are_bin_integrals_of_the_gaussian():
    """Each weight times the bin width is the quadrature of the occupation
    factors times the normal density of width sigma2 around d over the bin
    (the bin edges from the mesh factor a = 1/(1+r) of an exponential mesh),
    and bins at w = 0 are empty."""
    from scipy import integrate
    frq = _exp_mesh(12, 0.05, 1.3, zero=True)
    eo = np.array([[-0.4, -0.05]])
    eu = np.array([[0.02, 0.6, 1.1]])
    ef, s1, s2 = 0.0, 0.03, 0.08
    w = tetra.gauss4_weights(eo, eu, ef, (s1, s2), frq).reshape(2, 3, -1)
    assert (w[..., 0] == 0).all()
    a = 1 / (1 + 1.3)
    for b1 in range(2):
        for b2 in range(3):
            d = eu[0, b2] - eo[0, b1]
            occ = (1 + math.erf((ef - eo[0, b1]) / (math.sqrt(2) * s1))) / 2
            unocc = (1 + math.erf((eu[0, b2] - ef) / (math.sqrt(2) * s1))) / 2
            for i in range(1, len(frq)):
                df = a * (frq[i + 1] - frq[i]) if i < len(frq) - 1 else (1 - a) * (frq[i] - frq[i - 1])
                mass, _ = integrate.quad(
                    lambda x: math.exp(-(x - d) ** 2 / (2 * s2 ** 2)) / (math.sqrt(2 * math.pi) * s2),
                    frq[i] - df, frq[i] + df, epsabs=1e-15, epsrel=1e-13)
                np.testing.assert_allclose(w[b1, b2, i] * 2 * df, occ * unocc * mass, rtol=1e-10,
                                           atol=1e-16)


def test_gauss_weights_tile_the_mesh_and_sum_over_stars():
    """The bins of an exponential mesh tile [w1 - df1, w_n + df_n], so
    sum_i w(i) 2 df(i) is the Gaussian mass there times the occupations
    (divided by nk); the rows of a reduced set are the sums over their
    members, and mesh points without a parent drop out."""
    g = np.random.default_rng(81)
    nk, nocc, nunocc = 6, 2, 3
    eo = g.uniform(-1.0, -0.1, (nk, nocc))
    eu = g.uniform(0.1, 1.0, (nk, nunocc))
    frq = _exp_mesh(30, 0.01, 1.15)
    s = (0.02, 0.05)
    w = tetra.gauss4_weights(eo, eu, 0.0, s, frq).reshape(nk, nocc, nunocc, -1)
    a = 1 / (1 + 1.15)  # steps grow by r = 1.15
    df = np.append(a * np.diff(frq), (1 - a) * (frq[-1] - frq[-2]))
    lo, hi = frq[0] - df[0], frq[-1] + df[-1]
    np.testing.assert_allclose(frq[:-1] + df[:-1], frq[1:] - df[1:], rtol=1e-13)
    for k in range(nk):
        for b1 in range(nocc):
            for b2 in range(nunocc):
                d = eu[k, b2] - eo[k, b1]
                occ = (1 + math.erf(-eo[k, b1] / (math.sqrt(2) * s[0]))) / 2
                unocc = (1 + math.erf(eu[k, b2] / (math.sqrt(2) * s[0]))) / 2
                mass = (math.erf((hi - d) / (math.sqrt(2) * s[1]))
                        - math.erf((lo - d) / (math.sqrt(2) * s[1]))) / 2
                np.testing.assert_allclose((w[k, b1, b2] * 2 * df).sum(), occ * unocc * mass / nk,
                                           rtol=1e-12)
    kpt1p = np.array([0, 1, 1, -1, 0, 2])
    red = tetra.gauss4_weights(eo, eu, 0.0, s, frq, kpt1p, 3).reshape(3, nocc, nunocc, -1)
    for p in range(3):
        np.testing.assert_allclose(red[p], w[kpt1p == p].sum(0), rtol=1e-13, atol=1e-300)


_REF_CUT = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libcuttetra.so"


@pytest.mark.skipif(not _REF_CUT.exists(), reason="oracle/_ref not built (no reference sources)")
def test_cuttetraN_matches_the_reference_build_bit_for_bit():
    """The reference's own cuttetraN (tetrahedron.f:347-412, compiled by
    oracle/build_ref.sh) against the restatement: random corners in every
    regime, tied corner values (minloc order), and thresholds at the eta
    edges and exactly at corner values."""
    import ctypes as C
    lib = C.CDLL(str(_REF_CUT))
    f = lib.cuttetran_
    f.restype = None
    dp = C.POINTER(C.c_double)

    def ref(e, ef):
        w = (C.c_double * 4)()
        v = (C.c_double * 3)(-7.0, -7.0, -7.0)
        n = C.c_int(-1)
        ea = (C.c_double * 4)(*e)
        efa = C.c_double(ef)
        f(w, v, C.byref(n), ea, C.byref(efa), None)
        return np.array(w[:]), np.array(v[:]), n.value

    g = np.random.default_rng(91)
    cases = []
    for _ in range(3000):
        e = g.uniform(-1, 1, 4)
        cases.append((e, g.uniform(-1.2, 1.2)))
    for _ in range(500):
        e = np.round(g.uniform(-1, 1, 4) * 4) / 4  # ties
        for ef in (e.min(), e.max(), np.median(e), e.min() + 1e-14, e.max() - 1e-14,
                   e.min() + 5e-15, e.max() - 5e-15, g.uniform(-1, 1)):
            cases.append((e, float(ef)))
    nreg = set()
    for e, ef in cases:
        wr, vr, nr = ref(e, ef)
        wo, vo, no = tetra.cuttetraN(e, ef)
        assert nr == no, (e, ef)
        assert np.array_equal(wr, wo), (e, ef, wr, wo)
        assert np.array_equal(vr[:nr], np.asarray(vo)[:nr]), (e, ef)
        nreg.add((nr, int(np.sum(e < ef))))
    assert len(nreg) >= 5  # every regime was visited
