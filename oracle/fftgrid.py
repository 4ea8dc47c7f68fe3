"""numpy restatement of the reference's FFT grid setup -- TEST INFRASTRUCTURE.

Only tests/ may import it.  Restates:

* determine_imat (fft.f:163-238): among the unimodular integer matrices
  with entries in [-2, 2], the basis change imat that minimises the sum of
  squared cosines between the reciprocal basis vectors rlat*imat, scanned in
  the reference's loop order with its 1e-12 improvement threshold; imati is
  its adjugate;
* gcount_sphere (fft.f:101-158): the G points of rlat*imat in the sphere of
  radius gcutf, found shell by shell (|ix|+|iy|+|iz| = n) until a shell is
  empty, the half-widths doubled plus one and rounded up to 2^a 3^b 5^c 7^d;
* the step function on the grid (wavefproducts.f:576-590): cstep at the
  folded grid frequencies g (g > N/2 -> g - N) mapped by imat and cut at
  |rlat*imat*g| <= gmax, then an unnormalised backward FFT.

Conventions as the C ABI: rlat[:, j] is reciprocal basis vector j; imat
and imati are ordinary (row, column) integer matrices.
"""
import functools
import itertools

import numpy as np


def nonorth(r):
    """sum over the three pairs of basis vectors (columns) of cos^2 of their angle."""
    l2 = (r * r).sum(axis=-2)
    d01 = (r[..., :, 0] * r[..., :, 1]).sum(-1)
    d12 = (r[..., :, 1] * r[..., :, 2]).sum(-1)
    d20 = (r[..., :, 2] * r[..., :, 0]).sum(-1)
    return (d01**2 / (l2[..., 0] * l2[..., 1]) + d12**2 / (l2[..., 1] * l2[..., 2])
            + d20**2 / (l2[..., 2] * l2[..., 0]))


@functools.lru_cache(maxsize=2)
def unimodular_candidates(b=2):
    """Every 3x3 integer matrix with entries in [-b, b] and determinant 1, in
    the reference's loop order (i11 outermost ... i33 innermost)."""
    vals = np.arange(-b, b + 1)
    c = np.array(list(itertools.product(vals, repeat=9)), np.int64).reshape(-1, 3, 3)
    det = np.rint(np.linalg.det(c.astype(float))).astype(np.int64)
    out = c[det == 1]
    out.flags.writeable = False
    return out


def determine_imat(rlat):
    """(imat, imati) of fft.f:163-238."""
    rlat = np.asarray(rlat, float)
    cand = unimodular_candidates()
    vals = nonorth(np.einsum("ik,nkj->nij", rlat, cand.astype(float)))
    best, r0 = np.eye(3, dtype=np.int64), float(nonorth(rlat))
    for k in np.flatnonzero(vals < r0 - 1e-12):  # candidates that beat the start
        if vals[k] < r0 - 1e-12:
            r0, best = float(vals[k]), cand[k]
    imati = np.rint(np.linalg.inv(best.astype(float))).astype(np.int64)
    assert (best @ imati == np.eye(3, dtype=np.int64)).all()
    return best, imati


def smooth_size(n):
    """Smallest 2^a 3^b 5^c 7^d >= n (a <= 16, b <= 11, c <= 7, d <= 6)."""
    best = None
    for a in range(17):
        for b in range(12):
            for c in range(8):
                for d in range(7):
                    v = 2**a * 3**b * 5**c * 7**d
                    if v >= n and (best is None or v < best):
                        best = v
    return best


def gcount_sphere(rlat, imat, gcutf):
    """(nfft, number of G points in the sphere) of fft.f:36-51, 101-158."""
    r1 = np.asarray(rlat, float) @ np.asarray(imat, float)
    mx = np.array([-1000, -1000, -1000])
    count, n = 0, 0
    while True:
        pts = [(ix, iy, s * (n - abs(ix) - abs(iy)))
               for ix in range(-n, n + 1) for iy in range(-(n - abs(ix)), n - abs(ix) + 1)
               for s in ((1, -1) if n - abs(ix) - abs(iy) > 0 else (1,))]
        p = np.array(pts, float)
        v = p @ r1.T
        inside = (v * v).sum(1) <= gcutf**2
        if not inside.any():
            break
        count += int(inside.sum())
        mx = np.maximum(mx, np.array(pts)[inside].max(0))
        n += 1
    return tuple(smooth_size(2 * int(m) + 1) for m in mx), count


def step_grid(nfft, imat, rlat, gmax, cstep, cext):
    """theta(r) on the grid (x fastest, complex) from cstep(-c1:c1,-c2:c2,-c3:c3)
    given as an array [z][y][x] (wavefproducts.f:576-590)."""
    N = tuple(int(v) for v in nfft)
    box = np.zeros(N[::-1], complex)  # [z][y][x]
    g = [np.where(np.arange(n) > n // 2, np.arange(n) - n, np.arange(n)) for n in N]
    gx, gy, gz = np.meshgrid(g[0], g[1], g[2], indexing="ij")
    G = np.stack([gx, gy, gz], -1).reshape(-1, 3)  # x slowest here; reordered below
    G1 = G @ np.asarray(imat).T
    v = G1 @ np.asarray(rlat, float).T
    keep = (v * v).sum(1) <= gmax**2
    ix = np.stack(np.meshgrid(np.arange(N[0]), np.arange(N[1]), np.arange(N[2]), indexing="ij"),
                  -1).reshape(-1, 3)
    c = np.asarray(cstep).reshape(2 * cext[2] + 1, 2 * cext[1] + 1, 2 * cext[0] + 1)
    for (i, j, k), (a, b, d) in zip(ix[keep], G1[keep]):
        box[k, j, i] = c[d + cext[2], b + cext[1], a + cext[0]]
    return (np.fft.ifftn(box) * np.prod(N)).reshape(-1)


def step_direct(nfft, imat, rlat, gmax, cstep, cext, points):
    """The same truncated Fourier series summed directly at grid points
    (i, j, k): sum over the kept grid frequencies g of cstep(imat g)
    e^{+2 pi i g.r/N} -- no FFT."""
    N = np.array(nfft)
    c = np.asarray(cstep).reshape(2 * cext[2] + 1, 2 * cext[1] + 1, 2 * cext[0] + 1)
    out = []
    for p in points:
        s = 0j
        for gx in range(-((N[0] - 1) // 2), N[0] // 2 + 1):
            for gy in range(-((N[1] - 1) // 2), N[1] // 2 + 1):
                for gz in range(-((N[2] - 1) // 2), N[2] // 2 + 1):
                    g1 = np.asarray(imat) @ np.array([gx, gy, gz])
                    v = np.asarray(rlat, float) @ g1
                    if v @ v > gmax**2:
                        continue
                    ph = 2 * np.pi * (gx * p[0] / N[0] + gy * p[1] / N[1] + gz * p[2] / N[2])
                    s += c[g1[2] + cext[2], g1[1] + cext[1], g1[0] + cext[0]] * np.exp(1j * ph)
        out.append(s)
    return np.array(out)


def stepfunction_coeffs(rlat, vol, radii, cent, cext):
    """Fourier coefficients of the interstitial step function for spheres of
    the given radii at the given centres (overlap.f:141-172):
    theta(0) = 1 - sum 4 pi R^3/(3 vol), theta(G) = -sum 4 pi (sin gR -
    gR cos gR)/g^3/vol e^{-2 pi i c.G}.  Returns [z][y][x]."""
    c1, c2, c3 = cext
    z, y, x = np.meshgrid(np.arange(-c3, c3 + 1), np.arange(-c2, c2 + 1), np.arange(-c1, c1 + 1),
                          indexing="ij")
    G = np.stack([x, y, z], -1).astype(float)
    gn = np.sqrt(((G @ np.asarray(rlat, float).T) ** 2).sum(-1))
    out = np.zeros(gn.shape, complex)
    nz = gn > 0
    g = gn[nz]
    for r, c in zip(radii, cent):
        rr = g * r
        ph = np.exp(-2j * np.pi * (G[nz] @ np.asarray(c, float)))
        out[nz] -= 4 * np.pi * (np.sin(rr) - rr * np.cos(rr)) / g**3 / vol * ph
    out[~nz] = 1 - sum(4 * np.pi * r**3 / 3 / vol for r in radii)
    return out
