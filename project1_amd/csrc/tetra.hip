// Tetrahedron weights of the Hilbert mesh (the wtetra of the HILBERT pair
// sum) for a full cubic k mesh, insulators and metals.
//
// Reference: tetrahedron5_init (tetrahedron.f:903-1332).  Every mesh cube
// is split into six tetrahedra around its shortest diagonal (get_tetra,
// :219-271); the band-energy difference d = e_unocc(k+q) - e_occ(k) is
// interpolated linearly in each tetrahedron, and for every upper bin edge
// frq1(i) of the Hilbert mesh the cube adds to each of its 8 corner k points
// the integral of that corner's barycentric coordinate over {d < frq1(i)}
// (cuttetraN, :347-412; 1 per tetrahedron when the whole tetrahedron lies
// below).  Differences of consecutive edges over 2 dw(i) * 24 nkpt give the
// weight of bin i (:1224-1231).  The reference loops over cubes and
// scatters to corners; here one thread owns one transition (k, b1, b2) and
// gathers from the 8 cubes that have k as a corner.  Their corners are the
// 27 mesh points k + {-1,0,1}^3, whose d values the thread loads once; only
// edges between the smallest and largest of them need the tetrahedron
// cuts (below: nothing, above: the full 24).  In cubes whose bands the Fermi
// energy cuts (metals, case (c1)) each tetrahedron is first restricted to
// its occupied-unoccupied part, as cuttetraN sub-tetrahedra (nested when
// both bands are cut), and each of those is cut by d < frq1(i).  The output
// is written once, transition-fastest (coalesced): n x nfrqs doubles.
// The averaging over degenerate states is a separate pass (two
// out-of-place kernels: occupied groups, then unoccupied groups).
#include "common.hpp"
#include "tetra_mesh.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>

namespace {

constexpr int TW_NT = 64;  // one wave: its 27 x 64 LDS values (13.5 KB) leave room for 11 waves per CU
constexpr double TW_ETA = 1e-14;  // cuttetraN's tolerance


// Weight of corner position `q` (0..3) of one tetrahedron with corner values
// e_in and threshold ef: cuttetraN's wtetra(q) (tetrahedron.f:347-412).
__device__ double cut_corner(const double e_in[4], double ef, int q) {
  // rank of each corner, ties to the lower index (minloc picks the first)
  int r[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    int k = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) k += (e_in[b] < e_in[a]) || (e_in[b] == e_in[a] && b < a);
    r[a] = k;
  }
  double e[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
#pragma unroll
    for (int b = 0; b < 4; b++)
      if (r[b] == a) e[a] = e_in[b];
  }
  const int rq = r[q];  // the sorted position of corner q
  if (ef >= e[3] - TW_ETA) return 1.0;
  if (ef >= e[2]) {
    const double f0 = (ef - e[0]) / (e[3] - e[0]), f1 = (ef - e[1]) / (e[3] - e[1]),
                 f2 = (ef - e[2]) / (e[3] - e[2]);
    const double v0 = f1, v1 = f0 * (1 - f1) * (1 - f2), v2 = f2 * (1 - f1);
    switch (rq) {
      case 0: return v0 + v1 * (2 - f0) + v2;
      case 1: return v0 + (v0 + v1 + v2) * (1 - f1);
      case 2: return v0 + v1 * (1 - f2) + v2 * (2 - f2);
      default: return v0 * f1 + v1 * f0 + (v1 + v2) * (f1 + f2);
    }
  }
  if (ef > e[1]) {
    const double f0 = (ef - e[0]) / (e[2] - e[0]), f1 = (ef - e[0]) / (e[3] - e[0]),
                 f2 = (ef - e[1]) / (e[2] - e[1]), f3 = (ef - e[1]) / (e[3] - e[1]);
    const double v0 = f0 * f1, v1 = f0 * (1 - f1) * f3, v2 = f2 * (1 - f0) * f3;
    switch (rq) {
      case 0: return v0 * (3 - f0 - f1) + v1 * (2 - f0 - f1) + v2 * (1 - f0);
      case 1: return v0 + v1 * (2 - f3) + v2 * (3 - f2 - f3);
      case 2: return (v0 + v1 + v2) * f0 + v2 * f2;
      default: return (v0 + v1) * f1 + (v1 + v2) * f3;
    }
  }
  if (ef >= e[0] + TW_ETA) {
    const double f0 = (ef - e[0]) / (e[1] - e[0]), f1 = (ef - e[0]) / (e[2] - e[0]),
                 f2 = (ef - e[0]) / (e[3] - e[0]);
    const double v0 = f0 * f1 * f2;
    switch (rq) {
      case 0: return v0 * (4 - f0 - f1 - f2);
      case 1: return v0 * f0;
      case 2: return v0 * f1;
      default: return v0 * f2;
    }
  }
  return 0.0;
}

// Linear index of the mesh point k + (a%3-1, (a/3)%3-1, a/9-1), periodic.
__device__ inline long nbr(int i1, int i2, int i3, int a, int n1, int n2, int n3) {
  const int j1 = (i1 + a % 3 - 1 + n1) % n1, j2 = (i2 + (a / 3) % 3 - 1 + n2) % n2,
            j3 = (i3 + a / 9 - 1 + n3) % n3;
  return ((long)j1 * n2 + j2) * n3 + j3;
}

// The 27 d values of a thread live in LDS (d[a][thread]), so the loops over
// cubes and tetrahedra index them at run time without scratch memory.
// Sorted ranks of four values (ties to the lower index, as the reference's
// minloc ordering) and the sorted values.
__device__ inline void rank4(const double e_in[4], int r[4], double e[4]) {
#pragma unroll
  for (int a = 0; a < 4; a++) {
    int k = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) k += (e_in[b] < e_in[a]) || (e_in[b] == e_in[a] && b < a);
    r[a] = k;
  }
#pragma unroll
  for (int a = 0; a < 4; a++) {
#pragma unroll
    for (int b = 0; b < 4; b++)
      if (r[b] == a) e[a] = e_in[b];
  }
}

// Number of sub-tetrahedra of the part of a tetrahedron below ef
// (cuttetraN's ntetra).
__device__ inline int sub_count(const double e_in[4], double ef) {
  int r[4];
  double e[4];
  rank4(e_in, r, e);
  if (ef >= e[3] - TW_ETA) return 1;
  if (ef >= e[2] || ef > e[1]) return 3;
  if (ef >= e[0] + TW_ETA) return 1;
  return 0;
}

// Sub-tetrahedron m of the part below ef (cuttetraN's tetra(:,:,m) and
// vtetra(m), tetrahedron.f:366-407): T[i][j] = barycentric coordinate on
// input corner i of sub-corner j.  m < sub_count(e_in, ef).
__device__ inline double sub_tetra(const double e_in[4], double ef, int m, double T[4][4]) {
  int r[4];
  double e[4];
  rank4(e_in, r, e);
  double X[4][4] = {};  // X[s][j]: coordinate on the s-th smallest corner of sub-corner j
  double v;
  if (ef >= e[3] - TW_ETA) {
#pragma unroll
    for (int j = 0; j < 4; j++) X[j][j] = 1.0;
    v = 1.0;
  } else if (ef >= e[2]) {
    const double f0 = (ef - e[0]) / (e[3] - e[0]), f1 = (ef - e[1]) / (e[3] - e[1]),
                 f2 = (ef - e[2]) / (e[3] - e[2]);
    X[0][0] = 1.0;
    if (m == 0) {
      X[1][1] = 1.0; X[2][2] = 1.0; X[1][3] = 1 - f1; X[3][3] = f1;
      v = f1;
    } else if (m == 1) {
      X[0][1] = 1 - f0; X[3][1] = f0; X[1][2] = 1 - f1; X[3][2] = f1; X[2][3] = 1 - f2; X[3][3] = f2;
      v = f0 * (1 - f1) * (1 - f2);
    } else {
      X[2][1] = 1.0; X[1][2] = 1 - f1; X[3][2] = f1; X[2][3] = 1 - f2; X[3][3] = f2;
      v = f2 * (1 - f1);
    }
  } else if (ef > e[1]) {
    const double g0 = (ef - e[0]) / (e[2] - e[0]), g1 = (ef - e[0]) / (e[3] - e[0]),
                 g2 = (ef - e[1]) / (e[2] - e[1]), g3 = (ef - e[1]) / (e[3] - e[1]);
    if (m == 0) {
      X[0][0] = 1.0; X[1][1] = 1.0; X[0][2] = 1 - g0; X[2][2] = g0; X[0][3] = 1 - g1; X[3][3] = g1;
      v = g0 * g1;
    } else if (m == 1) {
      X[1][0] = 1.0; X[0][1] = 1 - g0; X[2][1] = g0; X[0][2] = 1 - g1; X[3][2] = g1;
      X[1][3] = 1 - g3; X[3][3] = g3;
      v = g0 * (1 - g1) * g3;
    } else {
      X[1][0] = 1.0; X[0][1] = 1 - g0; X[2][1] = g0; X[1][2] = 1 - g3; X[3][2] = g3;
      X[1][3] = 1 - g2; X[2][3] = g2;
      v = g2 * (1 - g0) * g3;
    }
  } else {
    const double h0 = (ef - e[0]) / (e[1] - e[0]), h1 = (ef - e[0]) / (e[2] - e[0]),
                 h2 = (ef - e[0]) / (e[3] - e[0]);
    X[0][0] = 1.0; X[0][1] = 1 - h0; X[1][1] = h0; X[0][2] = 1 - h1; X[2][2] = h1;
    X[0][3] = 1 - h2; X[3][3] = h2;
    v = h0 * h1 * h2;
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      double x = 0.0;
#pragma unroll
      for (int s2 = 0; s2 < 4; s2++) x = r[i] == s2 ? X[s2][j] : x;
      T[i][j] = x;
    }
  return v;
}

// One sub-tetrahedron's share for corner q at the bin edge f: rowq = row q of
// its matrix on the original corners, ds = d at its corners.
__device__ inline double sub_share(const double rowq[4], const double ds[4], double v, double f) {
  bool all_ge = true, all_lt = true;
  double sr = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    all_ge = all_ge && ds[j] >= f;
    all_lt = all_lt && ds[j] < f;
    sr += rowq[j];
  }
  if (all_ge) return 0.0;
  if (all_lt) return v * sr;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) acc += rowq[j] * cut_corner(ds, f, j);
  return v * acc;
}

// Corner q's weight of one tetrahedron of a cube whose bands the Fermi
// energy cuts (case (c1), tetrahedron.f:1100-1129 and :1176-1202): the part
// with e1 < ef (occupied) and e2 > ef (unoccupied) as cuttetraN
// sub-tetrahedra (nested when both bands are cut), each then cut by d < f.
__device__ double metal_corner(const double e1[4], const double e2[4], double ef, double f,
                               bool cut1, bool cut2, int q) {
  double dt[4], m2[4];
  bool all_ge = true;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    dt[j] = e2[j] - e1[j];
    all_ge = all_ge && dt[j] >= f;
    m2[j] = -e2[j];
  }
  if (all_ge) return 0.0;
  double w = 0.0, T[4][4], rowq[4], ds[4];
  if (!cut2) {  // occupied part only
    const int nt = sub_count(e1, ef);
    for (int m = 0; m < nt; m++) {
      const double v = sub_tetra(e1, ef, m, T);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        rowq[j] = T[q][j];
        ds[j] = dt[0] * T[0][j] + dt[1] * T[1][j] + dt[2] * T[2][j] + dt[3] * T[3][j];
      }
      w += sub_share(rowq, ds, v, f);
    }
    return w;
  }
  const int nt2 = sub_count(m2, -ef);
  for (int a = 0; a < nt2; a++) {  // unoccupied part (e2 > ef)
    const double v2 = sub_tetra(m2, -ef, a, T);
    double e1s[4], dt2[4], r2[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      e1s[j] = e1[0] * T[0][j] + e1[1] * T[1][j] + e1[2] * T[2][j] + e1[3] * T[3][j];
      dt2[j] = dt[0] * T[0][j] + dt[1] * T[1][j] + dt[2] * T[2][j] + dt[3] * T[3][j];
      r2[j] = T[q][j];
    }
    if (!cut1) {
      w += sub_share(r2, dt2, v2, f);
      continue;
    }
    const int nt1 = sub_count(e1s, ef);
    for (int b = 0; b < nt1; b++) {  // its occupied part
      double T1[4][4];
      const double v1 = sub_tetra(e1s, ef, b, T1);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        rowq[j] = r2[0] * T1[0][j] + r2[1] * T1[1][j] + r2[2] * T1[2][j] + r2[3] * T1[3][j];
        ds[j] = dt2[0] * T1[0][j] + dt2[1] * T1[1][j] + dt2[2] * T1[2][j] + dt2[3] * T1[3][j];
      }
      w += sub_share(rowq, ds, v2 * v1, f);
    }
  }
  return w;
}

template <int DEC>
__global__ __launch_bounds__(TW_NT) void tetra_weights_kernel(
    int n1, int n2, int n3, int nocc, int nunocc, const double *__restrict__ eo,
    const double *__restrict__ eu, double efermi, int nfrq, const double *__restrict__ frq1,
    const double *__restrict__ den, int zero_first, double *__restrict__ W, long ldw) {
  __shared__ double ds[27][TW_NT];
  const long nk = (long)n1 * n2 * n3;
  const long n = nk * nocc * nunocc;
  const int tid = threadIdx.x;
  const long t = (long)blockIdx.x * TW_NT + tid;
  if (t >= n) return;  // no barrier below: each thread reads only its own column of ds
  const int b2 = (int)(t % nunocc);
  const long r = t / nunocc;
  const int b1 = (int)(r % nocc);
  const int k = (int)(r / nocc);
  const int i3 = k % n3, i2 = (k / n3) % n2, i1 = k / (n2 * n3);
  // d at the 27 points k + (a1,a2,a3), a in {-1,0,1}: index (a1+1) + 3(a2+1) + 9(a3+1);
  // bit a of occ_hi / unocc_lo: the occupied energy lies above efermi / the
  // unoccupied one below it there
  double dlo = 1e300, dhi = -1e300;
  unsigned occ_hi = 0, unocc_lo = 0;
  for (int a = 0; a < 27; a++) {
    const long kk = nbr(i1, i2, i3, a, n1, n2, n3);
    const double e1 = eo[kk * nocc + b1], e2 = eu[kk * nunocc + b2];
    const double v = e2 - e1;
    ds[a][tid] = v;
    dlo = v < dlo ? v : dlo;
    dhi = v > dhi ? v : dhi;
    occ_hi |= (e1 > efermi ? 1u : 0u) << a;
    unocc_lo |= (e2 < efermi ? 1u : 0u) << a;
  }
  // per cube s (origin k - (s1,s2,s3)): its corners' bits in the 27-point masks
  unsigned cube_bits[8];
  bool any_cut = false;
#pragma unroll
  for (int s = 0; s < 8; s++) {
    unsigned b = 0;
#pragma unroll
    for (int c = 0; c < 8; c++)
      b |= 1u << (((c & 1) - (s & 1) + 1) + 3 * (((c >> 1) & 1) - ((s >> 1) & 1) + 1) +
                  9 * ((c >> 2) - (s >> 2) + 1));
    cube_bits[s] = b;
    any_cut = any_cut || (occ_hi & b) || (unocc_lo & b);
  }
  double prev = 0.0;
  for (int i = 0; i < nfrq; i++) {
    const double f = frq1[i];
    double w = 0.0;
    if (f > dhi && !any_cut) {
      w = 24.0;  // every tetrahedron of the 8 cubes below f: 24 corner shares
    } else if (f >= dlo) {
      // cube with origin k - (s1,s2,s3); this k is its corner p = s
      for (int s = 0; s < 8; s++) {
        const int s1 = s & 1, s2 = (s >> 1) & 1, s3 = s >> 2;
        const unsigned cb = cube_bits[s];
        if ((unocc_lo & cb) == cb || (occ_hi & cb) == cb) continue;  // no occupied-unoccupied part
        const bool cut1 = (occ_hi & cb) != 0, cut2 = (unocc_lo & cb) != 0;
        double cmin = 1e300, cmax = -1e300;
        for (int c = 0; c < 8; c++) {
          const double v = ds[((c & 1) - s1 + 1) + 3 * (((c >> 1) & 1) - s2 + 1) +
                              9 * ((c >> 2) - s3 + 1)][tid];
          cmin = v < cmin ? v : cmin;
          cmax = v > cmax ? v : cmax;
        }
        if (f < cmin) continue;  // (a) the cube adds nothing
        for (int it = 0; it < 6; it++) {
          const int q = corner_pos<DEC>(it, s);
          if (q < 0) continue;
          if (cut1 || cut2) {  // (c1): the energies at the tetrahedron's corners
            double e1[4], e2[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const int c = kDecomp[DEC][it][j];
              const long kk = nbr(i1, i2, i3,
                                  ((c & 1) - s1 + 1) + 3 * (((c >> 1) & 1) - s2 + 1) +
                                      9 * ((c >> 2) - s3 + 1),
                                  n1, n2, n3);
              e1[j] = eo[kk * nocc + b1];
              e2[j] = eu[kk * nunocc + b2];
            }
            w += metal_corner(e1, e2, efermi, f, cut1, cut2, q);
            continue;
          }
          if (f > cmax) {  // (b) the whole cube: wadd0
            w += 1.0;
            continue;
          }
          double et[4];
          bool all_ge = true, all_lt = true;
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int c = kDecomp[DEC][it][j];
            et[j] = ds[((c & 1) - s1 + 1) + 3 * (((c >> 1) & 1) - s2 + 1) +
                       9 * ((c >> 2) - s3 + 1)][tid];
            all_ge = all_ge && et[j] >= f;
            all_lt = all_lt && et[j] < f;
          }
          if (all_ge) continue;
          w += all_lt ? 1.0 : cut_corner(et, f, q);
        }
      }
    }
    const double out = (i == 0) ? (zero_first ? 0.0 : w / den[0]) : (w - prev) / den[i];
    W[t + ldw * i] = out;
    prev = w;
  }
}

// First and last band of b's degenerate group from a deg row (SPEX's
// convention: deg(b) is the group's last band if b is its first, else the
// first band).
// Returns false (and b's own band as its group) when the row is not a
// consistent deg array (an index outside the window, or first > b > last).
__device__ inline bool deg_group(const int *__restrict__ deg, int nb, int b, int &first,
                                 int &last) {
  const int d = deg[b];
  if (d >= b) {
    first = b;
    last = d;
  } else if (d >= 0) {
    first = d;
    last = deg[d];
  } else {
    first = last = b;
    return false;
  }
  if (last >= nb || last < b || first > b) {
    first = last = b;
    return false;
  }
  return true;
}

// out(t, i) = mean of in over the degenerate group of b1 (side 0, occupied
// states at k) or of b2 (side 1, unoccupied states at k+q), for every
// transition t = (k*nocc + b1)*nunocc + b2 and bin i (tetrahedron.f:
// 1233-1273).  Out of place: every thread reads only `in`.
__global__ void deg_average_kernel(int nk, int nocc, int nunocc, const int *__restrict__ deg,
                                   int side, int nfrq, const double *__restrict__ in, long ldi,
                                   double *__restrict__ out, long ldo, int *__restrict__ flag) {
  const long n = (long)nk * nocc * nunocc;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (t >= n || i >= nfrq) return;
  const int b2 = (int)(t % nunocc);
  const long r = t / nunocc;
  const int b1 = (int)(r % nocc);
  const int k = (int)(r / nocc);
  int first, last;
  double s = 0.0;
  bool ok;
  if (side == 0) {
    ok = deg_group(deg + (long)k * nocc, nocc, b1, first, last);
    for (int b = first; b <= last; b++) s += in[((long)k * nocc + b) * nunocc + b2 + ldi * i];
  } else {
    ok = deg_group(deg + (long)k * nunocc, nunocc, b2, first, last);
    for (int b = first; b <= last; b++) s += in[((long)k * nocc + b1) * nunocc + b + ldi * i];
  }
  if (!ok && i == 0) atomicOr(flag, 1);
  out[t + ldo * i] = s / (last - first + 1);
}

// Row of parent p = sum over its members (full-mesh points, ascending) of
// the full-mesh rows, times scale[p] when a scale table is given; one
// thread per reduced transition (p, b1, b2) and bin.
__global__ void star_sum_kernel(int nkpt1, int nocc, int nunocc, const int *__restrict__ off,
                                const int *__restrict__ mem, const double *__restrict__ scale,
                                int nfrq, const double *__restrict__ in, long ldi,
                                double *__restrict__ out, long ldo) {
  const long per = (long)nocc * nunocc;
  const long nr = (long)nkpt1 * per;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (t >= nr || i >= nfrq) return;
  const int p = (int)(t / per);
  const long bb = t - (long)p * per;
  double s = 0.0;
  for (int m = off[p]; m < off[p + 1]; m++) s += in[(long)mem[m] * per + bb + ldi * i];
  out[t + ldo * i] = scale ? s * scale[p] : s;
}

// Weights of window rows whose degenerate groups may reach beyond the
// window: the mean over the whole groups of the rows of the widened
// windows.  g1/g2 hold (first, last) per (row, window band), relative to
// the widened windows (nx1 x nx2 band pairs per row; window band b sits at
// b + off1 / b + off2); one thread per window transition (r, b1, b2) and
// bin.  clamp (ALIGNBD, where tetrahedron5_init returns before averaging,
// tetrahedron.f:1231): each window row keeps what the clamped band loop
// gave it instead (tetrahedron.f:1086, 1191-1194) -- its own weight, and at
// a window edge also the members of the edge group beyond the window.
__global__ void group_mean_kernel(int nrow, int nocc, int nunocc, int nx1, int nx2, int off1,
                                  int off2, int clamp, const int2 *__restrict__ g1,
                                  const int2 *__restrict__ g2, int nfrq,
                                  const double *__restrict__ in, long ldi,
                                  double *__restrict__ out, long ldo) {
  const long per = (long)nocc * nunocc;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (t >= (long)nrow * per || i >= nfrq) return;
  const int r = (int)(t / per);
  const int b1 = (int)((t - (long)r * per) / nunocc);
  const int b2 = (int)(t - (long)r * per - (long)b1 * nunocc);
  int2 G1 = g1[(long)r * nocc + b1], G2 = g2[(long)r * nunocc + b2];
  if (clamp) {
    G1 = make_int2(b1 == 0 ? G1.x : b1 + off1, b1 == nocc - 1 ? G1.y : b1 + off1);
    G2 = make_int2(b2 == 0 ? G2.x : b2 + off2, b2 == nunocc - 1 ? G2.y : b2 + off2);
  }
  const double *col = in + ldi * i + (long)r * nx1 * nx2;
  double s = 0.0;
  for (int m1 = G1.x; m1 <= G1.y; m1++) {
    double s2 = 0.0;
    for (int m2 = G2.x; m2 <= G2.y; m2++) s2 += col[(long)m1 * nx2 + m2];
    s += s2;
  }
  out[t + ldo * i] = clamp ? s : s / ((double)(G1.y - G1.x + 1) * (double)(G2.y - G2.x + 1));
}

// Mean over a group of equivalent k points: one thread per (group, band
// pair) and bin; members ascending, sum = sum + count*w unfused (the
// reference's rdum loop), divided by the total count and written to every
// member.  The groups of one launch are disjoint, so the update is in place.
__global__ void equiv_average_kernel(long per, int ngroup, const int *__restrict__ off,
                                     const int *__restrict__ mem, const int *__restrict__ cnt,
                                     int nfrq, double *__restrict__ w, long ldw) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (t >= (long)ngroup * per || i >= nfrq) return;
  const int g = (int)(t / per);
  const long bb = t - (long)g * per;
  double *col = w + ldw * i;
  double s = 0.0;
  int tot = 0;
  for (int m = off[g]; m < off[g + 1]; m++) {
    s = s + (double)cnt[m] * col[(long)mem[m] * per + bb];
    tot += cnt[m];
  }
  const double avg = s / (double)tot;
  for (int m = off[g]; m < off[g + 1]; m++) col[(long)mem[m] * per + bb] = avg;
}

// Rows of the reduced set from the full-mesh rows.  Without kpt1: the
// star sum over every point with a parent.  With kpt1 (TETRAF): only the
// points listed in kpt1 contribute, and the row of parent p is scaled by
// the number of mesh points whose parent is p.
int reduce_rows(spexhip_ctx *ctx, const char *name, int nk, int nocc, int nunocc, int nfrqs,
                const void *w_full, long ldw, const int *kpt1, const int *kpt1p, int nkpt1,
                void *w_red, long ldr) {
  SPX_CHECK(ctx, nk > 0 && nocc > 0 && nunocc > 0 && nfrqs > 0 && nkpt1 > 0 && kpt1p,
            "%s: bad arguments", name);
  const long n = (long)nk * nocc * nunocc, nr = (long)nkpt1 * nocc * nunocc;
  SPX_CHECK(ctx, ldw >= n && ldr >= nr, "%s: leading dimensions too small", name);
  std::vector<char> listed(nk, kpt1 ? 0 : 1);
  if (kpt1)
    for (int p = 0; p < nkpt1; p++) {
      SPX_CHECK(ctx, kpt1[p] >= 0 && kpt1[p] < nk, "%s: kpt1 out of range", name);
      listed[kpt1[p]] = 1;
    }
  // members in CSR form, ascending (the order of the reference's scatter)
  std::vector<int> off(nkpt1 + 1, 0), mem;
  std::vector<double> cnt(kpt1 ? nkpt1 : 0, 0.0);
  for (int k = 0; k < nk; k++) {
    SPX_CHECK(ctx, kpt1p[k] >= -1 && kpt1p[k] < nkpt1, "%s: kpt1p out of range", name);
    if (kpt1p[k] < 0) continue;
    if (listed[k]) off[kpt1p[k] + 1]++;
    if (kpt1) cnt[kpt1p[k]] += 1.0;
  }
  for (int p = 0; p < nkpt1; p++) off[p + 1] += off[p];
  mem.resize(off[nkpt1]);
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int k = 0; k < nk; k++)
    if (kpt1p[k] >= 0 && listed[k]) mem[fill[kpt1p[k]]++] = k;
  SPX_CHECK(ctx, spx_owns(ctx, w_full, 8 * ((size_t)ldw * (nfrqs - 1) + n)) &&
                     spx_owns(ctx, w_red, 8 * ((size_t)ldr * (nfrqs - 1) + nr)),
            "%s: arrays are not context allocations of the stated shapes", name);
  SPX_CHECK(ctx, w_full != w_red, "%s: output must not alias the input", name);
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *ws;
  const size_t ints = off.size() + mem.size();
  int rc = spx_scratch(ctx, 8 * cnt.size() + 4 * ints + 16, &ws);
  if (rc) return rc;
  double *dcnt = (double *)ws;
  int *doff = (int *)(dcnt + cnt.size()), *dmem = doff + off.size();
  if (!cnt.empty())
    SPX_HIP(ctx, hipMemcpyAsync(dcnt, cnt.data(), 8 * cnt.size(), hipMemcpyHostToDevice, ctx->stream));
  SPX_HIP(ctx, hipMemcpyAsync(doff, off.data(), 4 * off.size(), hipMemcpyHostToDevice, ctx->stream));
  if (!mem.empty())
    SPX_HIP(ctx, hipMemcpyAsync(dmem, mem.data(), 4 * mem.size(), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(star_sum_kernel, dim3((unsigned)((nr + 255) / 256), (unsigned)nfrqs),
                     dim3(256), 0, ctx->stream, nkpt1, nocc, nunocc, doff, dmem,
                     kpt1 ? (const double *)dcnt : nullptr, nfrqs, (const double *)w_full, ldw,
                     (double *)w_red, ldr);
  SPX_LAUNCHED(ctx);
  // the host tables are read by the copies above
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SPEXHIP_OK;
}

}  // namespace

extern "C" {

int spexhip_tetra_reduce_parents_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc, int nfrqs,
                                     const void *w_full, long ldw, const int *kpt1p, int nkpt1,
                                     void *w_red, long ldr) {
  if (!ctx) return SPEXHIP_E_INVALID;
  return reduce_rows(ctx, "tetra_reduce_parents", nk, nocc, nunocc, nfrqs, w_full, ldw, nullptr,
                     kpt1p, nkpt1, w_red, ldr);
}

int spexhip_tetra_reduce_tetraf_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc, int nfrqs,
                                    const void *w_full, long ldw, const int *kpt1, int nkpt1,
                                    const int *kpt1p, void *w_red, long ldr) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, kpt1 != nullptr, "tetra_reduce_tetraf: kpt1 is NULL");
  return reduce_rows(ctx, "tetra_reduce_tetraf", nk, nocc, nunocc, nfrqs, w_full, ldw, kpt1,
                     kpt1p, nkpt1, w_red, ldr);
}

int spexhip_tetra_average_equivalent_dev(spexhip_ctx *ctx, int nkpt1, int nocc, int nunocc,
                                         int nfrqs, const int *kpt1, int nk, const int *kpt1p,
                                         int nsym, const int *kptsym, int ikpt, void *wtetra,
                                         long ldwt) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nkpt1 > 0 && nocc > 0 && nunocc > 0 && nfrqs > 0 && nk > 0 && nsym > 0 &&
                     kpt1 && kpt1p && kptsym && ikpt >= 0 && ikpt < nk,
            "tetra_average_equivalent: bad arguments");
  const long per = (long)nocc * nunocc, n = (long)nkpt1 * per;
  SPX_CHECK(ctx, ldwt >= n, "tetra_average_equivalent: ldwt < nkpt1*nocc*nunocc");
  SPX_CHECK(ctx, spx_owns(ctx, wtetra, 8 * ((size_t)ldwt * (nfrqs - 1) + n)),
            "tetra_average_equivalent: wtetra is not a context allocation of the stated shape");
  for (int k = 0; k < nkpt1; k++)
    SPX_CHECK(ctx, kpt1[k] >= 0 && kpt1[k] < nk, "tetra_average_equivalent: kpt1 out of range");
  for (int k = 0; k < nk; k++)
    SPX_CHECK(ctx, kpt1p[k] >= -1 && kpt1p[k] < nkpt1,
              "tetra_average_equivalent: kpt1p out of range");
  for (long e = 0; e < (long)nk * nsym; e++)
    SPX_CHECK(ctx, kptsym[e] >= 0 && kptsym[e] < nk,
              "tetra_average_equivalent: kptsym out of range");
  // The groups in the reference's loop order (tetrahedron.f:1282-1318),
  // split into batches of disjoint groups: one launch per batch keeps the
  // reference's sequence where a later group reuses a member.
  std::vector<int> little;
  for (int s = 0; s < nsym; s++)
    if (kptsym[ikpt + (long)nk * s] == ikpt) little.push_back(s);
  std::vector<char> done(nkpt1, 0);
  std::vector<int> count(nkpt1), off(1, 0), mem, cnt, batch(1, 0), owner(nkpt1, -1);
  for (int k = 0; k < nkpt1; k++) {
    if (done[k]) continue;
    std::fill(count.begin(), count.end(), 0);
    bool moved = false;
    for (int s : little) {
      const int k1 = kpt1p[kptsym[kpt1[k] + (long)nk * s]];
      if (k1 >= 0) {
        count[k1]++;
        moved = moved || k1 != k;
      }
    }
    if (!moved) continue;
    const int g = (int)off.size() - 1;
    bool clash = false;
    for (int k1 = 0; k1 < nkpt1; k1++)
      if (count[k1] > 0 && owner[k1] >= batch.back()) clash = true;
    if (clash) batch.push_back(g);
    for (int k1 = 0; k1 < nkpt1; k1++)
      if (count[k1] > 0) {
        mem.push_back(k1);
        cnt.push_back(count[k1]);
        done[k1] = 1;
        owner[k1] = g;
      }
    off.push_back((int)mem.size());
  }
  const int ngroup = (int)off.size() - 1;
  if (ngroup == 0) return SPEXHIP_OK;
  batch.push_back(ngroup);
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *ws;
  int rc = spx_scratch(ctx, 4 * (off.size() + 2 * mem.size()) + 16, &ws);
  if (rc) return rc;
  int *doff = (int *)ws, *dmem = doff + off.size(), *dcnt = dmem + mem.size();
  SPX_HIP(ctx, hipMemcpyAsync(doff, off.data(), 4 * off.size(), hipMemcpyHostToDevice, ctx->stream));
  SPX_HIP(ctx, hipMemcpyAsync(dmem, mem.data(), 4 * mem.size(), hipMemcpyHostToDevice, ctx->stream));
  SPX_HIP(ctx, hipMemcpyAsync(dcnt, cnt.data(), 4 * cnt.size(), hipMemcpyHostToDevice, ctx->stream));
  for (size_t b = 0; b + 1 < batch.size(); b++) {
    const int g0 = batch[b], ng = batch[b + 1] - g0;
    hipLaunchKernelGGL(equiv_average_kernel,
                       dim3((unsigned)(((long)ng * per + 255) / 256), (unsigned)nfrqs), dim3(256),
                       0, ctx->stream, per, ng, doff + g0, dmem, dcnt, nfrqs, (double *)wtetra,
                       ldwt);
    SPX_LAUNCHED(ctx);
  }
  // the host group tables are read by the copies above
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SPEXHIP_OK;
}

int spexhip_tetra_average_degenerate_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc,
                                         const void *deg_occ, const void *deg_unocc, int nfrqs,
                                         void *wtetra, long ldwt) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nk > 0 && nocc > 0 && nunocc > 0 && nfrqs > 0, "tetra_average_degenerate: bad sizes");
  const long n = (long)nk * nocc * nunocc;
  SPX_CHECK(ctx, ldwt >= n, "tetra_average_degenerate: ldwt < n");
  SPX_CHECK(ctx, spx_owns(ctx, deg_occ, 4 * (size_t)nk * nocc) &&
                     spx_owns(ctx, deg_unocc, 4 * (size_t)nk * nunocc) &&
                     spx_owns(ctx, wtetra, 8 * ((size_t)ldwt * (nfrqs - 1) + n)),
            "tetra_average_degenerate: arrays are not context allocations of the stated shapes");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *tmp;
  int rc = spx_scratch(ctx, 8 * (size_t)n * nfrqs + 256, &tmp);
  if (rc) return rc;
  int *flag = (int *)((char *)tmp + 8 * (size_t)n * nfrqs);
  rc = spx_fill_zero(ctx, ctx->stream, flag, sizeof(int));
  if (rc) return rc;
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)nfrqs);
  hipLaunchKernelGGL(deg_average_kernel, grid, dim3(256), 0, ctx->stream, nk, nocc, nunocc,
                     (const int *)deg_occ, 0, nfrqs, (const double *)wtetra, ldwt, (double *)tmp,
                     n, flag);
  SPX_LAUNCHED(ctx);
  hipLaunchKernelGGL(deg_average_kernel, grid, dim3(256), 0, ctx->stream, nk, nocc, nunocc,
                     (const int *)deg_unocc, 1, nfrqs, (const double *)tmp, n, (double *)wtetra,
                     ldwt, flag);
  SPX_LAUNCHED(ctx);
  int h = 0;
  SPX_HIP(ctx, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h)
    return spx_seterr(ctx, SPEXHIP_E_INVALID,
                      "tetra_average_degenerate: a deg array is not consistent (the groups' "
                      "bands were left unaveraged)");
  return SPEXHIP_OK;
}

int spexhip_tetra_weights_dev(spexhip_ctx *ctx, const int *nkpt3, const double *rlat, int nocc,
                              int nunocc, const void *e_occ, const void *e_unocc, double efermi,
                              int nfrqs, const double *frqs, void *wtetra, long ldwt) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nkpt3 && rlat && frqs, "tetra_weights: null argument");
  SPX_CHECK(ctx, nkpt3[0] > 0 && nkpt3[1] > 0 && nkpt3[2] > 0 && nocc > 0 && nunocc > 0,
            "tetra_weights: bad sizes");
  const long nk = (long)nkpt3[0] * nkpt3[1] * nkpt3[2];
  const long n = nk * nocc * nunocc;
  SPX_CHECK(ctx, ldwt >= n, "tetra_weights: ldwt < n");
  // the Hilbert mesh: ascending, exponential, at least two points
  // (tetrahedron.f:951-975); upper bin edges frq1 and the denominators
  // 2 (frq1 - frq) * 24 nkpt of the bin differences (:1224-1231)
  SPX_CHECK(ctx, nfrqs >= 2, "tetra_weights: Number of Hilbert frequencies must be larger than one");
  for (int i = 1; i < nfrqs; i++)
    SPX_CHECK(ctx, frqs[i] > frqs[i - 1], "tetra_weights: input frequency mesh not properly ordered");
  double rdum = 0.0, a = 0.5;
  for (int i = 1; i < nfrqs - 1; i++) {
    const double ai = (frqs[i] - frqs[i - 1]) / (frqs[i + 1] - frqs[i - 1]);
    rdum += ai;
    SPX_CHECK(ctx, std::fabs(rdum / i - ai) <= 1e-10, "tetra_weights: Hilbert mesh not exponential");
  }
  if (nfrqs > 2) a = rdum / (nfrqs - 2);
  std::vector<double> host(2 * (size_t)nfrqs);
  double *f1 = host.data(), *den = f1 + nfrqs;
  for (int i = 0; i < nfrqs - 1; i++) f1[i] = frqs[i] + a * (frqs[i + 1] - frqs[i]);
  f1[nfrqs - 1] = frqs[nfrqs - 1] + (1 - a) * (frqs[nfrqs - 1] - frqs[nfrqs - 2]);
  const double norm = 4.0 * 6 * (double)nk;
  for (int i = 0; i < nfrqs; i++) den[i] = (2 * (f1[i] - frqs[i])) * norm;
  const int idmin = tetra_choice(nkpt3, rlat);
  SPX_CHECK(ctx, spx_owns(ctx, e_occ, 8 * (size_t)nk * nocc) &&
                     spx_owns(ctx, e_unocc, 8 * (size_t)nk * nunocc) &&
                     spx_owns(ctx, wtetra, 8 * ((size_t)ldwt * (nfrqs - 1) + n)),
            "tetra_weights: arrays are not context allocations of the stated shapes");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *ws;
  int rc = spx_scratch(ctx, 16 * (size_t)nfrqs, &ws);
  if (rc) return rc;
  double *dfrq = (double *)ws;
  SPX_HIP(ctx, hipMemcpyAsync(dfrq, host.data(), 16 * (size_t)nfrqs, hipMemcpyHostToDevice,
                              ctx->stream));
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)((n + TW_NT - 1) / TW_NT)), dim3(TW_NT), 0,
                       ctx->stream, nkpt3[0], nkpt3[1], nkpt3[2], nocc, nunocc,
                       (const double *)e_occ, (const double *)e_unocc, efermi, nfrqs, dfrq,
                       dfrq + nfrqs, frqs[0] == 0.0 ? 1 : 0, (double *)wtetra, ldwt);
  };
  switch (idmin) {
    case 0: go(tetra_weights_kernel<0>); break;
    case 1: go(tetra_weights_kernel<1>); break;
    case 2: go(tetra_weights_kernel<2>); break;
    default: go(tetra_weights_kernel<3>); break;
  }
  SPX_LAUNCHED(ctx);
  // the host frequency table is read by the copy above: wait before returning
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SPEXHIP_OK;
}

}  // extern "C"

namespace {

// Host drop-in for tetrahedron5_init (tetrahedron.f:903-1332) on host
// arrays: the full-mesh weights, then (kpt1p given) the rows of the reduced
// set -- the TETRAF rows when kpt1 is given too -- and (deg arrays given)
// the averaging over degenerate states.
int tetrahedron_weights_host(spexhip_ctx *ctx, const int *nkpt3, const double *rlat, int nocc,
                             int nunocc, const double *e_occ, const double *e_unocc,
                             double efermi, int nfrqs, const double *frqs, const int *kpt1,
                             const int *kpt1p, int nkpt1, const int *deg_occ,
                             const int *deg_unocc, double *wtetra) {
  SPX_CHECK(ctx, nkpt3 && e_occ && e_unocc && wtetra && nocc > 0 && nunocc > 0,
            "tetrahedron_weights: bad arguments");
  SPX_CHECK(ctx, (deg_occ == nullptr) == (deg_unocc == nullptr),
            "tetrahedron_weights: give both deg arrays or neither");
  const long nk = (long)nkpt3[0] * nkpt3[1] * nkpt3[2];
  const int nkr = kpt1p ? nkpt1 : (int)nk;
  SPX_CHECK(ctx, nk > 0 && nkr > 0, "tetrahedron_weights: bad sizes");
  const long n = nk * nocc * nunocc, nr = (long)nkr * nocc * nunocc;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *dEo = nullptr, *dEu = nullptr, *dW = nullptr, *dR = nullptr, *dDo = nullptr,
       *dDu = nullptr;
  int rc = spx_alloc(ctx, 8 * (size_t)nk * nocc, &dEo);
  if (!rc) rc = spx_alloc(ctx, 8 * (size_t)nk * nunocc, &dEu);
  if (!rc) rc = spx_alloc(ctx, 8 * (size_t)n * nfrqs, &dW);
  if (!rc && kpt1p) rc = spx_alloc(ctx, 8 * (size_t)nr * nfrqs, &dR);
  if (!rc && deg_occ) rc = spx_alloc(ctx, 4 * (size_t)nkr * nocc, &dDo);
  if (!rc && deg_occ) rc = spx_alloc(ctx, 4 * (size_t)nkr * nunocc, &dDu);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(dEo, e_occ, 8 * (size_t)nk * nocc, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dEu, e_unocc, 8 * (size_t)nk * nunocc, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && deg_occ)
      e = hipMemcpyAsync(dDo, deg_occ, 4 * (size_t)nkr * nocc, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && deg_occ)
      e = hipMemcpyAsync(dDu, deg_unocc, 4 * (size_t)nkr * nunocc, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "tetrahedron_weights: upload failed: %s",
                      hipGetErrorString(e));
  }
  if (!rc)
    rc = spexhip_tetra_weights_dev(ctx, nkpt3, rlat, nocc, nunocc, dEo, dEu, efermi, nfrqs, frqs,
                                   dW, n);
  void *res = dW;
  if (!rc && kpt1p) {
    rc = reduce_rows(ctx, kpt1 ? "tetra_reduce_tetraf" : "tetra_reduce_parents", (int)nk, nocc,
                     nunocc, nfrqs, dW, n, kpt1, kpt1p, nkr, dR, nr);
    res = dR;
  }
  if (!rc && deg_occ && !ctx->alignbd)
    rc = spexhip_tetra_average_degenerate_dev(ctx, nkr, nocc, nunocc, dDo, dDu, nfrqs, res, nr);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(wtetra, res, 8 * (size_t)nr * nfrqs, hipMemcpyDeviceToHost,
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "tetrahedron_weights: download failed: %s",
                      hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(ctx->stream);
  for (void *p : {dEo, dEu, dW, dR, dDo, dDu})
    if (p) spx_free(ctx, p);
  return rc;
}

// The degenerate group (first, last) of band b in one SPEX deg row of nb
// bands (0-based): deg[b] >= b marks the first band of a group and holds its
// last, else deg[b] is the group's first band.  False if the row is not
// consistent there.
bool host_group(const int *deg, int nb, int b, int &first, int &last) {
  const int d = deg[b];
  if (d >= b) {
    first = b;
    last = d;
  } else if (d >= 0) {
    first = d;
    last = deg[d];
  } else {
    return false;
  }
  return first >= 0 && first <= b && last >= b && last < nb;
}

}  // namespace

extern "C" {

int spexhip_set_alignbd(spexhip_ctx *ctx, int on) {
  if (!ctx) return SPEXHIP_E_INVALID;
  ctx->alignbd = on != 0;
  return SPEXHIP_OK;
}

int spexhip_tetrahedron_weights_bands(spexhip_ctx *ctx, const int *nkpt3, const double *rlat,
                                      int nb1, const double *ene1, const int *deg1, int b1low,
                                      int b1up, int nb2, const double *ene2, const int *deg2,
                                      int b2low, int b2up, double efermi, int nfrqs,
                                      const double *frqs, const int *kpt1, int nkpt1,
                                      const int *kpt1p, int tetraf, double *wtetra) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nkpt3 && rlat && ene1 && ene2 && deg1 && deg2 && frqs && wtetra,
            "tetrahedron_weights_bands: null argument");
  SPX_CHECK(ctx, nkpt3[0] > 0 && nkpt3[1] > 0 && nkpt3[2] > 0 && nfrqs > 0,
            "tetrahedron_weights_bands: bad sizes");
  SPX_CHECK(ctx, 0 <= b1low && b1low <= b1up && b1up < nb1 && 0 <= b2low && b2low <= b2up &&
                     b2up < nb2,
            "tetrahedron_weights_bands: band windows outside the band arrays");
  SPX_CHECK(ctx, !kpt1p || (kpt1 && nkpt1 > 0),
            "tetrahedron_weights_bands: a reduced set needs kpt1 and nkpt1");
  const long nk = (long)nkpt3[0] * nkpt3[1] * nkpt3[2];
  const int nocc = b1up - b1low + 1, nunocc = b2up - b2low + 1;
  // widened windows: the whole groups of the edge bands at every k
  // (tetrahedron.f:984-1012)
  int a1 = b1low, z1 = b1up, a2 = b2low, z2 = b2up;
  for (long k = 0; k < nk; k++) {
    int f, l;
    SPX_CHECK(ctx, host_group(deg1 + k * nb1, nb1, b1low, f, l),
              "tetrahedron_weights_bands: deg1 not consistent at k=%ld band %d", k, b1low);
    a1 = std::min(a1, f);
    SPX_CHECK(ctx, host_group(deg1 + k * nb1, nb1, b1up, f, l),
              "tetrahedron_weights_bands: deg1 not consistent at k=%ld band %d", k, b1up);
    z1 = std::max(z1, l);
    SPX_CHECK(ctx, host_group(deg2 + k * nb2, nb2, b2low, f, l),
              "tetrahedron_weights_bands: deg2 not consistent at k=%ld band %d", k, b2low);
    a2 = std::min(a2, f);
    SPX_CHECK(ctx, host_group(deg2 + k * nb2, nb2, b2up, f, l),
              "tetrahedron_weights_bands: deg2 not consistent at k=%ld band %d", k, b2up);
    z2 = std::max(z2, l);
  }
  const int nx1 = z1 - a1 + 1, nx2 = z2 - a2 + 1;
  const int nrow = kpt1p ? nkpt1 : (int)nk;
  // group tables of the rows (groups at kpt1(p) for a reduced set), relative
  // to the widened windows
  std::vector<int2> g1((size_t)nrow * nocc), g2((size_t)nrow * nunocc);
  for (int r = 0; r < nrow; r++) {
    const long kd = kpt1p ? kpt1[r] : r;
    SPX_CHECK(ctx, kd >= 0 && kd < nk, "tetrahedron_weights_bands: kpt1 out of range");
    for (int b = 0; b < nocc; b++) {
      int f, l;
      SPX_CHECK(ctx, host_group(deg1 + kd * nb1, nb1, b1low + b, f, l) && f >= a1 && l <= z1,
                "tetrahedron_weights_bands: deg1 not consistent at k=%ld band %d", kd, b1low + b);
      g1[(size_t)r * nocc + b] = make_int2(f - a1, l - a1);
    }
    for (int b = 0; b < nunocc; b++) {
      int f, l;
      SPX_CHECK(ctx, host_group(deg2 + kd * nb2, nb2, b2low + b, f, l) && f >= a2 && l <= z2,
                "tetrahedron_weights_bands: deg2 not consistent at k=%ld band %d", kd, b2low + b);
      g2[(size_t)r * nunocc + b] = make_int2(f - a2, l - a2);
    }
  }
  // energies of the widened windows
  std::vector<double> x1((size_t)nk * nx1), x2((size_t)nk * nx2);
  for (long k = 0; k < nk; k++) {
    for (int b = 0; b < nx1; b++) x1[(size_t)k * nx1 + b] = ene1[k * nb1 + a1 + b];
    for (int b = 0; b < nx2; b++) x2[(size_t)k * nx2 + b] = ene2[k * nb2 + a2 + b];
  }
  const long nxf = nk * nx1 * nx2, nxr = (long)nrow * nx1 * nx2, nout = (long)nrow * nocc * nunocc;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *dE1 = nullptr, *dE2 = nullptr, *dW = nullptr, *dR = nullptr, *dO = nullptr,
       *dG = nullptr;
  int rc = spx_alloc(ctx, 8 * x1.size(), &dE1);
  if (!rc) rc = spx_alloc(ctx, 8 * x2.size(), &dE2);
  if (!rc) rc = spx_alloc(ctx, 8 * (size_t)nxf * nfrqs, &dW);
  if (!rc && kpt1p) rc = spx_alloc(ctx, 8 * (size_t)nxr * nfrqs, &dR);
  if (!rc) rc = spx_alloc(ctx, 8 * (size_t)nout * nfrqs, &dO);
  if (!rc) rc = spx_alloc(ctx, 8 * (g1.size() + g2.size()), &dG);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(dE1, x1.data(), 8 * x1.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dE2, x2.data(), 8 * x2.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dG, g1.data(), 8 * g1.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync((int2 *)dG + g1.size(), g2.data(), 8 * g2.size(), hipMemcpyHostToDevice,
                         ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "tetrahedron_weights_bands: upload failed: %s",
                      hipGetErrorString(e));
  }
  if (!rc)
    rc = spexhip_tetra_weights_dev(ctx, nkpt3, rlat, nx1, nx2, dE1, dE2, efermi, nfrqs, frqs, dW,
                                   nxf);
  const void *rows = dW;
  long ldr = nxf;
  if (!rc && kpt1p) {
    rc = reduce_rows(ctx, tetraf ? "tetra_reduce_tetraf" : "tetra_reduce_parents", (int)nk, nx1,
                     nx2, nfrqs, dW, nxf, tetraf ? kpt1 : nullptr, kpt1p, nkpt1, dR, nxr);
    rows = dR;
    ldr = nxr;
  }
  if (!rc) {
    hipLaunchKernelGGL(group_mean_kernel, dim3((unsigned)((nout + 255) / 256), (unsigned)nfrqs),
                       dim3(256), 0, ctx->stream, nrow, nocc, nunocc, nx1, nx2, b1low - a1,
                       b2low - a2, ctx->alignbd, (const int2 *)dG, (const int2 *)dG + g1.size(),
                       nfrqs,
                       (const double *)rows, ldr, (double *)dO, nout);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(wtetra, dO, 8 * (size_t)nout * nfrqs, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "tetrahedron_weights_bands: %s",
                      hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(ctx->stream);
  for (void *p : {dE1, dE2, dW, dR, dO, dG})
    if (p) spx_free(ctx, p);
  return rc;
}

int spexhip_tetrahedron_weights(spexhip_ctx *ctx, const int *nkpt3, const double *rlat, int nocc,
                                int nunocc, const double *e_occ, const double *e_unocc,
                                double efermi, int nfrqs, const double *frqs, const int *kpt1p,
                                int nkpt1, const int *deg_occ, const int *deg_unocc,
                                double *wtetra) {
  if (!ctx) return SPEXHIP_E_INVALID;
  return tetrahedron_weights_host(ctx, nkpt3, rlat, nocc, nunocc, e_occ, e_unocc, efermi, nfrqs,
                                  frqs, nullptr, kpt1p, nkpt1, deg_occ, deg_unocc, wtetra);
}

int spexhip_tetrahedron_weights_tetraf(spexhip_ctx *ctx, const int *nkpt3, const double *rlat,
                                       int nocc, int nunocc, const double *e_occ,
                                       const double *e_unocc, double efermi, int nfrqs,
                                       const double *frqs, const int *kpt1, int nkpt1,
                                       const int *kpt1p, const int *deg_occ,
                                       const int *deg_unocc, double *wtetra) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, kpt1 && kpt1p, "tetrahedron_weights_tetraf: kpt1 and kpt1p are required");
  return tetrahedron_weights_host(ctx, nkpt3, rlat, nocc, nunocc, e_occ, e_unocc, efermi, nfrqs,
                                  frqs, kpt1, kpt1p, nkpt1, deg_occ, deg_unocc, wtetra);
}

int spexhip_tetra_average_equivalent(spexhip_ctx *ctx, int nkpt1, int nocc, int nunocc, int nfrqs,
                                     const int *kpt1, int nk, const int *kpt1p, int nsym,
                                     const int *kptsym, int ikpt, double *wtetra) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, wtetra && nkpt1 > 0 && nocc > 0 && nunocc > 0 && nfrqs > 0,
            "tetra_average_equivalent: bad arguments");
  if (ctx->alignbd) return SPEXHIP_OK;  // tetrahedron5_init returns first (tetrahedron.f:1231)
  const size_t bytes = 8 * (size_t)nkpt1 * nocc * nunocc * nfrqs;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *dW = nullptr;
  int rc = spx_alloc(ctx, bytes, &dW);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(dW, wtetra, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "tetra_average_equivalent: upload failed: %s",
                      hipGetErrorString(e));
  }
  if (!rc)
    rc = spexhip_tetra_average_equivalent_dev(ctx, nkpt1, nocc, nunocc, nfrqs, kpt1, nk, kpt1p,
                                              nsym, kptsym, ikpt, dW,
                                              (long)nkpt1 * nocc * nunocc);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(wtetra, dW, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "tetra_average_equivalent: download failed: %s",
                      hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (dW) spx_free(ctx, dW);
  return rc;
}

}  // extern "C"
