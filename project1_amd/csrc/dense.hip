// Dielectric step and its inverse in FP64 on the GPU.
//
// Reference: correlation.f:612-621 (eps = 1 - v^1/2 chi v^1/2 per imaginary
// frequency, packed, with p_diagmatdiag_dzpd1 wrapper.f:1264-1278 and
// p_plusdiag1_zp :1317-1328) and correlation.f:701-747 (Minverse per
// frequency -> inverse_zp1 ZHPTRF/ZHPTRI wrapper.f:2362-2374, or
// PZPOTRF/PZPOTRI w_scalapack.inc:752-761 under MPI).
//
// Device matrices are full column-major nb x nb per frequency (ld >= nb,
// frequency stride ld*nb).  v^1/2 is either diagonal (sqrt(coul), the
// Coulomb eigenbasis of the reference) or a dense Hermitian matrix; the
// dense products use an FP64 MFMA complex GEMM.  The inverse is an in-place
// blocked Gauss-Jordan elimination with 64-wide panels: one workgroup
// inverts the diagonal block in registers, an MFMA GEMM forms the row panel,
// and a rank-64 MFMA update sweeps the whole matrix.  eps is Hermitian
// positive definite on the imaginary axis (eigenvalues >= 1), so every
// leading block is HPD and no pivoting is needed there.  For such a matrix
// the working array stays Hermitian up to sign while the sweep runs: with S
// the columns already swept and U the rest, W(S,S) and W(U,U) are Hermitian
// and W(U,S) = -W(S,U)^H.  So only the tiles on and above the diagonal are
// updated (half the rank-64 work), the column and row panels of a step are
// read from the upper triangle (W(i,K) = conj(W(K,i)) below block K,
// W(K,j) = -conj(W(j,K)) left of it), and the lower triangle is filled from
// the upper one at the end.  The real-frequency
// eps (dielecc) is a general complex matrix, which the reference inverts by
// LU with partial pivoting (zgetrf/zgetri, wrapper.f:2357-2358); here each
// panel first picks its pivot rows by an LU with partial pivoting of the
// tall panel (a copy), the rows are interchanged in the whole working array,
// the panel is eliminated as above, and the columns of the result are
// interchanged back in reverse order at the end (Gauss-Jordan with row
// interchanges: inv(A) = inv(PA) P).
#include "gemm_f64.hpp"
#include "internal.hpp"

using namespace spxg;

namespace {

// C = alpha*A*B + beta*C + diag*I, batched over blockIdx.z (see ZgemmArgs).
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2))) void zgemm_kernel(ZgemmArgs a) {
  __shared__ GemmLds s;
  const int b = blockIdx.z;
  const int i0 = blockIdx.x * TM, j0 = blockIdx.y * TM;
  if (a.upper && i0 > j0) return;  // tile wholly below the diagonal
  TileAcc acc;
  tile_zero(acc);
  gemm_tile(a.A + (b / a.dA) * a.sA, a.lda, a.B + (b / a.dB) * a.sB, a.ldb, a.m, a.n, a.k, i0,
            j0, acc, s);
  double2 *Cb = a.C + (b / a.dC) * a.sC;
  tile_foreach(acc, i0, j0, [&](int i, int j, double re, double im) {
    if (i >= a.m || j >= a.n) return;
    double2 *p = Cb + i + a.ldc * (long)j;
    double2 o = make_double2(a.alpha.x * re - a.alpha.y * im, a.alpha.x * im + a.alpha.y * re);
    if (a.beta != 0.0) {
      const double2 old = *p;
      o.x += a.beta * old.x;
      o.y += a.beta * old.y;
    }
    if (i == j) o.x += a.diag;
    *p = o;
  });
}

// Lower triangle <- conj(upper) (I > J), optional real diagonal.  One
// workgroup per 32x32 tile (bi, bj) on or below the diagonal (blockIdx.x =
// bi, blockIdx.y = bj, blockIdx.z = matrix): it reads the mirrored upper
// tile column by column and writes the lower one column by column, the
// transposition going through LDS, so both sides are coalesced.  On a
// diagonal tile the elements read (upper) and written (lower, diagonal)
// are disjoint or read before the barrier.
constexpr int HT = 32;
__global__ __launch_bounds__(256) void hermitize_kernel(int nb, double2 *__restrict__ X, long ldx,
                                                        long sX, int real_diag) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi < bj) return;  // whole workgroup, before the barrier
  __shared__ double2 t[HT][HT + 1];
  double2 *Xb = X + sX * blockIdx.z;
  const int tx = threadIdx.x % HT, ty = threadIdx.x / HT;
  // t[c][r] = X(bj*HT + r, bi*HT + c)
  for (int c = ty; c < HT; c += 256 / HT) {
    const int row = bj * HT + tx, col = bi * HT + c;
    if (row < nb && col < nb) t[c][tx] = Xb[row + ldx * (long)col];
  }
  __syncthreads();
  // X(bi*HT + a, bj*HT + b) = conj(t[a][b]) for row > col
  for (int b = ty; b < HT; b += 256 / HT) {
    const int row = bi * HT + tx, col = bj * HT + b;
    if (row >= nb || col >= nb) continue;
    if (row > col) {
      const double2 v = t[tx][b];
      Xb[row + ldx * (long)col] = make_double2(v.x, -v.y);
    } else if (row == col && real_diag) {
      Xb[row + ldx * (long)row].y = 0.0;
    }
  }
}

// eps = 1 - diag(s) chi diag(s) on the full matrix (correlation.f:616-620).
__global__ void eps_diag_kernel(int nb, const double *__restrict__ sq,
                                const double2 *__restrict__ X, long ldx, long sX,
                                double2 *__restrict__ E, long lde, long sE) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nb * nb) return;
  const int j = (int)(g / nb), i = (int)(g % nb);
  const double2 v = X[sX * blockIdx.y + i + ldx * (long)j];
  const double f = sq[i] * sq[j];
  double2 o = make_double2(-f * v.x, -f * v.y);
  if (i == j) o.x += 1.0;
  E[sE * blockIdx.y + i + lde * (long)j] = o;
}

// ----------------------------------------------- Gauss-Jordan inversion --
constexpr int GB = 64;  // panel width

// P = inverse of the GB x GB diagonal block at (k0,k0) (rows/cols beyond nb
// are treated as identity).  One workgroup of GB/GJ_CW waves per matrix
// (blockIdx.y): wave v keeps columns GJ_CW v .. GJ_CW v + GJ_CW-1 of the
// block in registers, lane = row; several waves per SIMD hide each other's
// barrier, LDS and division latencies.  Pivot step p publishes column p
// (its owner wave) and row p (lane p of every wave) through LDS,
// double-buffered so that one barrier per step suffices: a buffer is
// rewritten two steps later, after every wave has passed the barrier that
// follows its reads.
// logdet (optional): logdet[blockIdx.y] = sum of log(pivot) over this block
// -- the pivots of Gauss-Jordan on an HPD matrix are the Schur-complement
// diagonals, so summed over all blocks they give log det.
// GEN: a general complex matrix (real-frequency eps) whose rows were
// interchanged by gj_pivot_kernel: complex pivots; a zero or non-finite
// pivot sets the flag.
constexpr int GJ_CW = 4;                   // columns per wave
constexpr int GJ_NT = 64 * (GB / GJ_CW);  // 16 waves
template <bool GEN>
__global__ __launch_bounds__(GJ_NT) void gj_diag_kernel(int nb, const double2 *__restrict__ X,
                                                      long ldx, long sX, int k0,
                                                      double2 *__restrict__ P, int *__restrict__ flag,
                                                      double *__restrict__ logdet) {
  static_assert(GB == 64 && GB % GJ_CW == 0, "one lane per row, GJ_CW columns per wave");
  __shared__ double2 colb[2][GB], rowb[2][GB];
  __shared__ double piv[GB];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double2 *Xb = X + sX * blockIdx.y;
  double2 a[GJ_CW];
#pragma unroll
  for (int c = 0; c < GJ_CW; c++) {
    const int j = wv * GJ_CW + c;
    double2 v = make_double2(lane == j ? 1.0 : 0.0, 0.0);
    if (k0 + lane < nb && k0 + j < nb) v = Xb[(k0 + lane) + ldx * (long)(k0 + j)];
    a[c] = v;
  }
  bool bad = false;
  for (int pb = 0; pb < GB / GJ_CW; pb++) {
#pragma unroll
    for (int pc = 0; pc < GJ_CW; pc++) {
      const int p = pb * GJ_CW + pc, buf = p & 1;
      if (wv == pb) colb[buf][lane] = a[pc];
      if (lane == p) {
#pragma unroll
        for (int c = 0; c < GJ_CW; c++) rowb[buf][wv * GJ_CW + c] = a[c];
      }
      __syncthreads();
      const double2 f = colb[buf][lane];
      double2 pr;  // 1 / pivot
      if constexpr (GEN) {
        const double2 d = colb[buf][p];
        const double m2 = d.x * d.x + d.y * d.y;
        bad |= !(m2 > 0.0) || !isfinite(m2);
        pr = make_double2(d.x / m2, -d.y / m2);
      } else {
        // Hermitian positive definite: the pivot is real and positive.
        const double d = colb[buf][p].x;
        bad |= !(d > 0.0);
        if (threadIdx.x == 0) piv[p] = d;
        pr = make_double2(1.0 / d, 0.0);
      }
#pragma unroll
      for (int c = 0; c < GJ_CW; c++) {
        const int j = wv * GJ_CW + c;
        if (j == p) {
          a[c] = lane == p ? pr
                           : make_double2(-(f.x * pr.x - f.y * pr.y), -(f.x * pr.y + f.y * pr.x));
        } else {
          const double2 r = rowb[buf][j];
          const double xr = r.x * pr.x - r.y * pr.y, xi = r.x * pr.y + r.y * pr.x;
          if (lane == p) {
            a[c] = make_double2(xr, xi);
          } else {
            a[c].x -= f.x * xr - f.y * xi;
            a[c].y -= f.x * xi + f.y * xr;
          }
        }
      }
    }
  }
  double2 *Pb = P + (long)GB * GB * blockIdx.y;
#pragma unroll
  for (int c = 0; c < GJ_CW; c++) Pb[lane + GB * (wv * GJ_CW + c)] = a[c];
  if (threadIdx.x == 0 && bad) flag[0] = 1;
  if (!GEN && logdet) {
    __syncthreads();
    if (wv == 0) piv[lane] = log(piv[lane]);
    __syncthreads();
    if (threadIdx.x == 0) {
      double ld = 0.0;
      for (int p = 0; p < GB; p++) ld += piv[p];
      logdet[blockIdx.y] = ld;
    }
  }
}

// Pivot rows of panel k0 (general matrices): LU with partial pivoting of
// the tall panel X(k0:nb, k0:k0+w) on a copy in C (ld ldc), one workgroup
// of 16 waves per matrix (blockIdx.y).  Column p's pivot is the first row
// of largest |Re| + |Im| at or below p (izamax, as zgetrf chooses it);
// piv[k0+p] = its absolute row.  Only the pivot choice is kept: the
// multipliers are not stored, since the Gauss-Jordan step recomputes the
// elimination on the interchanged matrix.  A zero or non-finite pivot sets
// the flag.
__global__ __launch_bounds__(1024) void gj_pivot_kernel(int nb, const double2 *__restrict__ X,
                                                        long ldx, long sX, int k0,
                                                        double2 *__restrict__ Cw, long ldc,
                                                        int *__restrict__ piv, long nbp,
                                                        int *__restrict__ flag) {
  constexpr int NTP = 1024, NWV = NTP / 64;
  __shared__ double bv[NWV];
  __shared__ int bi[NWV];
  __shared__ int prow;
  const double2 *Xb = X + sX * blockIdx.y;
  double2 *C = Cw + ldc * GB * blockIdx.y;
  int *pv = piv + nbp * blockIdx.y;
  const int m = nb - k0, w = m < GB ? m : GB;
  for (long e = threadIdx.x; e < (long)m * w; e += NTP) {
    const int c = (int)(e / m), r = (int)(e % m);
    C[r + ldc * c] = Xb[(k0 + r) + ldx * (long)(k0 + c)];
  }
  __syncthreads();
  bool bad = false;
  for (int p = 0; p < w; p++) {
    double best = -1.0;
    int bidx = p;
    for (int r = p + (int)threadIdx.x; r < m; r += NTP) {
      const double2 z = C[r + ldc * p];
      const double v = fabs(z.x) + fabs(z.y);
      if (v > best) {
        best = v;
        bidx = r;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bidx, off);
      if (ov > best || (ov == best && oi < bidx)) {
        best = ov;
        bidx = oi;
      }
    }
    if ((threadIdx.x & 63) == 0) {
      bv[threadIdx.x >> 6] = best;
      bi[threadIdx.x >> 6] = bidx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double bb = bv[0];
      int ii = bi[0];
      for (int k = 1; k < NWV; k++)
        if (bv[k] > bb || (bv[k] == bb && bi[k] < ii)) {
          bb = bv[k];
          ii = bi[k];
        }
      prow = ii;
      pv[k0 + p] = k0 + ii;
    }
    __syncthreads();
    const int q = prow;
    if (q != p && (int)threadIdx.x < w) {
      const double2 t = C[p + ldc * threadIdx.x];
      C[p + ldc * threadIdx.x] = C[q + ldc * threadIdx.x];
      C[q + ldc * threadIdx.x] = t;
    }
    __syncthreads();
    const double2 d = C[p + ldc * p];
    const double m2 = d.x * d.x + d.y * d.y;
    bad |= !(m2 > 0.0) || !isfinite(m2);
    const double2 inv = m2 > 0.0 ? make_double2(d.x / m2, -d.y / m2) : make_double2(0.0, 0.0);
    for (int r = p + 1 + (int)threadIdx.x; r < m; r += NTP) {
      const double2 a = C[r + ldc * p];
      const double2 l = make_double2(a.x * inv.x - a.y * inv.y, a.x * inv.y + a.y * inv.x);
      for (int c = p + 1; c < w; c++) {
        const double2 u = C[p + ldc * c];
        double2 v = C[r + ldc * c];
        v.x -= l.x * u.x - l.y * u.y;
        v.y -= l.x * u.y + l.y * u.x;
        C[r + ldc * c] = v;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && bad) flag[0] = 1;
}

// Row interchanges of panel k0 in every column of the working array.
__global__ void gj_swap_rows_kernel(int nb, double2 *__restrict__ X, long ldx, long sX, int k0,
                                    const int *__restrict__ piv, long nbp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  double2 *col = X + sX * blockIdx.y + ldx * (long)j;
  const int *pv = piv + nbp * blockIdx.y;
  const int w = nb - k0 < GB ? nb - k0 : GB;
  for (int p = k0; p < k0 + w; p++) {
    const int q = pv[p];
    if (q != p) {
      const double2 t = col[p];
      col[p] = col[q];
      col[q] = t;
    }
  }
}

// inv(A) = inv(PA) P: column interchanges in reverse order, one thread per row.
__global__ void gj_unswap_cols_kernel(int nb, double2 *__restrict__ X, long ldx, long sX,
                                      const int *__restrict__ piv, long nbp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  double2 *Xb = X + sX * blockIdx.y;
  const int *pv = piv + nbp * blockIdx.y;
  for (int p = nb - 1; p >= 0; p--) {
    const int q = pv[p];
    if (q != p) {
      const double2 t = Xb[i + ldx * (long)p];
      Xb[i + ldx * (long)p] = Xb[i + ldx * (long)q];
      Xb[i + ldx * (long)q] = t;
    }
  }
}

// W = diag(s) X diag(s) (correlation.f:868-872, p_diagmatdiag with sqcoul).
__global__ void scale_diag_kernel(int nb, const double *__restrict__ sq,
                                  const double2 *__restrict__ X, long ldx, long sX,
                                  double2 *__restrict__ W, long ldw, long sW) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nb * nb) return;
  const int j = (int)(g / nb), i = (int)(g % nb);
  const double2 v = X[sX * blockIdx.y + i + ldx * (long)j];
  const double f = sq[i] * sq[j];
  W[sW * blockIdx.y + i + ldw * (long)j] = make_double2(f * v.x, f * v.y);
}

// tr[w] = sum_i Re X(i,i,w), one workgroup per matrix, fixed-order tree.
__global__ __launch_bounds__(NT) void trace_kernel(int nb, const double2 *__restrict__ X, long ldx,
                                                   long sX, double *__restrict__ tr) {
  __shared__ double s[NT];
  const double2 *Xb = X + sX * blockIdx.y;
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += NT) a += Xb[i + ldx * (long)i].x;
  s[threadIdx.x] = a;
  __syncthreads();
  for (int h = NT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) tr[blockIdx.y] = s[0];
}

// out[w] = weight * (log det eps + nb - tr eps) = weight * sum(log(e) + 1 - e)
// over the eigenvalues e of eps (correlation.f:686-690).
__global__ void rpa_sum_kernel(int nblk, int nw, int nb, const double *__restrict__ logdet,
                               const double *__restrict__ tr, double weight,
                               double *__restrict__ out) {
  const int w = threadIdx.x;
  if (w >= nw) return;
  double ld = 0.0;
  for (int k = 0; k < nblk; k++) ld += logdet[(long)k * nw + w];
  out[w] = weight * (ld + (double)nb - tr[w]);
}

// Copy the old column panel X(:, k0:k0+GB) into Cp (nb x GB, ld nbp).
__global__ void gj_copy_col_kernel(int nb, const double2 *__restrict__ X, long ldx, long sX,
                                   int k0, double2 *__restrict__ Cp, long ldcp) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nb * GB) return;
  const int c = (int)(g / nb), i = (int)(g % nb);
  double2 v = make_double2(0, 0);
  if (k0 + c < nb) v = X[sX * blockIdx.y + i + ldx * (long)(k0 + c)];
  Cp[(long)ldcp * GB * blockIdx.y + i + ldcp * (long)c] = v;
}

// The column panel Cp = W(:,K) (nb x GB, ld ldcp) and the row panel
// Rs = W(K,:) (GB x nb, ld GB) of the Hermitian sweep, from the upper
// triangle and the diagonal tiles of W.  blockIdx.z = 0: thread per (i, c),
// i fastest, rows i above block K's end (Cp direct; Rs(c,i) = -conj(W(i,k))
// left of K); 1: thread per (c, j), c fastest, columns j >= k0 (Rs direct;
// Cp(j,c) = conj(W(k,j)) below K).  Columns of K past nb are zero.
__global__ void gj_herm_panels_kernel(int nb, const double2 *__restrict__ X, long ldx, long sX,
                                      int k0, double2 *__restrict__ Cp, long ldcp,
                                      double2 *__restrict__ Rs, long sR) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const double2 *Xb = X + sX * b;
  double2 *Cb = Cp + ldcp * (long)GB * b, *Rb = Rs + sR * b;
  const int kend = k0 + GB < nb ? k0 + GB : nb;
  if (blockIdx.z == 0) {
    if (g >= (long)kend * GB) return;
    const int c = (int)(g / kend), i = (int)(g % kend), k = k0 + c;
    const double2 v = k < nb ? Xb[i + ldx * (long)k] : make_double2(0.0, 0.0);
    Cb[i + ldcp * (long)c] = v;
    if (i < k0) Rb[c + (long)GB * i] = make_double2(-v.x, v.y);
  } else {
    const long cnt = (long)(nb - k0) * GB;
    if (g >= cnt) return;
    const int j = k0 + (int)(g / GB), c = (int)(g % GB), k = k0 + c;
    const double2 v = k < nb ? Xb[k + ldx * (long)j] : make_double2(0.0, 0.0);
    Rb[c + (long)GB * j] = v;
    if (j >= kend) Cb[j + ldcp * (long)c] = make_double2(v.x, -v.y);
  }
}

// Rank-GB sweep of Gauss-Jordan step k0 (see file header):
//   A(K,K) = P;  A(K,J) = R(:,J);  A(I,K) = -Cp(I,:) P;  A(I,J) -= Cp(I,:) R(:,J)
// herm: tiles below the diagonal are left alone (the Hermitian sweep).
__global__ __launch_bounds__(NT) void gj_update_kernel(int nb, double2 *__restrict__ X, long ldx,
                                                       long sX, int k0,
                                                       const double2 *__restrict__ P,
                                                       const double2 *__restrict__ R, long ldr, long sR,
                                                       const double2 *__restrict__ Cp, long ldcp,
                                                       int herm) {
  __shared__ GemmLds s;
  const int b = blockIdx.z;
  const int i0 = blockIdx.x * TM, j0 = blockIdx.y * TM;
  if (herm && i0 > j0) return;
  const bool rowK = (i0 == k0), colK = (j0 == k0);
  double2 *Xb = X + sX * b;
  const double2 *Pb = P + (long)GB * GB * b;
  const double2 *Rb = R + sR * b;
  const double2 *Cb = Cp + (long)ldcp * GB * b;
  if (rowK) {
    // A(K, J) = R(:, J)   (for J == K, R(:,K) == P by construction)
    for (int e = threadIdx.x; e < TM * TM; e += NT) {
      const int i = e % TM, j = e / TM;
      if (i0 + i < nb && j0 + j < nb) Xb[(i0 + i) + ldx * (long)(j0 + j)] = colK ? Pb[i + GB * j] : Rb[i + ldr * (long)(j0 + j)];
    }
    return;
  }
  TileAcc acc;
  tile_zero(acc);
  if (colK)
    gemm_tile(Cb, ldcp, Pb, GB, nb, GB, GB, i0, 0, acc, s);
  else
    gemm_tile(Cb, ldcp, Rb, ldr, nb, nb, GB, i0, j0, acc, s);
  tile_foreach(acc, i0, j0, [&](int i, int j, double re, double im) {
    if (i >= nb || j >= nb) return;
    double2 *p = Xb + i + ldx * (long)j;
    if (colK) {
      *p = make_double2(-re, -im);
    } else {
      double2 o = *p;
      o.x -= re;
      o.y -= im;
      *p = o;
    }
  });
}

}  // namespace

// ------------------------------------------------------------ launchers --
int spx_zgemm_ex(spexhip_ctx *ctx, hipStream_t st, const ZgemmArgs &a, int batch) {
  if (a.m <= 0 || a.n <= 0 || batch <= 0) return SPEXHIP_OK;
  dim3 grid((a.m + TM - 1) / TM, (a.n + TM - 1) / TM, batch);
  hipLaunchKernelGGL(zgemm_kernel, grid, dim3(NT), 0, st, a);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spx_zgemm(spexhip_ctx *ctx, hipStream_t st, int m, int n, int k, double alpha,
              const void *A, long lda, long sA, const void *B, long ldb, long sB, double beta,
              void *C, long ldc, long sC, double diag, int batch) {
  ZgemmArgs a{m, n, k, make_double2(alpha, 0.0), (const double2 *)A, lda, sA, 1,
              (const double2 *)B, ldb, sB, 1, beta, (double2 *)C, ldc, sC, 1, diag};
  return spx_zgemm_ex(ctx, st, a, batch);
}

int spx_hermitize(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, void *X, long ldx,
                  int real_diag) {
  const unsigned nt = (unsigned)((nb + HT - 1) / HT);
  hipLaunchKernelGGL(hermitize_kernel, dim3(nt, nt, nw), dim3(256), 0, st, nb, (double2 *)X, ldx,
                     ldx * (long)nb, real_diag);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

// Workspace (bytes) of spx_inverse for nb, nw; the last 256 bytes are free
// for the caller's flag.
size_t spx_inverse_ws(int nb, int nw) {
  const long nbp = ((nb + GB - 1) / GB) * GB;
  const size_t piv = ((4 * (size_t)nw * nbp) + 255) & ~(size_t)255;
  return 16 * (size_t)nw * ((size_t)GB * GB + 3 * (size_t)GB * nbp) + piv + 256;
}

// In-place inverse of nw matrices (full storage): HPD (general = 0) or
// general with partial pivoting (general != 0).  ws: spx_inverse_ws bytes;
// flag: device int, set non-zero if a pivot is not positive (HPD) or is
// zero (general).
int spx_inverse(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, void *X, long ldx, void *ws,
                int *flag, double *logdet_parts, int general) {
  const long nbp = ((nb + GB - 1) / GB) * GB;
  double2 *P = (double2 *)ws;
  double2 *R = P + (long)nw * GB * GB;        // GB x nbp per matrix (ld GB)
  double2 *Cp = R + (long)nw * GB * nbp;      // nbp x GB per matrix (ld nbp)
  double2 *Rs = Cp + (long)nw * nbp * GB;     // GB x nbp per matrix (ld GB): W(K,:), HPD
  int *piv = (int *)(Rs + (long)nw * GB * nbp);  // nbp per matrix (general)
  const long sX = ldx * (long)nb;
  const int nt = (int)(nbp / GB);
  for (int kb = 0; kb < nt; kb++) {
    const int k0 = kb * GB;
    if (general) {
      // Cp is free until gj_copy_col_kernel below: the pivot search's copy
      hipLaunchKernelGGL(gj_pivot_kernel, dim3(1, nw), dim3(1024), 0, st, nb, (const double2 *)X,
                         ldx, sX, k0, Cp, nbp, piv, nbp, flag);
      SPX_LAUNCHED(ctx);
      hipLaunchKernelGGL(gj_swap_rows_kernel, dim3((nb + 255) / 256, nw), dim3(256), 0, st, nb,
                         (double2 *)X, ldx, sX, k0, (const int *)piv, nbp);
      SPX_LAUNCHED(ctx);
      hipLaunchKernelGGL(gj_diag_kernel<true>, dim3(1, nw), dim3(GJ_NT), 0, st, nb,
                         (const double2 *)X, ldx, sX, k0, P, flag, nullptr);
    } else
      hipLaunchKernelGGL(gj_diag_kernel<false>, dim3(1, nw), dim3(GJ_NT), 0, st, nb,
                         (const double2 *)X, ldx, sX, k0, P, flag,
                         logdet_parts ? logdet_parts + (long)kb * nw : nullptr);
    SPX_LAUNCHED(ctx);
    const int kk = nb - k0 < GB ? nb - k0 : GB;
    if (general) {
      // R = P * X(K, :)   (GB x nb), batched; then the old column panel
      int rc = spx_zgemm(ctx, st, GB, nb, kk, 1.0, P, GB, (long)GB * GB, (const double2 *)X + k0,
                         ldx, sX, 0.0, R, GB, (long)GB * nbp, 0.0, nw);
      if (rc) return rc;
      const long tot = (long)nb * GB;
      hipLaunchKernelGGL(gj_copy_col_kernel, dim3((tot + 255) / 256, nw), dim3(256), 0, st, nb,
                         (const double2 *)X, ldx, sX, k0, Cp, nbp);
      SPX_LAUNCHED(ctx);
    } else {
      // both panels from the upper triangle, then R = P * W(K, :)
      const long tot = (long)nb * GB;
      hipLaunchKernelGGL(gj_herm_panels_kernel, dim3((unsigned)((tot + 255) / 256), nw, 2),
                         dim3(256), 0, st, nb, (const double2 *)X, ldx, sX, k0, Cp, nbp, Rs,
                         (long)GB * nbp);
      SPX_LAUNCHED(ctx);
      int rc = spx_zgemm(ctx, st, GB, nb, kk, 1.0, P, GB, (long)GB * GB, Rs, GB, (long)GB * nbp,
                         0.0, R, GB, (long)GB * nbp, 0.0, nw);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(gj_update_kernel, dim3(nt, nt, nw), dim3(NT), 0, st, nb, (double2 *)X, ldx,
                       sX, k0, (const double2 *)P, (const double2 *)R, (long)GB, (long)GB * nbp,
                       (const double2 *)Cp, nbp, general ? 0 : 1);
    SPX_LAUNCHED(ctx);
  }
  if (general) {
    hipLaunchKernelGGL(gj_unswap_cols_kernel, dim3((nb + 255) / 256, nw), dim3(256), 0, st, nb,
                       (double2 *)X, ldx, sX, (const int *)piv, nbp);
    SPX_LAUNCHED(ctx);
  } else {
    // the lower triangle from the upper one
    int rc = spx_hermitize(ctx, st, nb, nw, X, ldx, 0);
    if (rc) return rc;
  }
  return SPEXHIP_OK;
}

// eps = 1 - Vh chi Vh.  vkind 0: vh = real vector sqrt(coul); 1: dense
// complex nb x nb (column-major, ld nb).  chi is Hermitized first, with a
// real diagonal: the reference's diagonal is real by construction (re(t) =
// |cprod|^2, im(t) = 0 for i = j, susceptibility.f:1195-1201), while a split
// form leaves rounding in Im chi(I,I) (about 1e-10 of max|chi| in FP32
// pieces) that eps^-1 would amplify by its condition number.  E may alias X
// only for vkind 0.  tmp: nb*nb*nw complex (vkind 1).
int spx_dielectric(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, void *X, long ldx, int vkind,
                   const void *vh, void *E, long lde, void *tmp) {
  int rc = spx_hermitize(ctx, st, nb, nw, X, ldx, 1);
  if (rc) return rc;
  const long nn = (long)nb * nb;
  if (vkind == 0) {
    hipLaunchKernelGGL(eps_diag_kernel, dim3((nn + 255) / 256, nw), dim3(256), 0, st, nb,
                       (const double *)vh, (const double2 *)X, ldx, ldx * (long)nb, (double2 *)E,
                       lde, lde * (long)nb);
    SPX_LAUNCHED(ctx);
    return SPEXHIP_OK;
  }
  // tmp = chi * Vh ; E = I - Vh * tmp on the tiles on and above the
  // diagonal (E is Hermitian), then the lower triangle from the upper
  rc = spx_zgemm(ctx, st, nb, nb, nb, 1.0, X, ldx, ldx * (long)nb, vh, nb, 0, 0.0, tmp, nb, nn, 0.0, nw);
  if (rc) return rc;
  ZgemmArgs e{nb, nb, nb, make_double2(-1.0, 0.0), (const double2 *)vh, nb, 0, 1,
              (const double2 *)tmp, nb, nn, 1, 0.0, (double2 *)E, lde, lde * (long)nb, 1, 1.0};
  e.upper = 1;
  rc = spx_zgemm_ex(ctx, st, e, nw);
  if (rc) return rc;
  return spx_hermitize(ctx, st, nb, nw, E, lde, 0);
}

// Real-frequency eps = 1 - Vh chi Vh of a general (non-Hermitian) chi
// (correlation.f:632-636: p_diagmatdiag(-sqcoul, suscepc, sqcoul),
// p_plusdiag): no Hermitization, both triangles computed.  E may alias X
// only for vkind 0.  tmp: nb*nb*nw complex (vkind 1).
int spx_dielectric_general(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, const void *X,
                           long ldx, int vkind, const void *vh, void *E, long lde, void *tmp) {
  const long nn = (long)nb * nb;
  if (vkind == 0) {
    hipLaunchKernelGGL(eps_diag_kernel, dim3((nn + 255) / 256, nw), dim3(256), 0, st, nb,
                       (const double *)vh, (const double2 *)X, ldx, ldx * (long)nb, (double2 *)E,
                       lde, lde * (long)nb);
    SPX_LAUNCHED(ctx);
    return SPEXHIP_OK;
  }
  int rc = spx_zgemm(ctx, st, nb, nb, nb, 1.0, X, ldx, ldx * (long)nb, vh, nb, 0, 0.0, tmp, nb, nn,
                     0.0, nw);
  if (rc) return rc;
  ZgemmArgs e{nb, nb, nb, make_double2(-1.0, 0.0), (const double2 *)vh, nb, 0, 1,
              (const double2 *)tmp, nb, nn, 1, 0.0, (double2 *)E, lde, lde * (long)nb, 1, 1.0};
  return spx_zgemm_ex(ctx, st, e, nw);
}

// W = Vh Einv Vh (correlation.f:868-877).  vkind 0: vh = sqrt(coul)
// vector; 1: dense nb x nb.  tmp: nb*nb*nw complex (vkind 1); W != Einv.
int spx_screen(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, const void *Einv, long lde,
               int vkind, const void *vh, void *W, long ldw, void *tmp) {
  const long nn = (long)nb * nb;
  if (vkind == 0) {
    hipLaunchKernelGGL(scale_diag_kernel, dim3((nn + 255) / 256, nw), dim3(256), 0, st, nb,
                       (const double *)vh, (const double2 *)Einv, lde, lde * (long)nb,
                       (double2 *)W, ldw, ldw * (long)nb);
    SPX_LAUNCHED(ctx);
    return SPEXHIP_OK;
  }
  int rc = spx_zgemm(ctx, st, nb, nb, nb, 1.0, Einv, lde, lde * (long)nb, vh, nb, 0, 0.0, tmp, nb,
                     nn, 0.0, nw);
  if (rc) return rc;
  return spx_zgemm(ctx, st, nb, nb, nb, 1.0, vh, nb, 0, tmp, nb, nn, 0.0, W, ldw, ldw * (long)nb,
                   0.0, nw);
}

extern "C" {

int spexhip_screened_w_dev(spexhip_ctx *ctx, int nb, int nw, const void *einv, int lde, int vkind,
                           const void *vhalf, void *W, int ldw) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && lde >= nb && ldw >= nb, "screened W: bad sizes");
  SPX_CHECK(ctx, vkind == 0 || vkind == 1, "screened W: vkind must be 0 (diagonal) or 1 (dense)");
  SPX_CHECK(ctx, einv != W, "screened W: W must not alias eps^-1");
  SPX_CHECK(ctx, spx_owns(ctx, einv, 16 * (size_t)lde * nb * nw) && spx_owns(ctx, W, 16 * (size_t)ldw * nb * nw),
            "screened W: eps^-1/W not context allocations of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, vhalf, vkind ? 16 * (size_t)nb * nb : 8 * (size_t)nb),
            "screened W: vhalf not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *tmp = nullptr;
  if (vkind) {
    int rc = spx_scratch(ctx, 16 * (size_t)nb * nb * nw, &tmp);
    if (rc) return rc;
  }
  return spx_screen(ctx, ctx->stream, nb, nw, einv, lde, vkind, vhalf, W, ldw, tmp);
}

int spexhip_rpa_integrand_dev(spexhip_ctx *ctx, int nb, int nw, const void *eps, int lde,
                              double weight, double *out) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && nw <= 1024 && lde >= nb && out, "rpa integrand: bad arguments");
  SPX_CHECK(ctx, spx_owns(ctx, eps, 16 * (size_t)lde * nb * nw),
            "rpa integrand: eps not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const int nblk = (nb + GB - 1) / GB;
  const size_t cb = 16 * (size_t)nb * nb * nw, ib = spx_inverse_ws(nb, nw);
  const size_t tail = 8 * ((size_t)nblk * nw + 2 * (size_t)nw) + 256;
  void *ws;
  int rc = spx_scratch(ctx, cb + ib + tail, &ws);
  if (rc) return rc;
  double2 *E = (double2 *)ws;  // copy of eps: the inversion is destructive
  double *parts = (double *)((char *)ws + cb + ib);
  double *tr = parts + (size_t)nblk * nw, *res = tr + nw;
  int *flag = (int *)(res + nw);
  SPX_HIP(ctx, hipMemcpy2DAsync(E, 16 * (size_t)nb, eps, 16 * (size_t)lde, 16 * (size_t)nb,
                                (size_t)nb * nw, hipMemcpyDeviceToDevice, ctx->stream));
  SPX_HIP(ctx, hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL(trace_kernel, dim3(1, nw), dim3(NT), 0, ctx->stream, nb, (const double2 *)E,
                     (long)nb, (long)nb * nb, tr);
  SPX_LAUNCHED(ctx);
  rc = spx_inverse(ctx, ctx->stream, nb, nw, E, nb, (char *)ws + cb, flag, parts);
  if (rc) return rc;
  hipLaunchKernelGGL(rpa_sum_kernel, dim3(1), dim3(1024), 0, ctx->stream, nblk, nw, nb, parts, tr,
                     weight, res);
  SPX_LAUNCHED(ctx);
  int h = 0;
  SPX_HIP(ctx, hipMemcpyAsync(out, res, 8 * (size_t)nw, hipMemcpyDeviceToHost, ctx->stream));
  SPX_HIP(ctx, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h) return spx_seterr(ctx, SPEXHIP_E_NOTPD, "rpa integrand: eps not positive definite");
  return SPEXHIP_OK;
}

// Host drop-ins on packed upper storage (the reference dielec / screen).
int spexhip_screened_w_packed(spexhip_ctx *ctx, int nb, int nw, const double *einv_packed,
                              int vkind, const double *vhalf, double *w_packed) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && einv_packed && vhalf && w_packed && (vkind == 0 || vkind == 1),
            "screened W (packed): bad arguments");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t mb = 16 * (size_t)nb * nb * nw, vb = vkind ? 16 * (size_t)nb * nb : 8 * (size_t)nb;
  void *E = nullptr, *W = nullptr, *V = nullptr;
  int rc = spx_alloc(ctx, mb, &E);
  if (!rc) rc = spx_alloc(ctx, mb, &W);
  if (!rc) rc = spx_alloc(ctx, vb, &V);
  if (!rc && hipMemcpy(V, vhalf, vb, hipMemcpyHostToDevice) != hipSuccess)
    rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "screened W: upload failed");
  if (!rc) rc = spexhip_unpack_h2d(ctx, nb, nw, einv_packed, E, nb);
  if (!rc) rc = spexhip_screened_w_dev(ctx, nb, nw, E, nb, vkind, V, W, nb);
  if (!rc) rc = spexhip_pack_upper_d2h(ctx, nb, nw, W, nb, w_packed, 0);
  (void)hipStreamSynchronize(ctx->stream);
  spx_free(ctx, E);
  spx_free(ctx, W);
  spx_free(ctx, V);
  return rc;
}

int spexhip_rpa_integrand_packed(spexhip_ctx *ctx, int nb, int nw, const double *eps_packed,
                                 double weight, double *out) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && eps_packed && out, "rpa integrand (packed): bad arguments");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *E = nullptr;
  int rc = spx_alloc(ctx, 16 * (size_t)nb * nb * nw, &E);
  if (!rc) rc = spexhip_unpack_h2d(ctx, nb, nw, eps_packed, E, nb);
  if (!rc) rc = spexhip_rpa_integrand_dev(ctx, nb, nw, E, nb, weight, out);
  (void)hipStreamSynchronize(ctx->stream);
  spx_free(ctx, E);
  return rc;
}

int spexhip_dielectric_dev(spexhip_ctx *ctx, int nb, int nw, void *X, int ldx, int vkind,
                           const void *vhalf, void *E, int lde) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && ldx >= nb && lde >= nb, "dielectric: bad sizes");
  SPX_CHECK(ctx, vkind == 0 || vkind == 1, "dielectric: vkind must be 0 (diagonal) or 1 (dense)");
  SPX_CHECK(ctx, spx_owns(ctx, X, 16 * (size_t)ldx * nb * nw) && spx_owns(ctx, E, 16 * (size_t)lde * nb * nw),
            "dielectric: X/E not context allocations of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, vhalf, vkind ? 16 * (size_t)nb * nb : 8 * (size_t)nb),
            "dielectric: vhalf not a context allocation of the stated shape");
  SPX_CHECK(ctx, vkind == 0 || X != E, "dielectric: dense v^1/2 needs E != X");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *tmp = nullptr;
  if (vkind) {
    int rc = spx_scratch(ctx, 16 * (size_t)nb * nb * nw, &tmp);
    if (rc) return rc;
  }
  return spx_dielectric(ctx, ctx->stream, nb, nw, X, ldx, vkind, vhalf, E, lde, tmp);
}

int spexhip_inverse_hpd_dev(spexhip_ctx *ctx, int nb, int nw, void *X, int ldx) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && ldx >= nb, "inverse: bad sizes");
  SPX_CHECK(ctx, spx_owns(ctx, X, 16 * (size_t)ldx * nb * nw), "inverse: X not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *ws;
  int rc = spx_scratch(ctx, spx_inverse_ws(nb, nw), &ws);
  if (rc) return rc;
  int *flag = (int *)((char *)ws + spx_inverse_ws(nb, nw) - 256);
  SPX_HIP(ctx, hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
  rc = spx_inverse(ctx, ctx->stream, nb, nw, X, ldx, ws, flag, nullptr);
  if (rc) return rc;
  int h = 0;
  SPX_HIP(ctx, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h) return spx_seterr(ctx, SPEXHIP_E_NOTPD, "inverse: matrix not positive definite");
  return SPEXHIP_OK;
}

int spexhip_dielectric_general_dev(spexhip_ctx *ctx, int nb, int nw, const void *X, int ldx,
                                   int vkind, const void *vhalf, void *E, int lde) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && ldx >= nb && lde >= nb, "dielectric_general: bad sizes");
  SPX_CHECK(ctx, vkind == 0 || vkind == 1, "dielectric_general: vkind must be 0 or 1");
  SPX_CHECK(ctx, spx_owns(ctx, X, 16 * (size_t)ldx * nb * nw) && spx_owns(ctx, E, 16 * (size_t)lde * nb * nw),
            "dielectric_general: X/E not context allocations of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, vhalf, vkind ? 16 * (size_t)nb * nb : 8 * (size_t)nb),
            "dielectric_general: vhalf not a context allocation of the stated shape");
  SPX_CHECK(ctx, vkind == 0 || X != E, "dielectric_general: dense v^1/2 needs E != X");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *tmp = nullptr;
  if (vkind) {
    int rc = spx_scratch(ctx, 16 * (size_t)nb * nb * nw, &tmp);
    if (rc) return rc;
  }
  return spx_dielectric_general(ctx, ctx->stream, nb, nw, X, ldx, vkind, vhalf, E, lde, tmp);
}

int spexhip_inverse_general_dev(spexhip_ctx *ctx, int nb, int nw, void *X, int ldx) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && ldx >= nb, "inverse_general: bad sizes");
  SPX_CHECK(ctx, spx_owns(ctx, X, 16 * (size_t)ldx * nb * nw),
            "inverse_general: X not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *ws;
  int rc = spx_scratch(ctx, spx_inverse_ws(nb, nw), &ws);
  if (rc) return rc;
  int *flag = (int *)((char *)ws + spx_inverse_ws(nb, nw) - 256);
  rc = spx_fill_zero(ctx, ctx->stream, flag, sizeof(int));
  if (rc) return rc;
  rc = spx_inverse(ctx, ctx->stream, nb, nw, X, ldx, ws, flag, nullptr, 1);
  if (rc) return rc;
  int h = 0;
  SPX_HIP(ctx, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h) return spx_seterr(ctx, SPEXHIP_E_NOTPD, "inverse_general: zero or non-finite pivot");
  return SPEXHIP_OK;
}

// Host drop-in for the real-frequency branch (dielecc, correlation.f:
// 632-636 and :701-747): suscepc (nb x nb x nw, full) -> its eps^-1 (same
// layout), v^1/2 = sqcoul (vkind 0) or dense (vkind 1).
int spexhip_dielectric_inverse_general(spexhip_ctx *ctx, int nb, int nw, const double *chi,
                                       int vkind, const double *vhalf, double *einv) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && chi && vhalf && einv && (vkind == 0 || vkind == 1),
            "dielectric_inverse_general: bad arguments");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t xb = 16 * (size_t)nb * nb * nw;
  const size_t vb = vkind ? 16 * (size_t)nb * nb : 8 * (size_t)nb;
  void *X = nullptr, *E = nullptr, *V = nullptr;
  int rc = spx_alloc(ctx, xb, &X);
  if (!rc) rc = spx_alloc(ctx, xb, &E);
  if (!rc) rc = spx_alloc(ctx, vb, &V);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(X, chi, xb, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(V, vhalf, vb, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "upload: %s", hipGetErrorString(e));
  }
  if (!rc) rc = spexhip_dielectric_general_dev(ctx, nb, nw, X, nb, vkind, V, E, nb);
  if (!rc) rc = spexhip_inverse_general_dev(ctx, nb, nw, E, nb);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(einv, E, xb, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "download: %s", hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(ctx->stream);
  spx_free(ctx, X);
  spx_free(ctx, E);
  spx_free(ctx, V);
  return rc;
}

int spexhip_dielectric_inverse(spexhip_ctx *ctx, int nb, int nw, const double *chi_packed,
                               int vkind, const double *vhalf, double *einv_packed) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && chi_packed && vhalf && einv_packed, "dielectric_inverse: bad arguments");
  SPX_CHECK(ctx, vkind == 0 || vkind == 1, "dielectric_inverse: vkind must be 0 or 1");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t xb = 16 * (size_t)nb * nb * nw;
  const size_t vb = vkind ? 16 * (size_t)nb * nb : 8 * (size_t)nb;
  void *X = nullptr, *E = nullptr, *V = nullptr;
  int rc = spx_alloc(ctx, xb, &X);
  if (!rc) rc = spx_alloc(ctx, xb, &E);
  if (!rc) rc = spx_alloc(ctx, vb, &V);
  if (!rc) rc = spexhip_unpack_h2d(ctx, nb, nw, chi_packed, X, nb);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(V, vhalf, vb, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "upload: %s", hipGetErrorString(e));
  }
  if (!rc) rc = spexhip_dielectric_dev(ctx, nb, nw, X, nb, vkind, V, E, nb);
  if (!rc) rc = spexhip_inverse_hpd_dev(ctx, nb, nw, E, nb);
  if (!rc) rc = spexhip_pack_upper_d2h(ctx, nb, nw, E, nb, einv_packed, 0);
  (void)hipStreamSynchronize(ctx->stream);
  spx_free(ctx, X);
  spx_free(ctx, E);
  spx_free(ctx, V);
  return rc;
}

int spexhip_inverse_hpd_packed(spexhip_ctx *ctx, int nb, int count, double *packed) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && count > 0 && packed, "inverse_hpd_packed: bad arguments");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *X = nullptr;
  int rc = spx_alloc(ctx, 16 * (size_t)nb * nb * count, &X);
  if (!rc) rc = spexhip_unpack_h2d(ctx, nb, count, packed, X, nb);
  if (!rc) rc = spexhip_inverse_hpd_dev(ctx, nb, count, X, nb);
  if (!rc) rc = spexhip_pack_upper_d2h(ctx, nb, count, X, nb, packed, 0);
  (void)hipStreamSynchronize(ctx->stream);
  spx_free(ctx, X);
  return rc;
}

}  // extern "C"
