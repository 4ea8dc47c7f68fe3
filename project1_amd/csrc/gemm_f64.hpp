// FP64 complex GEMM building blocks on v_mfma_f64_16x16x4f64 (gfx950).
// A 256-thread workgroup (4 waves in a 2x2 layout) owns a 64x64 complex
// output tile; operands are staged through LDS in 16-deep K slices with
// double buffering.
#pragma once
#include "common.hpp"

namespace spxg {

constexpr int TM = 64, KC = 16, NT = 256;

// ------------------------------------------------- complex FP64 GEMM tile --
// acc += A(i0:i0+64, 0:kdim) * B(0:kdim, j0:j0+64); A, B column-major.
// Rows/cols outside [0,m) / [0,n) and k outside [0,kdim) read as zero.
// Three real MFMAs per complex multiply-add (Gauss): the accumulators hold
// t1 = sum ar.br, t2 = sum ai.bi, t3 = sum (ar + ai)(br + bi), and
// tile_foreach hands out re = t1 - t2, im = t3 - t1 - t2.
struct TileAcc {
  d4_t r[2][2], i[2][2], s[2][2];
};

__device__ inline void tile_zero(TileAcc &a) {
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) {
      a.r[x][y] = d4_t{0, 0, 0, 0};
      a.i[x][y] = d4_t{0, 0, 0, 0};
      a.s[x][y] = d4_t{0, 0, 0, 0};
    }
}

struct GemmLds {
  double Ar[2][KC][TM], Ai[2][KC][TM], Br[2][KC][TM], Bi[2][KC][TM];
};

__device__ inline void gemm_tile(const double2 *__restrict__ A, long lda,
                          const double2 *__restrict__ B, long ldb, int m, int n,
                          int kdim, int i0, int j0, TileAcc &acc, GemmLds &s) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1, lr = lane & 15, lk = lane >> 4;
  double2 va[4], vb[4];
  const int nst = (kdim + KC - 1) / KC;
  auto fetch = [&](int st) {
    const int k0 = st * KC;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = tid + NT * r;
      {  // A: i fastest (coalesced down a column)
        const int i = e % TM, k = e / TM;
        const int gi = i0 + i, gk = k0 + k;
        va[r] = (gi < m && gk < kdim) ? A[gi + lda * (long)gk] : make_double2(0, 0);
      }
      {  // B: k fastest
        const int k = e % KC, j = e / KC;
        const int gj = j0 + j, gk = k0 + k;
        vb[r] = (gj < n && gk < kdim) ? B[gk + ldb * (long)gj] : make_double2(0, 0);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = tid + NT * r;
      {
        const int i = e % TM, k = e / TM;
        s.Ar[buf][k][i] = va[r].x;
        s.Ai[buf][k][i] = va[r].y;
      }
      {
        const int k = e % KC, j = e / KC;
        s.Br[buf][k][j] = vb[r].x;
        s.Bi[buf][k][j] = vb[r].y;
      }
    }
  };
  __syncthreads();  // LDS may still be in use by a previous call
  fetch(0);
  store(0);
  __syncthreads();
  for (int st = 0; st < nst; st++) {
    const int buf = st & 1;
    if (st + 1 < nst) fetch(st + 1);
#pragma unroll
    for (int kk = 0; kk < KC / 4; kk++) {
      const int k = kk * 4 + lk;
      double ar[2], ai[2], as[2], br[2], bi[2], bs[2];
#pragma unroll
      for (int x = 0; x < 2; x++) {
        ar[x] = s.Ar[buf][k][wr * 32 + x * 16 + lr];
        ai[x] = s.Ai[buf][k][wr * 32 + x * 16 + lr];
        br[x] = s.Br[buf][k][wc * 32 + x * 16 + lr];
        bi[x] = s.Bi[buf][k][wc * 32 + x * 16 + lr];
        as[x] = ar[x] + ai[x];
        bs[x] = br[x] + bi[x];
      }
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
          acc.r[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], acc.r[x][y], 0, 0, 0);
          acc.i[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], bi[y], acc.i[x][y], 0, 0, 0);
          acc.s[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[x], bs[y], acc.s[x][y], 0, 0, 0);
        }
    }
    if (st + 1 < nst) store(buf ^ 1);
    __syncthreads();
  }
}

// Visit the tile's elements owned by this lane: f(i, j, re, im).  The
// v_mfma_f64_16x16x4f64 result holds column j = lane%16, rows lane/16 + 4*r.
template <class F>
__device__ inline void tile_foreach(const TileAcc &acc, int i0, int j0, F f) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = wv >> 1, wc = wv & 1, lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const double t1 = acc.r[x][y][r], t2 = acc.i[x][y][r];
        f(i0 + wr * 32 + x * 16 + lk + 4 * r, j0 + wc * 32 + y * 16 + lr, t1 - t2,
          acc.s[x][y][r] - t1 - t2);
      }
}


// C_b = alpha*A_b*B_b + beta*C_b + diag*I for batch b = blockIdx.z, with
// X_b = X + (b / dX) * sX for X in {A, B, C}; all column-major.
struct ZgemmArgs {
  int m, n, k;
  double2 alpha;
  const double2 *A; long lda, sA; int dA;
  const double2 *B; long ldb, sB; int dB;
  double beta;
  double2 *C; long ldc, sC; int dC;
  double diag;
  int upper;  // non-zero: only tiles that touch the upper triangle (C Hermitian)
};

}  // namespace spxg

int spx_zgemm_ex(spexhip_ctx *ctx, hipStream_t st, const spxg::ZgemmArgs &a, int batch);
