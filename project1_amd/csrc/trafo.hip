// The products in the eigenbasis of the Coulomb matrix, replacing
// susceptibility.f:806-832 (noinv build):
//
//   cprod = conjg(cprod);  cprod(:, :dim) = cprod . T;  cprod = conjg(cprod)
//
// i.e. M'(t, j) = sum_I M(t, I) conj(T(I, j)) for the dim <= nbasm leading
// eigenvectors.  T is ctrafo for products <M phi|phi> (fullpw), else
// ctrafo1: ctrafo with every interstitial column passed through olapvec
// (overlap.f:298-315) over the mixed-basis plane waves of the q point,
// which spexhip_ctrafo_olap_dev forms on the device.
//
// Transition-major M (t*ldm + I) read column-major is X(I, t), so the
// transformation is the GEMM Y = T^H X (dim x n): T^H is formed once
// (ctrafo_h_kernel), then the FP64 MFMA GEMM of dense.hip runs over
// blocks of transitions.  In place (out == M) each block goes through a
// scratch buffer before it is copied back, so a block only reads its own
// rows.
#include <algorithm>
#include <vector>

#include "gemm_f64.hpp"

int spx_zgemm(spexhip_ctx *ctx, hipStream_t st, int m, int n, int k, double alpha,
              const void *A, long lda, long sA, const void *B, long ldb, long sB, double beta,
              void *C, long ldc, long sC, double diag, int batch);

namespace {

constexpr int NT = 256;

// H(j, I) = conj(T(I, j)), H dim x nbasm column-major (ld dim).
__global__ void ctrafo_h_kernel(int nbasm, int dim, const double2 *__restrict__ T, long ldt,
                                double2 *__restrict__ H) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nbasm * dim) return;
  const int I = (int)(g % nbasm), j = (int)(g / nbasm);
  const double2 v = T[I + ldt * (long)j];
  H[j + (long)dim * I] = make_double2(v.x, -v.y);
}

// out[t*ldo + j] = Y(j, t - t0) for the block, Y dim x nt column-major.
__global__ void copy_rows_kernel(long nt, int dim, const double2 *__restrict__ Y,
                                 double2 *__restrict__ out, long ldo) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nt * dim) return;
  const long t = g / dim;
  const int j = (int)(g % dim);
  out[t * ldo + j] = Y[g];
}

// ctrafo1 (susceptibility.f:815-819): rows < nbasp copied, interstitial rows
//   T1(nbasp + i, c) = sum_j conj(cstep(G_j - G_i)) T(nbasp + j, c)
// One thread per (row i, eight columns); 64 plane waves of G vectors and
// column values at a time in LDS.
constexpr int OC = 8;
__global__ __launch_bounds__(NT) void ctrafo_olap_kernel(int nbasp, int ngptm, int dim,
                                                         const int *__restrict__ G,
                                                         const double2 *__restrict__ cstep, int c1,
                                                         int c2, int c3,
                                                         const double2 *__restrict__ T, long ldt,
                                                         double2 *__restrict__ T1, long ldt1) {
  __shared__ int Gs[64][3];
  __shared__ double2 Cs[64][OC];
  const int i = blockIdx.x * NT + threadIdx.x;
  const int cb = blockIdx.y * OC;
  int gi[3] = {0, 0, 0};
  if (i < ngptm)
    for (int d = 0; d < 3; d++) gi[d] = G[3 * i + d];
  double2 acc[OC];
#pragma unroll
  for (int q = 0; q < OC; q++) acc[q] = make_double2(0.0, 0.0);
  const int w0 = 2 * c1 + 1, w1 = 2 * c2 + 1;
  for (int j0 = 0; j0 < ngptm; j0 += 64) {
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 3; e += NT) {
      const int jj = j0 + e / 3;
      Gs[e / 3][e % 3] = jj < ngptm ? G[3 * jj + e % 3] : 0;
    }
    for (int e = threadIdx.x; e < 64 * OC; e += NT) {
      const int jj = j0 + e / OC, c = cb + e % OC;
      Cs[e / OC][e % OC] = (jj < ngptm && c < dim) ? T[nbasp + jj + ldt * (long)c]
                                                   : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (i < ngptm) {
      const int len = ngptm - j0 < 64 ? ngptm - j0 : 64;
      for (int jj = 0; jj < len; jj++) {
        const int dx = Gs[jj][0] - gi[0], dy = Gs[jj][1] - gi[1], dz = Gs[jj][2] - gi[2];
        const double2 s = cstep[((long)(dz + c3) * w1 + (dy + c2)) * w0 + (dx + c1)];
#pragma unroll
        for (int q = 0; q < OC; q++) {  // conj(s) * v
          const double2 v = Cs[jj][q];
          acc[q].x += s.x * v.x + s.y * v.y;
          acc[q].y += s.x * v.y - s.y * v.x;
        }
      }
    }
  }
  if (i >= ngptm) return;
#pragma unroll
  for (int q = 0; q < OC; q++) {
    const int c = cb + q;
    if (c < dim) T1[nbasp + i + ldt1 * (long)c] = acc[q];
  }
}

__global__ void copy_mt_rows_kernel(int nbasp, int dim, const double2 *__restrict__ T, long ldt,
                                    double2 *__restrict__ T1, long ldt1) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nbasp * dim) return;
  const int I = (int)(g % nbasp), c = (int)(g / nbasp);
  T1[I + ldt1 * (long)c] = T[I + ldt * (long)c];
}

}  // namespace

extern "C" {

int spexhip_coulomb_trafo_dev(spexhip_ctx *ctx, long n, int nbasm, void *M, long ldm,
                              const void *T, long ldt, int dim, void *out, long ldo) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, n >= 0 && nbasm > 0 && dim > 0 && dim <= nbasm && ldm >= nbasm && ldt >= nbasm,
            "coulomb_trafo: bad sizes");
  const bool inplace = out == nullptr || out == M;
  if (inplace) {
    out = M;
    ldo = ldm;
  }
  SPX_CHECK(ctx, ldo >= dim, "coulomb_trafo: ldo < dim");
  if (n == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, spx_owns(ctx, M, 16 * ((size_t)ldm * (n - 1) + nbasm)), "coulomb_trafo: M is not a context allocation of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, T, 16 * ((size_t)ldt * (dim - 1) + nbasm)), "coulomb_trafo: T is not a context allocation of the stated shape");
  SPX_CHECK(ctx, inplace || spx_owns(ctx, out, 16 * ((size_t)ldo * (n - 1) + dim)), "coulomb_trafo: out is not a context allocation of the stated shape");
  if (!inplace) {  // the output must not overlap the input
    const char *a = (const char *)M, *b = (const char *)out;
    const size_t la = 16 * ((size_t)ldm * (n - 1) + nbasm), lb = 16 * ((size_t)ldo * (n - 1) + dim);
    SPX_CHECK(ctx, a + la <= b || b + lb <= a, "coulomb_trafo: out overlaps M (pass out = M for in place)");
  }
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const long blk = inplace ? std::min<long>(n, 16384) : 0;
  const size_t hb = 16 * (size_t)dim * nbasm, yb = 16 * (size_t)dim * blk;
  void *ws = nullptr;
  int rc = spx_scratch(ctx, hb + yb, &ws);
  if (rc) return rc;
  double2 *H = (double2 *)ws, *Y = (double2 *)((char *)ws + hb);
  const long ne = (long)nbasm * dim;
  hipLaunchKernelGGL(ctrafo_h_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, nbasm,
                     dim, (const double2 *)T, ldt, H);
  SPX_LAUNCHED(ctx);
  const long step = inplace ? blk : 65535L * 64;  // grid.y of the GEMM stays below 65536
  for (long t0 = 0; t0 < n; t0 += step) {
    const long nt = std::min(step, n - t0);
    const double2 *X = (const double2 *)M + t0 * ldm;
    if (inplace) {
      rc = spx_zgemm(ctx, st, dim, (int)nt, nbasm, 1.0, H, dim, 0, X, ldm, 0, 0.0, Y, dim, 0, 0.0, 1);
      if (rc) return rc;
      hipLaunchKernelGGL(copy_rows_kernel, dim3((unsigned)((nt * dim + 255) / 256)), dim3(256), 0, st,
                         nt, dim, (const double2 *)Y, (double2 *)M + t0 * ldm, ldm);
      SPX_LAUNCHED(ctx);
    } else {
      rc = spx_zgemm(ctx, st, dim, (int)nt, nbasm, 1.0, H, dim, 0, X, ldm, 0, 0.0,
                     (double2 *)out + t0 * ldo, ldo, 0, 0.0, 1);
      if (rc) return rc;
    }
  }
  return SPEXHIP_OK;
}

int spexhip_ctrafo_olap_dev(spexhip_ctx *ctx, int nbasp, int ngptm, int dim, const void *T,
                            long ldt, const int *gptm, const double *cstep, const int *cext,
                            void *T1, long ldt1) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nbasp >= 0 && ngptm >= 0 && dim > 0 && ldt >= nbasp + ngptm &&
                     ldt1 >= nbasp + ngptm && (ngptm == 0 || (gptm && cstep && cext)),
            "ctrafo_olap: bad arguments");
  const int nbasm = nbasp + ngptm;
  SPX_CHECK(ctx, spx_owns(ctx, T, 16 * ((size_t)ldt * (dim - 1) + nbasm)), "ctrafo_olap: T is not a context allocation of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, T1, 16 * ((size_t)ldt1 * (dim - 1) + nbasm)), "ctrafo_olap: T1 is not a context allocation of the stated shape");
  SPX_CHECK(ctx, T != T1, "ctrafo_olap: T1 must not be T");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (nbasp > 0) {
    const long ne = (long)nbasp * dim;
    hipLaunchKernelGGL(copy_mt_rows_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st,
                       nbasp, dim, (const double2 *)T, ldt, (double2 *)T1, ldt1);
    SPX_LAUNCHED(ctx);
  }
  if (ngptm == 0) return SPEXHIP_OK;
  // every G_j - G_i must lie inside cstep
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  for (int i = 0; i < ngptm; i++)
    for (int x = 0; x < 3; x++) {
      lo[x] = std::min(lo[x], gptm[3 * i + x]);
      hi[x] = std::max(hi[x], gptm[3 * i + x]);
    }
  for (int x = 0; x < 3; x++)
    SPX_CHECK(ctx, cext[x] >= 0 && hi[x] - lo[x] <= cext[x],
              "ctrafo_olap: G differences reach %d > cext[%d] = %d", hi[x] - lo[x], x, cext[x]);
  const size_t gb = sizeof(int) * 3 * (size_t)ngptm;
  const size_t ncs = (size_t)(2 * cext[0] + 1) * (2 * cext[1] + 1) * (2 * cext[2] + 1);
  const size_t gbp = (gb + 255) / 256 * 256;
  void *ws = nullptr;
  int rc = spx_scratch(ctx, gbp + 16 * ncs, &ws);
  if (rc) return rc;
  int *Gd = (int *)ws;
  double2 *Cd = (double2 *)((char *)ws + gbp);
  SPX_HIP(ctx, hipMemcpyAsync(Gd, gptm, gb, hipMemcpyHostToDevice, st));
  SPX_HIP(ctx, hipMemcpyAsync(Cd, cstep, 16 * ncs, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(ctrafo_olap_kernel, dim3((ngptm + NT - 1) / NT, (dim + OC - 1) / OC), dim3(NT),
                     0, st, nbasp, ngptm, dim, (const int *)Gd, (const double2 *)Cd, cext[0],
                     cext[1], cext[2], (const double2 *)T, ldt, (double2 *)T1, ldt1);
  SPX_LAUNCHED(ctx);
  SPX_HIP(ctx, hipStreamSynchronize(st));  // gptm and cstep were copied from the caller's arrays
  return SPEXHIP_OK;
}

}  // extern "C"
