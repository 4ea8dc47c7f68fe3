// HILBERT frequency weights and the real-frequency (complex-weighted) chi0.
//
// Reference: with a Hilbert mesh (HILBERT, the default for metals and for
// real frequencies) the pair loop does not use one weight per transition.
// Per pair it forms the spectral function on the mesh,
//     S(s) = sum_{t: wtetra(t,s) > wghtthr} re(t) wtetra(t,s)
// (susceptibility.f:1066-1077 index, :1793-1815 spectral_function), and
// then the Hilbert sums sus(w) = sum_s S(s) wghtSr(s,w) (:1838-1841), with
// wtetra from the tetrahedron method and wghtS from getwghtS
// (hilbert.f:193-300).  Summing over s first gives one weight per transition,
//     D(t,w) = sum_{s: wtetra(t,s) > wghtthr} wtetra(t,s) wghtS(s,w),
// and the pair sum becomes the same rank-k update as HILBERT NONE.  That
// fold is an n x nfrqs by nfrqs x nfrq product with a tiny inner size:
// it streams wtetra once from HBM (the bound), with the wghtS block of the
// workgroup in LDS and one transition per thread, summing s in increasing
// order (deterministic).
//
// Real frequencies (suscepc, susceptibility.f:1231-1236) carry complex
// weights Dr + i Di and fill the whole matrix:
//     chi(I,J,w) = sum_t (Dr + i Di)(t,w) M(t,I) conj(M(t,J)).
// With A_x(I,J) = sum_t D_x M(t,I) conj(M(t,J)) (Hermitian) that is
// chi = A_r + i A_i: two rank-k updates on the existing kernels, then one
// pass that writes both triangles.
#include "common.hpp"
#include "internal.hpp"

namespace {

constexpr int HB_NT = 256;
constexpr int HB_MAXF = 32;  // frequencies per fold call
constexpr int HB_SBLK = 64;  // mesh points per LDS stage

__global__ __launch_bounds__(HB_NT) void weights_fold_kernel(long n, int nfrqs,
                                                             const double *__restrict__ wt,
                                                             long ldwt, double thr, int nfrq,
                                                             const double *__restrict__ H,
                                                             int ldh, double *__restrict__ D,
                                                             long ldd) {
  __shared__ double Hs[HB_SBLK][HB_MAXF];
  const long t = (long)blockIdx.x * HB_NT + threadIdx.x;
  double acc[HB_MAXF];
#pragma unroll
  for (int w = 0; w < HB_MAXF; w++) acc[w] = 0.0;
  for (int s0 = 0; s0 < nfrqs; s0 += HB_SBLK) {
    const int ns = nfrqs - s0 < HB_SBLK ? nfrqs - s0 : HB_SBLK;
    __syncthreads();  // the previous stage has been read
    for (int e = threadIdx.x; e < HB_SBLK * HB_MAXF; e += HB_NT) {
      const int s = e / HB_MAXF, w = e % HB_MAXF;
      Hs[s][w] = (s < ns && w < nfrq) ? H[(s0 + s) + (long)ldh * w] : 0.0;
    }
    __syncthreads();
    if (t < n) {
      for (int s = 0; s < ns; s++) {
        const double v = wt[t + ldwt * (long)(s0 + s)];
        if (v > thr) {  // the reference's index test (strict, signed)
#pragma unroll
          for (int w = 0; w < HB_MAXF; w++) acc[w] += v * Hs[s][w];
        }
      }
    }
  }
  if (t >= n) return;
#pragma unroll
  for (int w = 0; w < HB_MAXF; w++)
    if (w < nfrq) D[t + ldd * w] = acc[w];
}

// X holds A_r and T holds A_i (upper triangles valid, column-major, frequency
// stride ld*nb).  Writes X(I,J) = A_r(I,J) + i A_i(I,J) for I <= J and
// X(J,I) = conj(A_r(I,J)) + i conj(A_i(I,J)); the diagonal takes the real
// parts (A_x is Hermitian).  One thread per (I <= J) pair.
__global__ void cplx_combine_kernel(int nb, double2 *__restrict__ X, const double2 *__restrict__ T,
                                    long ld) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long npack = (long)nb * (nb + 1) / 2;
  if (g >= npack) return;
  // g = J(J+1)/2 + I, I <= J
  int J = (int)((sqrt(8.0 * (double)g + 1.0) - 1.0) / 2.0);
  while ((long)J * (J + 1) / 2 > g) J--;
  while ((long)(J + 1) * (J + 2) / 2 <= g) J++;
  const int I = (int)(g - (long)J * (J + 1) / 2);
  const long off = ld * nb * (long)blockIdx.y;
  double2 *Xw = X + off;
  const double2 *Tw = T + off;
  const double2 ar = Xw[I + ld * (long)J], ai = Tw[I + ld * (long)J];
  if (I == J) {
    Xw[I + ld * (long)I] = make_double2(ar.x, ai.x);
    return;
  }
  Xw[I + ld * (long)J] = make_double2(ar.x - ai.y, ar.y + ai.x);   // ar + i ai
  Xw[J + ld * (long)I] = make_double2(ar.x + ai.y, -ar.y + ai.x);  // conj(ar) + i conj(ai)
}

}  // namespace

extern "C" {

int spexhip_weights_hilbert_dev(spexhip_ctx *ctx, long n, int nfrqs, const void *wtetra, long ldwt,
                                double wghtthr, int nfrq, const void *wghtS, int ldws, void *D,
                                long ldd) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, n >= 0 && nfrqs >= 1 && nfrq >= 1 && ldwt >= n && ldws >= nfrqs && ldd >= n,
            "weights_hilbert: bad sizes");
  SPX_CHECK(ctx, nfrq <= HB_MAXF, "weights_hilbert: at most %d frequencies per call (got %d)",
            HB_MAXF, nfrq);
  if (n == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, spx_owns(ctx, wtetra, 8 * ((size_t)ldwt * (nfrqs - 1) + n)),
            "weights_hilbert: wtetra is not a context allocation of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, wghtS, 8 * ((size_t)ldws * (nfrq - 1) + nfrqs)),
            "weights_hilbert: wghtS is not a context allocation of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, D, 8 * ((size_t)ldd * (nfrq - 1) + n)),
            "weights_hilbert: D is not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(weights_fold_kernel, dim3((unsigned)((n + HB_NT - 1) / HB_NT)), dim3(HB_NT),
                     0, ctx->stream, n, nfrqs, (const double *)wtetra, ldwt, wghtthr, nfrq,
                     (const double *)wghtS, ldws, (double *)D, ldd);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_chi0_cplx_dev(spexhip_ctx *ctx, long n, int nb, const void *M, long ldm, int colmajor,
                          int nw, const void *Dr, const void *Di, int precision, double alpha,
                          void *chi, int ldchi) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && n >= 0 && ldchi >= nb, "chi0_cplx: bad sizes");
  SPX_CHECK(ctx, precision == SPEXHIP_PREC_FP64 || precision == SPEXHIP_PREC_SPLIT_FP32,
            "chi0_cplx: FP64 or split-FP32 only (real-frequency weights change sign)");
  if (n > 0) {
    const size_t mbytes = colmajor ? 16 * ((size_t)ldm * (nb - 1) + n)
                                   : 16 * ((size_t)ldm * (n - 1) + nb);
    SPX_CHECK(ctx, spx_owns(ctx, M, mbytes), "chi0_cplx: M is not a context allocation of the stated shape");
    SPX_CHECK(ctx, spx_owns(ctx, Dr, sizeof(double) * n * nw) && spx_owns(ctx, Di, sizeof(double) * n * nw),
              "chi0_cplx: Dr, Di are not context allocations of n*nw doubles");
  }
  const size_t cbytes = 16 * (size_t)ldchi * nb * nw;
  SPX_CHECK(ctx, spx_owns(ctx, chi, cbytes), "chi0_cplx: chi is not a context allocation of ldchi*nb*nw");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *T;
  int rc = spx_scratch(ctx, cbytes, &T);
  if (rc) return rc;
  rc = spx_chi0_launch(ctx, ctx->stream, n, nb, M, ldm, colmajor, nw, Dr, precision, alpha, 0.0,
                       chi, ldchi);
  if (!rc)
    rc = spx_chi0_launch(ctx, ctx->stream, n, nb, M, ldm, colmajor, nw, Di, precision, alpha, 0.0,
                         T, ldchi);
  if (rc) return rc;
  const long npack = (long)nb * (nb + 1) / 2;
  hipLaunchKernelGGL(cplx_combine_kernel, dim3((unsigned)((npack + 255) / 256), nw), dim3(256), 0,
                     ctx->stream, nb, (double2 *)chi, (const double2 *)T, (long)ldchi);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

}  // extern "C"
