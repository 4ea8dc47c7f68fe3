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

namespace {

// MULTDIFF ON (multdiff_ = 1) inputs of the host drop-in: the Hilbert mesh
// and the band energies of the packet's transitions.
struct MultDiff {
  const double *frqs;
  int nk, nocc, nunocc;
  const double *e_occ, *e_unocc;
};

// The HILBERT pair sum of one memory packet on host arrays; md non-null adds
// MULTDIFF's first row and column (susceptibility.f:912-925, 1193).
int susc_hilbert(spexhip_ctx *ctx, long n, int nb, const double *cprod, long ldc, int nfrqs,
                 const double *wtetra, long ldwt, double wghtthr, int nfrqh, int nfrqc,
                 const double *wghtSr, const double *wghtSi, int precision, double *susceph,
                 double *suscepc, const MultDiff *md) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, n >= 0 && nb > 0 && nfrqs >= 1 && nfrqh >= 0 && nfrqc >= 0 && ldc >= n &&
                     ldwt >= n,
            "susceptibility_hilbert: bad arguments");
  SPX_CHECK(ctx, (nfrqh == 0 || susceph) && (nfrqc == 0 || (suscepc && wghtSi)),
            "susceptibility_hilbert: missing output or wghtSi");
  if (n == 0 || nfrqh + nfrqc == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, cprod && wtetra && wghtSr, "susceptibility_hilbert: null input");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const int nf = nfrqh + nfrqc;
  const int FB = HB_MAXF;  // frequencies per fold / pair-sum batch
  const size_t mb = 16 * (size_t)n * nb, tb = 8 * (size_t)n * nfrqs, sb = 8 * (size_t)nfrqs * nf,
               db = 8 * (size_t)n * FB, cb = 16 * (size_t)nb * nb * FB;
  void *dM = nullptr, *dT = nullptr, *dSr = nullptr, *dSi = nullptr, *dDr = nullptr,
       *dDi = nullptr, *dC = nullptr;
  int rc = spx_alloc(ctx, mb, &dM);
  if (!rc) rc = spx_alloc(ctx, tb, &dT);
  if (!rc) rc = spx_alloc(ctx, sb, &dSr);
  if (!rc && nfrqc) rc = spx_alloc(ctx, sb, &dSi);
  if (!rc) rc = spx_alloc(ctx, db, &dDr);
  if (!rc && nfrqc) rc = spx_alloc(ctx, db, &dDi);
  if (!rc) rc = spx_alloc(ctx, cb, &dC);
  // MULTDIFF: wghtSr and wghtSi divided by frqs^l (l = 1, 2) where frqs > 0,
  // their folds, and the energies (the plain weights D0 are dDr / dDi)
  void *dS1 = nullptr, *dS2 = nullptr, *dS1i = nullptr, *dS2i = nullptr, *dD1 = nullptr,
       *dD2 = nullptr, *dD1i = nullptr, *dD2i = nullptr, *dEo = nullptr, *dEu = nullptr;
  if (md) {
    const size_t eb = 8 * (size_t)md->nk * md->nocc, ub = 8 * (size_t)md->nk * md->nunocc;
    if (!rc) rc = spx_alloc(ctx, sb, &dS1);
    if (!rc) rc = spx_alloc(ctx, sb, &dS2);
    if (!rc && nfrqc) rc = spx_alloc(ctx, sb, &dS1i);
    if (!rc && nfrqc) rc = spx_alloc(ctx, sb, &dS2i);
    if (!rc) rc = spx_alloc(ctx, db, &dD1);
    if (!rc) rc = spx_alloc(ctx, db, &dD2);
    if (!rc && nfrqc) rc = spx_alloc(ctx, db, &dD1i);
    if (!rc && nfrqc) rc = spx_alloc(ctx, db, &dD2i);
    if (!rc) rc = spx_alloc(ctx, eb, &dEo);
    if (!rc) rc = spx_alloc(ctx, ub, &dEu);
    std::vector<double> tab((size_t)nfrqs * nf);
    auto put = [&](const double *S, int l, void *dst) {
      for (int f = 0; f < nf; f++)
        for (int i = 0; i < nfrqs; i++) {
          const size_t ix = i + (size_t)nfrqs * f;
          const double r = md->frqs[i] > 0 ? md->frqs[i] : 1.0;
          tab[ix] = l == 1 ? S[ix] / r : S[ix] / (r * r);
        }
      return hipMemcpy(dst, tab.data(), sb, hipMemcpyHostToDevice);
    };
    if (!rc) {
      hipError_t e = put(wghtSr, 1, dS1);
      if (e == hipSuccess) e = put(wghtSr, 2, dS2);
      if (e == hipSuccess && nfrqc) e = put(wghtSi, 1, dS1i);
      if (e == hipSuccess && nfrqc) e = put(wghtSi, 2, dS2i);
      if (e == hipSuccess) e = hipMemcpy(dEo, md->e_occ, eb, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(dEu, md->e_unocc, ub, hipMemcpyHostToDevice);
      if (e != hipSuccess)
        rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "susceptibility_hilbert: upload failed: %s",
                        hipGetErrorString(e));
    }
  }
  std::vector<double> full;
  if (!rc) {
    hipError_t e = hipMemcpy2DAsync(dM, 16 * n, cprod, 16 * ldc, 16 * n, nb,
                                    hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpy2DAsync(dT, 8 * n, wtetra, 8 * ldwt, 8 * n, nfrqs, hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dSr, wghtSr, sb, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && nfrqc)
      e = hipMemcpyAsync(dSi, wghtSi, sb, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "susceptibility_hilbert: upload failed: %s",
                      hipGetErrorString(e));
  }
  // imaginary frequencies: fold, Hermitian pair sum, packed accumulate
  for (int f0 = 0; !rc && f0 < nfrqh; f0 += FB) {
    const int k = nfrqh - f0 < FB ? nfrqh - f0 : FB;
    rc = spexhip_weights_hilbert_dev(ctx, n, nfrqs, dT, n, wghtthr, k,
                                     (const double *)dSr + (size_t)nfrqs * f0, nfrqs, dDr, n);
    if (!rc) rc = spx_chi0_launch(ctx, ctx->stream, n, nb, dM, n, 1, k, dDr, precision, 1.0, 0.0, dC, nb);
    if (!rc && md) {
      const size_t off = (size_t)nfrqs * f0;
      rc = spexhip_weights_hilbert_dev(ctx, n, nfrqs, dT, n, wghtthr, k,
                                       (const double *)dS1 + off, nfrqs, dD1, n);
      if (!rc)
        rc = spexhip_weights_hilbert_dev(ctx, n, nfrqs, dT, n, wghtthr, k,
                                         (const double *)dS2 + off, nfrqs, dD2, n);
      if (!rc)
        rc = spexhip_multdiff_dev(ctx, md->nk, md->nocc, md->nunocc, dEo, dEu, nb, dM, n, 1, k,
                                  dDr, dD1, dD2, n, 1.0, dC, nb);
    }
    if (!rc)
      rc = spexhip_pack_upper_d2h(ctx, nb, k, dC, nb, susceph + 2 * (size_t)nb * (nb + 1) / 2 * f0, 1);
  }
  // real frequencies: complex weights, whole matrix
  for (int f0 = 0; !rc && f0 < nfrqc; f0 += FB) {
    const int k = nfrqc - f0 < FB ? nfrqc - f0 : FB;
    const size_t col = (size_t)nfrqs * (nfrqh + f0);
    rc = spexhip_weights_hilbert_dev(ctx, n, nfrqs, dT, n, wghtthr, k, (const double *)dSr + col,
                                     nfrqs, dDr, n);
    if (!rc)
      rc = spexhip_weights_hilbert_dev(ctx, n, nfrqs, dT, n, wghtthr, k, (const double *)dSi + col,
                                       nfrqs, dDi, n);
    if (!rc)
      rc = spexhip_chi0_cplx_dev(ctx, n, nb, dM, n, 1, k, dDr, dDi, precision, 1.0, dC, nb);
    if (!rc && md) {
      const struct {
        void *S, *D;
      } fold[4] = {{dS1, dD1}, {dS1i, dD1i}, {dS2, dD2}, {dS2i, dD2i}};
      for (int j = 0; !rc && j < 4; j++)
        rc = spexhip_weights_hilbert_dev(ctx, n, nfrqs, dT, n, wghtthr, k,
                                         (const double *)fold[j].S + col, nfrqs, fold[j].D, n);
      if (!rc)
        rc = spexhip_multdiff_cplx_dev(ctx, md->nk, md->nocc, md->nunocc, dEo, dEu, nb, dM, n, 1,
                                       k, dDr, dDi, dD1, dD1i, dD2, dD2i, n, 1.0, dC, nb);
    }
    if (!rc) {
      full.resize(2 * (size_t)nb * nb * k);
      hipError_t e = hipMemcpyAsync(full.data(), dC, 16 * (size_t)nb * nb * k,
                                    hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess)
        rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "susceptibility_hilbert: download failed: %s",
                        hipGetErrorString(e));
      double *dst = suscepc + 2 * (size_t)nb * nb * f0;
      for (size_t i = 0; !rc && i < full.size(); i++) dst[i] += full[i];
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  for (void *p : {dM, dT, dSr, dSi, dDr, dDi, dC, dS1, dS2, dS1i, dS2i, dD1, dD2, dD1i,
                  dD2i, dEo, dEu})
    if (p) spx_free(ctx, p);
  return rc;
}

}  // namespace

extern "C" {

// Host drop-in for the HILBERT pair sum of one memory packet: the index
// and threshold (susceptibility.f:1066-1077), the pair loop with
// spectral_function (:1159-1251, :1793-1849) and its accumulation into
// susceph (imaginary frequencies, packed upper) and suscepc (real
// frequencies, whole matrix, :1231-1236).  All arrays are host arrays in the
// reference layouts; susceph and suscepc are accumulated.
int spexhip_susceptibility_hilbert(spexhip_ctx *ctx, long n, int nb, const double *cprod,
                                   long ldc, int nfrqs, const double *wtetra, long ldwt,
                                   double wghtthr, int nfrqh, int nfrqc, const double *wghtSr,
                                   const double *wghtSi, int precision, double *susceph,
                                   double *suscepc) {
  return susc_hilbert(ctx, n, nb, cprod, ldc, nfrqs, wtetra, ldwt, wghtthr, nfrqh, nfrqc, wghtSr,
                      wghtSi, precision, susceph, suscepc, nullptr);
}

// The same with MULTDIFF ON (multdiff_ = 1, susceptibility.f:588-592,
// 912-925, 1193, 1817-1824): cprod is the unscaled packet, frqs the Hilbert
// mesh, e_occ [nk][nocc] at k and e_unocc [nk][nunocc] at k+q of its
// n = nk*nocc*nunocc transitions.
int spexhip_susceptibility_hilbert_multdiff(spexhip_ctx *ctx, long n, int nb, const double *cprod,
                                            long ldc, int nfrqs, const double *wtetra, long ldwt,
                                            double wghtthr, int nfrqh, int nfrqc,
                                            const double *wghtSr, const double *wghtSi,
                                            int precision, const double *frqs, int nk, int nocc,
                                            int nunocc, const double *e_occ,
                                            const double *e_unocc, double *susceph,
                                            double *suscepc) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, frqs && e_occ && e_unocc && nk > 0 && nocc > 0 && nunocc > 0 &&
                     (long)nk * nocc * nunocc == n,
            "susceptibility_hilbert_multdiff: frqs, energies or n = nk*nocc*nunocc missing");
  const MultDiff md{frqs, nk, nocc, nunocc, e_occ, e_unocc};
  return susc_hilbert(ctx, n, nb, cprod, ldc, nfrqs, wtetra, ldwt, wghtthr, nfrqh, nfrqc, wghtSr,
                      wghtSi, precision, susceph, suscepc, &md);
}

}  // extern "C"
