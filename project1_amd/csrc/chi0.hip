// chi0 pair sum as a Hermitian rank-k update on the CDNA4 matrix cores.
//
// Reference: susceptibility.f:1159-1251 (pair loop) with spectral_function,
// HILBERT NONE branch :1780-1788, i.e.
//     susceph(ij,w) += sum_t wghtSr(t,w) * cprod(t,i) * conjg(cprod(t,j)),
// one DDOT of length n per (i<=j, w) in the reference.  Here the same sum
// is C_w = L_w^T conj(M) with L_w(t,I) = wght(t,w) M(t,I), computed per
// 64x64 output tile of the upper triangle with MFMA:
//   * FP64 form: v_mfma_f64_16x16x4f64, 4 real products per complex step;
//   * split-FP32 form: L and M are split into hi = fp32(x), lo = fp32(x-hi);
//     hi*hi, hi*lo and lo*hi are issued as v_mfma_f32_16x16x4f32 into two
//     FP32 accumulators (hi*hi and the cross terms) that are flushed into
//     FP64 accumulators after every K-chunk of 16 transitions.
#include "common.hpp"

namespace {

constexpr int TM = 64;    // output tile edge (complex elements)
constexpr int KC = 16;    // transitions per pipeline stage
constexpr int NT = 256;   // threads per workgroup (4 waves, 2x2 layout)

__device__ inline void tri_index(int x, int &bi, int &bj) {
  int j = (int)((sqrtf(8.f * (float)x + 1.f) - 1.f) * 0.5f);
  while ((j + 1) * (j + 2) / 2 <= x) j++;
  while (j * (j + 1) / 2 > x) j--;
  bj = j;
  bi = x - j * (j + 1) / 2;
}

// Element e = tid + NT*r of a KC x TM stage; mapping keeps global loads
// coalesced for either operand layout.
template <bool COLMAJOR>
__device__ inline void stage_coords(int e, int &k, int &i) {
  if (COLMAJOR) {
    k = e % KC;
    i = e / KC;
  } else {
    k = e / TM;
    i = e % TM;
  }
}

template <bool COLMAJOR>
__device__ inline void load_stage(const double2 *__restrict__ M, long ldm, long n,
                                  int nb, long t0, int c0, int tid, double2 v[4]) {
#pragma unroll
  for (int r = 0; r < 4; r++) {
    int k, i;
    stage_coords<COLMAJOR>(tid + NT * r, k, i);
    long t = t0 + k;
    int c = c0 + i;
    double2 z = make_double2(0.0, 0.0);
    if (t < n && c < nb) z = COLMAJOR ? M[t + ldm * (long)c] : M[t * ldm + c];
    v[r] = z;
  }
}

// ---------------------------------------------------------------- FP64 --
template <bool COLMAJOR>
__global__ __launch_bounds__(NT) void chi0_f64_kernel(
    long n, int nb, const double2 *__restrict__ M, long ldm,
    const double *__restrict__ wght, int nw, double2 *__restrict__ chi,
    long ldc, double alpha, double beta) {
  __shared__ double Lr[2][KC][TM], Li[2][KC][TM], Rr[2][KC][TM], Ri[2][KC][TM];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1;
  int bi, bj;
  tri_index(blockIdx.x, bi, bj);
  const int w = blockIdx.y;
  const int I0 = bi * TM, J0 = bj * TM;
  const double *D = wght + (long)w * n;

  d4_t accr[2][2], acci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      accr[a][b] = d4_t{0, 0, 0, 0};
      acci[a][b] = d4_t{0, 0, 0, 0};
    }

  double2 va[4], vb[4];
  double dv[4];
  const long nst = (n + KC - 1) / KC;

  auto fetch = [&](long s) {
    long t0 = s * KC;
    load_stage<COLMAJOR>(M, ldm, n, nb, t0, I0, tid, va);
    load_stage<COLMAJOR>(M, ldm, n, nb, t0, J0, tid, vb);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int k, i;
      stage_coords<COLMAJOR>(tid + NT * r, k, i);
      long t = t0 + k;
      dv[r] = t < n ? D[t] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int k, i;
      stage_coords<COLMAJOR>(tid + NT * r, k, i);
      Lr[buf][k][i] = dv[r] * va[r].x;
      Li[buf][k][i] = dv[r] * va[r].y;
      Rr[buf][k][i] = vb[r].x;
      Ri[buf][k][i] = vb[r].y;
    }
  };

  fetch(0);
  store(0);
  __syncthreads();
  const int lr = lane & 15, lk = lane >> 4;
  for (long s = 0; s < nst; s++) {
    const int buf = s & 1;
    if (s + 1 < nst) fetch(s + 1);
#pragma unroll
    for (int kk = 0; kk < KC / 4; kk++) {
      const int k = kk * 4 + lk;
      double ar[2], ai[2], br[2], bim[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        ar[a] = Lr[buf][k][wr * 32 + a * 16 + lr];
        ai[a] = Li[buf][k][wr * 32 + a * 16 + lr];
        br[a] = Rr[buf][k][wc * 32 + a * 16 + lr];
        bim[a] = Ri[buf][k][wc * 32 + a * 16 + lr];
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          // (ar + i ai)(br - i bim) = ar br + ai bim + i(ai br - ar bim)
          accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], accr[a][b], 0, 0, 0);
          accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bim[b], accr[a][b], 0, 0, 0);
          acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], acci[a][b], 0, 0, 0);
          acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[a], bim[b], acci[a][b], 0, 0, 0);
        }
    }
    if (s + 1 < nst) store(buf ^ 1);
    __syncthreads();
  }

  // v_mfma_f64_16x16x4f64 result layout: lane holds column j = lane%16 and
  // rows i = lane/16 + 4*r (r = 0..3).
  double2 *C = chi + (long)w * ldc * nb;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int J = J0 + wc * 32 + b * 16 + lr;
      if (J >= nb) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int I = I0 + wr * 32 + a * 16 + lk + 4 * r;
        if (I >= nb) continue;
        double2 *p = C + I + ldc * (long)J;
        double2 o = make_double2(alpha * accr[a][b][r], alpha * acci[a][b][r]);
        if (beta != 0.0) {
          double2 old = *p;
          o.x += beta * old.x;
          o.y += beta * old.y;
        }
        *p = o;
      }
    }
}

// ------------------------------------------------------------ split FP32 --
__device__ inline void split2(double x, float &hi, float &lo) {
  hi = (float)x;
  lo = (float)(x - (double)hi);
}

template <bool COLMAJOR>
__global__ __launch_bounds__(NT) void chi0_split_kernel(
    long n, int nb, const double2 *__restrict__ M, long ldm,
    const double *__restrict__ wght, int nw, double2 *__restrict__ chi,
    long ldc, double alpha, double beta) {
  // hi/lo planes of L (weighted left operand) and R (right operand)
  __shared__ float Lrh[2][KC][TM], Lrl[2][KC][TM], Lih[2][KC][TM], Lil[2][KC][TM];
  __shared__ float Rrh[2][KC][TM], Rrl[2][KC][TM], Rih[2][KC][TM], Ril[2][KC][TM];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1;
  int bi, bj;
  tri_index(blockIdx.x, bi, bj);
  const int w = blockIdx.y;
  const int I0 = bi * TM, J0 = bj * TM;
  const double *D = wght + (long)w * n;

  d4_t accr[2][2], acci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      accr[a][b] = d4_t{0, 0, 0, 0};
      acci[a][b] = d4_t{0, 0, 0, 0};
    }

  double2 va[4], vb[4];
  double dv[4];
  const long nst = (n + KC - 1) / KC;

  auto fetch = [&](long s) {
    long t0 = s * KC;
    load_stage<COLMAJOR>(M, ldm, n, nb, t0, I0, tid, va);
    load_stage<COLMAJOR>(M, ldm, n, nb, t0, J0, tid, vb);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int k, i;
      stage_coords<COLMAJOR>(tid + NT * r, k, i);
      long t = t0 + k;
      dv[r] = t < n ? D[t] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int k, i;
      stage_coords<COLMAJOR>(tid + NT * r, k, i);
      float h, l;
      split2(dv[r] * va[r].x, h, l);
      Lrh[buf][k][i] = h;
      Lrl[buf][k][i] = l;
      split2(dv[r] * va[r].y, h, l);
      Lih[buf][k][i] = h;
      Lil[buf][k][i] = l;
      split2(vb[r].x, h, l);
      Rrh[buf][k][i] = h;
      Rrl[buf][k][i] = l;
      split2(vb[r].y, h, l);
      Rih[buf][k][i] = h;
      Ril[buf][k][i] = l;
    }
  };

  fetch(0);
  store(0);
  __syncthreads();
  const int lr = lane & 15, lk = lane >> 4;
  for (long s = 0; s < nst; s++) {
    const int buf = s & 1;
    if (s + 1 < nst) fetch(s + 1);
    f4_t hr[2][2], hi_[2][2], xr[2][2], xi[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        hr[a][b] = f4_t{0, 0, 0, 0};
        hi_[a][b] = f4_t{0, 0, 0, 0};
        xr[a][b] = f4_t{0, 0, 0, 0};
        xi[a][b] = f4_t{0, 0, 0, 0};
      }
#pragma unroll
    for (int kk = 0; kk < KC / 4; kk++) {
      const int k = kk * 4 + lk;
      float arh[2], arl[2], aih[2], ail[2], brh[2], brl[2], bih[2], bil[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int ia = wr * 32 + a * 16 + lr, jb = wc * 32 + a * 16 + lr;
        arh[a] = Lrh[buf][k][ia];
        arl[a] = Lrl[buf][k][ia];
        aih[a] = Lih[buf][k][ia];
        ail[a] = Lil[buf][k][ia];
        brh[a] = Rrh[buf][k][jb];
        brl[a] = Rrl[buf][k][jb];
        bih[a] = Rih[buf][k][jb];
        bil[a] = Ril[buf][k][jb];
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          // hi*hi
          hr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(arh[a], brh[b], hr[a][b], 0, 0, 0);
          hr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[a], bih[b], hr[a][b], 0, 0, 0);
          hi_[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[a], brh[b], hi_[a][b], 0, 0, 0);
          hi_[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-arh[a], bih[b], hi_[a][b], 0, 0, 0);
          // hi*lo + lo*hi
          xr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(arh[a], brl[b], xr[a][b], 0, 0, 0);
          xr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(arl[a], brh[b], xr[a][b], 0, 0, 0);
          xr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[a], bil[b], xr[a][b], 0, 0, 0);
          xr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ail[a], bih[b], xr[a][b], 0, 0, 0);
          xi[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[a], brl[b], xi[a][b], 0, 0, 0);
          xi[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ail[a], brh[b], xi[a][b], 0, 0, 0);
          xi[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-arh[a], bil[b], xi[a][b], 0, 0, 0);
          xi[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-arl[a], bih[b], xi[a][b], 0, 0, 0);
        }
    }
    // FP64 correction pass for this K-chunk
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          accr[a][b][r] += (double)hr[a][b][r] + (double)xr[a][b][r];
          acci[a][b][r] += (double)hi_[a][b][r] + (double)xi[a][b][r];
        }
    if (s + 1 < nst) store(buf ^ 1);
    __syncthreads();
  }

  // v_mfma_f32_16x16x4f32 result layout: lane holds column j = lane%16 and
  // rows i = 4*(lane/16) + r.
  double2 *C = chi + (long)w * ldc * nb;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int J = J0 + wc * 32 + b * 16 + lr;
      if (J >= nb) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int I = I0 + wr * 32 + a * 16 + 4 * lk + r;
        if (I >= nb) continue;
        double2 *p = C + I + ldc * (long)J;
        double2 o = make_double2(alpha * accr[a][b][r], alpha * acci[a][b][r]);
        if (beta != 0.0) {
          double2 old = *p;
          o.x += beta * old.x;
          o.y += beta * old.y;
        }
        *p = o;
      }
    }
}

// ------------------------------------------------------------ weights --
__global__ void weights_none_kernel(int nk, int nocc, int nunocc,
                                    const double *__restrict__ eo,
                                    const double *__restrict__ eu, int nkpt,
                                    int nsym1, const int *__restrict__ nkpts,
                                    int nw, const double *__restrict__ om,
                                    double *__restrict__ wght) {
  const long n = (long)nk * nocc * nunocc;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int b2 = (int)(t % nunocc);
  const long r = t / nunocc;
  const int b1 = (int)(r % nocc);
  const int k = (int)(r / nocc);
  const double d = eu[(long)k * nunocc + b2] - eo[(long)k * nocc + b1];
  const double kw = (double)nkpts[k];
  for (int f = 0; f < nw; f++) {
    // same operation order and roundings as the reference expression (no FMA)
#pragma clang fp contract(off)
    const double w = om[f];
    wght[t + n * f] = -2 * d / (w * w + d * d) / nkpt / nsym1 * kw;
  }
}

// ----------------------------------------------------------- packing --
__global__ void pack_upper_kernel(int nb, int nw, const double2 *__restrict__ X,
                                  long ldx, double2 *__restrict__ P) {
  const long npack = (long)nb * (nb + 1) / 2;
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npack * nw) return;
  const int w = (int)(g / npack);
  const long ij = g - (long)w * npack;
  long j = (long)((sqrt(8.0 * (double)ij + 1.0) - 1.0) * 0.5);
  while ((j + 1) * (j + 2) / 2 <= ij) j++;
  while (j * (j + 1) / 2 > ij) j--;
  const long i = ij - j * (j + 1) / 2;
  P[g] = X[(long)w * ldx * nb + i + ldx * j];
}

__global__ void unpack_kernel(int nb, int nw, const double2 *__restrict__ P,
                              double2 *__restrict__ X, long ldx) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nb * nb * nw) return;
  const int w = (int)(g / ((long)nb * nb));
  const long r = g - (long)w * nb * nb;
  const int j = (int)(r / nb), i = (int)(r % nb);
  const long npack = (long)nb * (nb + 1) / 2;
  double2 v;
  if (i <= j) {
    v = P[(long)w * npack + (long)j * (j + 1) / 2 + i];
  } else {
    v = P[(long)w * npack + (long)i * (i + 1) / 2 + j];
    v.y = -v.y;
  }
  X[(long)w * ldx * nb + i + ldx * (long)j] = v;
}

}  // namespace

int spx_chi0_launch(spexhip_ctx *ctx, hipStream_t st, long n, int nb,
                    const void *M, long ldm, int colmajor, int nw,
                    const void *wght, int precision, double alpha, double beta,
                    void *chi, int ldchi) {
  SPX_CHECK(ctx, n >= 0 && nb > 0 && nw > 0, "chi0: bad sizes n=%ld nb=%d nw=%d", n, nb, nw);
  SPX_CHECK(ctx, ldchi >= nb, "chi0: ldchi=%d < nb=%d", ldchi, nb);
  SPX_CHECK(ctx, colmajor ? ldm >= n : ldm >= nb, "chi0: ldm=%ld too small", ldm);
  SPX_CHECK(ctx, precision == SPEXHIP_PREC_FP64 || precision == SPEXHIP_PREC_SPLIT_FP32 ||
                     precision == SPEXHIP_PREC_OZAKI_INT8,
            "chi0: unknown precision %d", precision);
  if (precision == SPEXHIP_PREC_OZAKI_INT8)
    return spx_chi0_ozaki(ctx, st, n, nb, M, ldm, colmajor, nw, wght, alpha, beta, chi, ldchi,
                          ctx->ozaki_ws_limit);
  const int nbt = (nb + TM - 1) / TM;
  dim3 grid(nbt * (nbt + 1) / 2, nw), block(NT);
  if (n == 0) {
    // empty transition set: chi = beta*chi
    alpha = 0.0;
  }
  const double2 *Mp = (const double2 *)M;
  const double *W = (const double *)wght;
  double2 *C = (double2 *)chi;
  if (precision == SPEXHIP_PREC_FP64) {
    if (colmajor)
      hipLaunchKernelGGL(chi0_f64_kernel<true>, grid, block, 0, st, n, nb, Mp, ldm, W, nw, C, (long)ldchi, alpha, beta);
    else
      hipLaunchKernelGGL(chi0_f64_kernel<false>, grid, block, 0, st, n, nb, Mp, ldm, W, nw, C, (long)ldchi, alpha, beta);
  } else {
    if (colmajor)
      hipLaunchKernelGGL(chi0_split_kernel<true>, grid, block, 0, st, n, nb, Mp, ldm, W, nw, C, (long)ldchi, alpha, beta);
    else
      hipLaunchKernelGGL(chi0_split_kernel<false>, grid, block, 0, st, n, nb, Mp, ldm, W, nw, C, (long)ldchi, alpha, beta);
  }
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

extern "C" {

int spexhip_weights_none_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc,
                             const void *e_occ, const void *e_unocc, int nkpt,
                             int nsym1, const void *nkpts, int nw,
                             const void *omega, void *wght) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nk >= 0 && nocc >= 0 && nunocc >= 0 && nw > 0 && nkpt > 0 && nsym1 > 0,
            "weights: bad sizes");
  const long n = (long)nk * nocc * nunocc;
  if (n == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, spx_owns(ctx, wght, sizeof(double) * n * nw), "weights: output not a context allocation of n*nw doubles");
  SPX_CHECK(ctx, spx_owns(ctx, e_occ, sizeof(double) * nk * nocc) &&
                 spx_owns(ctx, e_unocc, sizeof(double) * nk * nunocc) &&
                 spx_owns(ctx, nkpts, sizeof(int) * nk) &&
                 spx_owns(ctx, omega, sizeof(double) * nw),
            "weights: input not a context allocation of the stated size");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const int bs = 256;
  hipLaunchKernelGGL(weights_none_kernel, dim3((n + bs - 1) / bs), dim3(bs), 0, ctx->stream,
                     nk, nocc, nunocc, (const double *)e_occ, (const double *)e_unocc, nkpt,
                     nsym1, (const int *)nkpts, nw, (const double *)omega, (double *)wght);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_chi0_dev(spexhip_ctx *ctx, long n, int nb, const void *M, long ldm,
                     int colmajor, int nw, const void *wght, int precision,
                     double alpha, double beta, void *chi, int ldchi) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && n >= 0 && ldchi >= nb, "chi0: bad sizes");
  if (n > 0) {
    const size_t mbytes = colmajor ? 16 * ((size_t)ldm * (nb - 1) + n)
                                   : 16 * ((size_t)ldm * (n - 1) + nb);
    SPX_CHECK(ctx, spx_owns(ctx, M, mbytes), "chi0: M is not a context allocation of the stated shape");
    SPX_CHECK(ctx, spx_owns(ctx, wght, sizeof(double) * n * nw), "chi0: wght is not a context allocation of n*nw doubles");
  }
  SPX_CHECK(ctx, spx_owns(ctx, chi, 16 * ((size_t)ldchi * nb * (nw - 1) + (size_t)ldchi * (nb - 1) + nb)),
            "chi0: chi is not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  return spx_chi0_launch(ctx, ctx->stream, n, nb, M, ldm, colmajor, nw, wght, precision,
                         alpha, beta, chi, ldchi);
}

int spexhip_pack_upper_d2h(spexhip_ctx *ctx, int nb, int nw, const void *X, int ldx,
                           double *packed, int accumulate) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && ldx >= nb && packed, "pack: bad arguments");
  SPX_CHECK(ctx, spx_owns(ctx, X, 16 * ((size_t)ldx * nb * nw)), "pack: X is not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const long npack = (long)nb * (nb + 1) / 2;
  const size_t bytes = 16 * (size_t)npack * nw;
  void *P;
  int rc = spx_scratch(ctx, bytes, &P);
  if (rc) return rc;
  const int bs = 256;
  hipLaunchKernelGGL(pack_upper_kernel, dim3((npack * nw + bs - 1) / bs), dim3(bs), 0,
                     ctx->stream, nb, nw, (const double2 *)X, (long)ldx, (double2 *)P);
  SPX_LAUNCHED(ctx);
  if (!accumulate) {
    SPX_HIP(ctx, hipMemcpyAsync(packed, P, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SPEXHIP_OK;
  }
  double *tmp = (double *)malloc(bytes);
  SPX_CHECK(ctx, tmp, "pack: host allocation failed");
  hipError_t e = hipMemcpyAsync(tmp, P, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free(tmp);
    return spx_seterr(ctx, SPEXHIP_E_DEVICE, "pack: copy failed: %s", hipGetErrorString(e));
  }
  for (size_t q = 0; q < bytes / 8; q++) packed[q] += tmp[q];
  free(tmp);
  return SPEXHIP_OK;
}

int spexhip_unpack_h2d(spexhip_ctx *ctx, int nb, int nw, const double *packed, void *X,
                       int ldx) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && ldx >= nb && packed, "unpack: bad arguments");
  SPX_CHECK(ctx, spx_owns(ctx, X, 16 * ((size_t)ldx * nb * nw)), "unpack: X is not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const long npack = (long)nb * (nb + 1) / 2;
  const size_t bytes = 16 * (size_t)npack * nw;
  void *P;
  int rc = spx_scratch(ctx, bytes, &P);
  if (rc) return rc;
  SPX_HIP(ctx, hipMemcpyAsync(P, packed, bytes, hipMemcpyHostToDevice, ctx->stream));
  const int bs = 256;
  const long tot = (long)nb * nb * nw;
  hipLaunchKernelGGL(unpack_kernel, dim3((tot + bs - 1) / bs), dim3(bs), 0, ctx->stream, nb, nw,
                     (const double2 *)P, (double2 *)X, (long)ldx);
  SPX_LAUNCHED(ctx);
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SPEXHIP_OK;
}

int spexhip_susceptibility_none(spexhip_ctx *ctx, long n, int nb, const double *cprod,
                                long ldc, int nw, const double *wght, int precision,
                                double *susceph) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, n >= 0 && nb > 0 && nw > 0 && ldc >= n && susceph, "susceptibility: bad arguments");
  if (n == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, cprod && wght, "susceptibility: null input");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t mb = 16 * (size_t)n * nb, wb = 8 * (size_t)n * nw,
               cb = 16 * (size_t)nb * nb * nw;
  void *dM = nullptr, *dW = nullptr, *dC = nullptr;
  int rc = spx_alloc(ctx, mb, &dM);
  if (!rc) rc = spx_alloc(ctx, wb, &dW);
  if (!rc) rc = spx_alloc(ctx, cb, &dC);
  if (!rc) {
    hipError_t e = hipMemcpy2DAsync(dM, 16 * n, cprod, 16 * ldc, 16 * n, nb,
                                    hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dW, wght, wb, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "susceptibility: upload failed: %s", hipGetErrorString(e));
  }
  if (!rc)
    rc = spx_chi0_launch(ctx, ctx->stream, n, nb, dM, n, 1, nw, dW, precision, 1.0, 0.0, dC, nb);
  if (!rc) rc = spexhip_pack_upper_d2h(ctx, nb, nw, dC, nb, susceph, 1);
  (void)hipStreamSynchronize(ctx->stream);
  if (dM) spx_free(ctx, dM);
  if (dW) spx_free(ctx, dW);
  if (dC) spx_free(ctx, dC);
  return rc;
}

}  // extern "C"
