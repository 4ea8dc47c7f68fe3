// chi0 pair sum as a Hermitian rank-k update on the CDNA4 matrix cores.
//
// Reference: susceptibility.f:1159-1251 (pair loop) with spectral_function,
// HILBERT NONE branch :1780-1788, i.e.
//     susceph(ij,w) += sum_t wghtSr(t,w) * cprod(t,i) * conjg(cprod(t,j)),
// one DDOT of length n per (i<=j, w) in the reference.  Here the same sum
// is C_w = L_w^T conj(M) with L_w(t,I) = wght(t,w) M(t,I), computed per
// 64x64 output tile of the upper triangle with MFMA:
//   * FP64 form: v_mfma_f64_16x16x4f64, 4 real products per complex step;
//   * split-FP32 form: both sides are sqrt|D| M on a grid relative to their
//     column maximum, split into an integer part h (<= 10 bits, so h*h and
//     its K-chunk sums are exact in FP32) and an FP32 rest l; h*h goes into
//     one FP32 accumulator, h*l + l*fp32(h+l) into another, both flushed
//     into FP64 accumulators after every K-chunk of 16 transitions
//     (chi0_tile.hpp).  A scale pass (colscale.hpp) reads M once first.
#include <algorithm>
#include <type_traits>

#include "colscale.hpp"
#include "common.hpp"

#include "chi0_tile.hpp"

namespace {

using chi0t::Lds;
using chi0t::Problem;

template <int TM, int KC, bool COLMAJOR, bool SPLIT>
__global__ __launch_bounds__(chi0t::NT) void chi0_kernel(Problem P) {
  __shared__ Lds<TM, KC, SPLIT> s;
  int bi, bj;
  chi0t::tri_index(blockIdx.x, bi, bj);
  chi0t::tile<TM, KC, COLMAJOR, SPLIT>(P, bi, bj, blockIdx.y, s);
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

template <int TM, int KC>
static void launch_cfg(hipStream_t st, const Problem &P, int colmajor, int split) {
  const int nbt = (P.nb + TM - 1) / TM;
  dim3 grid(nbt * (nbt + 1) / 2, P.nw), block(chi0t::NT);
  if (split) {
    if (colmajor) hipLaunchKernelGGL((chi0_kernel<TM, KC, true, true>), grid, block, 0, st, P);
    else hipLaunchKernelGGL((chi0_kernel<TM, KC, false, true>), grid, block, 0, st, P);
  } else {
    if (colmajor) hipLaunchKernelGGL((chi0_kernel<TM, KC, true, false>), grid, block, 0, st, P);
    else hipLaunchKernelGGL((chi0_kernel<TM, KC, false, false>), grid, block, 0, st, P);
  }
}

// ------------------------------------------------------- column scales --
namespace colscale {

// scale[w][I] via atomicMax on the bit pattern (non-negative doubles order
// like unsigned integers); the caller zeroes scale first.  One thread per
// column I and tchunk transitions.  The weights of a stage of 64
// transitions are the same for every thread: the workgroup stages them in
// LDS (NWP = nw rounded up to 4, rows w >= nw read as 0) and each thread
// reads them back as broadcasts, eight transitions at a time.
template <bool COLMAJOR, int NWP, bool FROM_D>
__global__ __launch_bounds__(256) void col_scale_kernel(long n, int nb,
                                                        const double2 *__restrict__ M, long ldm,
                                                        int nw, const double *__restrict__ W,
                                                        long tchunk,
                                                        unsigned long long *__restrict__ scale) {
  constexpr int TS = 64;
  __shared__ double ss[NWP][TS];
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = I < nb;
  const long t0 = (long)blockIdx.y * tchunk;
  const long t1 = t0 + tchunk < n ? t0 + tchunk : n;
  double m[NWP];
#pragma unroll
  for (int w = 0; w < NWP; w++) m[w] = 0.0;
  for (long tb = t0; tb < t1; tb += TS) {
    const int len = (int)(t1 - tb < TS ? t1 - tb : TS);
    __syncthreads();  // the previous stage has been read
    for (int e = threadIdx.x; e < NWP * TS; e += blockDim.x) {
      const int w = e / TS, j = e % TS;
      double v = 0.0;
      if (w < nw && j < len) {
        v = W[(long)w * n + tb + j];
        if (FROM_D) v = sqrt(fabs(v));
      }
      ss[w][j] = v;
    }
    __syncthreads();
    if (live) {
      int j0 = 0;
      for (; j0 + 8 <= len; j0 += 8) {  // no branches: eight loads in flight
        double2 v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const long t = tb + j0 + j;
          v[j] = COLMAJOR ? M[t + ldm * (long)I] : M[t * ldm + I];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the eight loads ahead of their uses
        double a[8];
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = fmax(fabs(v[j].x), fabs(v[j].y));
#pragma unroll
        for (int w = 0; w < NWP; w++) {
          double x = a[0] * ss[w][j0];
#pragma unroll
          for (int j = 1; j < 8; j++) x = fmax(x, a[j] * ss[w][j0 + j]);
          m[w] = fmax(m[w], x);
        }
      }
      for (; j0 < len; j0++) {
        const long t = tb + j0;
        const double2 v = COLMAJOR ? M[t + ldm * (long)I] : M[t * ldm + I];
        const double a = fmax(fabs(v.x), fabs(v.y));
#pragma unroll
        for (int w = 0; w < NWP; w++) m[w] = fmax(m[w], a * ss[w][j0]);
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int w = 0; w < NWP; w++)
    if (w < nw) atomicMax(scale + (long)w * nb + I, (unsigned long long)__double_as_longlong(m[w]));
}

__global__ void zero_u64_kernel(unsigned long long *__restrict__ p, long count) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < count) p[g] = 0ull;
}

}  // namespace colscale

size_t spx_split_ws_bytes(int nb, int nw) { return 8 * (size_t)nb * nw; }

// Column scales of spexhip_chi0_noalloc_dev: a fixed array in device memory
// (one copy per GPU, held until the process ends) instead of a context
// allocation.  Host code never passes `scratch` itself to a kernel (on the
// host its name is a stand-in, not a device address); spx_noalloc_setup
// looks up the device address once per context, with the context's device
// current, and the entry point passes that address.
static __device__ __align__(256) double scratch[SPEXHIP_NOALLOC_MAX_SCALES];

int spx_noalloc_setup(spexhip_ctx *ctx) {
  void *ptr = nullptr;
  SPX_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  SPX_CHECK(ctx, ptr != nullptr, "no-alloc scratch: null device address");
  ctx->noalloc_scales = ptr;
  return SPEXHIP_OK;
}

int spx_col_scales(spexhip_ctx *ctx, hipStream_t st, long n, int nb, const void *M, long ldm,
                   int colmajor, int nw, const double *W, int from_d, double *scale) {
  const long cnt = (long)nb * nw;
  hipLaunchKernelGGL(colscale::zero_u64_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0,
                     st, (unsigned long long *)scale, cnt);
  SPX_LAUNCHED(ctx);
  if (n == 0) return SPEXHIP_OK;
  const long tchunk = 1024;  // enough workgroups to keep HBM busy; few atomics per column
  dim3 grid((nb + 255) / 256, (unsigned)((n + tchunk - 1) / tchunk));
  for (int w0 = 0; w0 < nw; w0 += 16) {
    const int g = std::min(16, nw - w0);
    const double *Wg = W + (long)w0 * n;  // rows of n transitions
    unsigned long long *Sg = (unsigned long long *)scale + (long)w0 * nb;
    auto go = [&](auto cm, auto np, auto fd) {
      hipLaunchKernelGGL((colscale::col_scale_kernel<decltype(cm)::value, decltype(np)::value,
                                                     decltype(fd)::value>),
                         grid, dim3(256), 0, st, n, nb, (const double2 *)M, ldm, g, Wg, tchunk,
                         Sg);
    };
    auto go_nw = [&](auto cm, auto fd) {
      switch ((g + 3) / 4 * 4) {
        case 4: go(cm, std::integral_constant<int, 4>{}, fd); break;
        case 8: go(cm, std::integral_constant<int, 8>{}, fd); break;
        case 12: go(cm, std::integral_constant<int, 12>{}, fd); break;
        default: go(cm, std::integral_constant<int, 16>{}, fd); break;
      }
    };
    auto go_fd = [&](auto cm) {
      if (from_d) go_nw(cm, std::true_type{});
      else go_nw(cm, std::false_type{});
    };
    if (colmajor) go_fd(std::true_type{});
    else go_fd(std::false_type{});
    SPX_LAUNCHED(ctx);
  }
  return SPEXHIP_OK;
}

int spx_chi0_launch(spexhip_ctx *ctx, hipStream_t st, long n, int nb,
                    const void *M, long ldm, int colmajor, int nw,
                    const void *wght, int precision, double alpha, double beta,
                    void *chi, int ldchi, void *scale_ws) {
  SPX_CHECK(ctx, n >= 0 && nb > 0 && nw > 0, "chi0: bad sizes n=%ld nb=%d nw=%d", n, nb, nw);
  SPX_CHECK(ctx, ldchi >= nb, "chi0: ldchi=%d < nb=%d", ldchi, nb);
  SPX_CHECK(ctx, colmajor ? ldm >= n : ldm >= nb, "chi0: ldm=%ld too small", ldm);
  SPX_CHECK(ctx, precision == SPEXHIP_PREC_FP64 || precision == SPEXHIP_PREC_SPLIT_FP32 ||
                     precision == SPEXHIP_PREC_OZAKI_INT8,
            "chi0: unknown precision %d", precision);
  if (precision == SPEXHIP_PREC_OZAKI_INT8)
    return spx_chi0_ozaki(ctx, st, n, nb, M, ldm, colmajor, nw, wght, alpha, beta, chi, ldchi,
                          ctx->ozaki_ws_limit);
  const int split = precision == SPEXHIP_PREC_SPLIT_FP32 && n > 0;
  double *scale = nullptr;
  if (split) {
    if (scale_ws) {
      scale = (double *)scale_ws;
    } else {
      SPX_CHECK(ctx, st == ctx->stream, "chi0: split form off the context stream needs a workspace");
      void *p = nullptr;
      int rc = spx_colscale_buf(ctx, spx_split_ws_bytes(nb, nw), &p);
      if (rc) return rc;
      scale = (double *)p;
    }
    int rc = spx_col_scales(ctx, st, n, nb, M, ldm, colmajor, nw, (const double *)wght, 1, scale);
    if (rc) return rc;
  }
  Problem P{n, nb, nw, (const double2 *)M, ldm, (const double *)wght, (double2 *)chi,
            (long)ldchi, n == 0 ? 0.0 : alpha, beta, scale};
  launch_cfg<64, 16>(st, P, colmajor, split);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spx_weights_launch(spexhip_ctx *ctx, hipStream_t st, int nk, int nocc, int nunocc,
                       const void *e_occ, const void *e_unocc, int nkpt, int nsym1,
                       const void *nkpts, int nw, const void *omega, void *wght) {
  const long n = (long)nk * nocc * nunocc;
  if (n == 0) return SPEXHIP_OK;
  const int bs = 256;
  hipLaunchKernelGGL(weights_none_kernel, dim3((n + bs - 1) / bs), dim3(bs), 0, st, nk, nocc,
                     nunocc, (const double *)e_occ, (const double *)e_unocc, nkpt, nsym1,
                     (const int *)nkpts, nw, (const double *)omega, (double *)wght);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spx_pack_upper(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, const void *X, long ldx,
                   void *P) {
  const long npack = (long)nb * (nb + 1) / 2;
  const int bs = 256;
  hipLaunchKernelGGL(pack_upper_kernel, dim3((npack * nw + bs - 1) / bs), dim3(bs), 0, st, nb, nw,
                     (const double2 *)X, ldx, (double2 *)P);
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

int spexhip_chi0_noalloc_dev(spexhip_ctx *ctx, long n, int nb, const void *M, long ldm,
                             int colmajor, int nw, const void *wght, int precision,
                             double alpha, double beta, void *chi, int ldchi) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, precision == SPEXHIP_PREC_FP64 || precision == SPEXHIP_PREC_SPLIT_FP32,
            "chi0_noalloc: precision %d not available without allocation (the Ozaki form "
            "needs a slice workspace; use spexhip_chi0_dev)", precision);
  SPX_CHECK(ctx, nb > 0 && nw > 0 && n >= 0 && ldchi >= nb, "chi0_noalloc: bad sizes");
  SPX_CHECK(ctx, (long)nb * nw <= SPEXHIP_NOALLOC_MAX_SCALES,
            "chi0_noalloc: nb*nw = %ld exceeds the fixed scratch of %d column scales",
            (long)nb * nw, SPEXHIP_NOALLOC_MAX_SCALES);
  SPX_CHECK(ctx, ctx->noalloc_scales != nullptr, "chi0_noalloc: scratch address not set up");
  if (n > 0) {
    const size_t mbytes = colmajor ? 16 * ((size_t)ldm * (nb - 1) + n)
                                   : 16 * ((size_t)ldm * (n - 1) + nb);
    SPX_CHECK(ctx, spx_owns(ctx, M, mbytes), "chi0_noalloc: M is not a context allocation of the stated shape");
    SPX_CHECK(ctx, spx_owns(ctx, wght, sizeof(double) * n * nw), "chi0_noalloc: wght is not a context allocation of n*nw doubles");
  }
  SPX_CHECK(ctx, spx_owns(ctx, chi, 16 * ((size_t)ldchi * nb * (nw - 1) + (size_t)ldchi * (nb - 1) + nb)),
            "chi0_noalloc: chi is not a context allocation of the stated shape");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  return spx_chi0_launch(ctx, ctx->stream, n, nb, M, ldm, colmajor, nw, wght, precision,
                         alpha, beta, chi, ldchi, ctx->noalloc_scales);
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
