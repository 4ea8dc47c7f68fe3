// Synthetic workload generator on the device and HIP-event timing helpers.
//
// No reference counterpart: SPEX reads its inputs from DFT files.  The
// generator is counter-based (splitmix64 of (seed, index)) so that a value
// depends only on its index and the seed -- identical for any launch
// shape, and reproducible on the host (project1_amd/synth.py mirrors it).
#include "common.hpp"

namespace {

__host__ __device__ inline unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Uniform in (0, 1]: 53 random bits.
__device__ inline double u01(unsigned long long h) {
  return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

// Pairs (2i, 2i+1) are one Box-Muller draw from counter i.
__global__ void normal_kernel(double *__restrict__ out, long count, unsigned long long seed,
                              double scale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * i >= count) return;
  const unsigned long long base = splitmix64(seed ^ 0xD1B54A32D192ED03ull);
  const double u1 = u01(splitmix64(base + 2 * (unsigned long long)i));
  const double u2 = u01(splitmix64(base + 2 * (unsigned long long)i + 1));
  const double r = sqrt(-2.0 * log(u1)) * scale;
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  out[2 * i] = r * c;
  if (2 * i + 1 < count) out[2 * i + 1] = r * s;
}

// Complex (rows, cols, nmat) column-major array; element i is the draw of
// counter i (as normal_kernel) times scale[row + rows * mat].
__global__ void normal_rows_kernel(double2 *__restrict__ out, long rows, long cols, long nmat,
                                   unsigned long long seed, const double *__restrict__ scale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols * nmat) return;
  const unsigned long long base = splitmix64(seed ^ 0xD1B54A32D192ED03ull);
  const double u1 = u01(splitmix64(base + 2 * (unsigned long long)i));
  const double u2 = u01(splitmix64(base + 2 * (unsigned long long)i + 1));
  const long r = i % rows, m = i / (rows * cols);
  const double rr = sqrt(-2.0 * log(u1)) * scale[r + rows * m];
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  out[i] = make_double2(rr * c, rr * s);
}

}  // namespace

extern "C" {

int spexhip_synth_normal_rows_dev(spexhip_ctx *ctx, void *out, long rows, long cols, long nmat,
                                  unsigned long long seed, const void *scale_dev) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, rows >= 0 && cols >= 0 && nmat >= 0, "synth_normal_rows: negative extent");
  const long count = rows * cols * nmat;
  if (count == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, spx_owns(ctx, out, 16 * (size_t)count),
            "synth_normal_rows: output not a context allocation of rows*cols*nmat complex");
  SPX_CHECK(ctx, spx_owns(ctx, scale_dev, 8 * (size_t)rows * nmat),
            "synth_normal_rows: scale not a context allocation of rows*nmat doubles");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(normal_rows_kernel, dim3((count + 255) / 256), dim3(256), 0, ctx->stream,
                     (double2 *)out, rows, cols, nmat, seed, (const double *)scale_dev);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_synth_normal_dev(spexhip_ctx *ctx, void *out, long count, unsigned long long seed,
                             double scale) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, count >= 0, "synth_normal: negative count");
  if (count == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, spx_owns(ctx, out, 8 * (size_t)count), "synth_normal: output not a context allocation of count doubles");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const long pairs = (count + 1) / 2;
  hipLaunchKernelGGL(normal_kernel, dim3((pairs + 255) / 256), dim3(256), 0, ctx->stream,
                     (double *)out, count, seed, scale);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_event_create(spexhip_ctx *ctx, void **ev) {
  if (!ctx || !ev) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  hipEvent_t e;
  SPX_HIP(ctx, hipEventCreate(&e));
  *ev = (void *)e;
  return SPEXHIP_OK;
}

int spexhip_event_record(spexhip_ctx *ctx, void *ev) {
  if (!ctx || !ev) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipEventRecord((hipEvent_t)ev, ctx->stream));
  return SPEXHIP_OK;
}

int spexhip_event_elapsed(spexhip_ctx *ctx, void *a, void *b, float *ms) {
  if (!ctx || !a || !b || !ms) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipEventSynchronize((hipEvent_t)b));
  SPX_HIP(ctx, hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return SPEXHIP_OK;
}

int spexhip_event_destroy(spexhip_ctx *ctx, void *ev) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (ev) SPX_HIP(ctx, hipEventDestroy((hipEvent_t)ev));
  return SPEXHIP_OK;
}

}  // extern "C"
