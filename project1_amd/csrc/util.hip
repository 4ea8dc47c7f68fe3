// Layout conversion for the host drop-in entry points: device products are
// transition-major (t*ldm + I); the reference's cprod is column-major
// (t fastest, wavefproducts.f:120,490).
#include <algorithm>

#include "common.hpp"

namespace {

__global__ void transpose_kernel(const double2 *__restrict__ M, long rows, int cols, long ldm,
                                 double2 *__restrict__ T) {
  __shared__ double2 tile[32][33];
  const long r0 = (long)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int j = ty; j < 32; j += 8) {
    const long r = r0 + j;
    const int c = c0 + tx;
    if (r < rows && c < cols) tile[j][tx] = M[r * ldm + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j;
    const long r = r0 + tx;
    if (r < rows && c < cols) T[r + rows * (long)c] = tile[tx][j];
  }
}

// Plain streaming copy, 16 B per thread per step (grid-stride).
__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long n16) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void fill_zero_kernel(unsigned int *__restrict__ p, long count) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < count) p[g] = 0u;
}

}  // namespace

// Zero `bytes` (a multiple of 4) with a kernel rather than a memset node, so
// that the call may be captured into a graph.
int spx_fill_zero(spexhip_ctx *ctx, hipStream_t st, void *p, size_t bytes) {
  const long cnt = (long)(bytes / 4);
  if (cnt == 0) return SPEXHIP_OK;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((cnt + 255) / 256), dim3(256), 0, st,
                     (unsigned int *)p, cnt);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spx_transpose_d2h(spexhip_ctx *ctx, hipStream_t st, const void *M, long rows, int cols,
                      long ldm, double *host) {
  void *T;
  const size_t bytes = 16 * (size_t)rows * cols;
  int rc = spx_scratch(ctx, bytes, &T);
  if (rc) return rc;
  dim3 grid((rows + 31) / 32, (cols + 31) / 32);
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, st, (const double2 *)M, rows, cols, ldm,
                     (double2 *)T);
  SPX_LAUNCHED(ctx);
  SPX_HIP(ctx, hipMemcpyAsync(host, T, bytes, hipMemcpyDeviceToHost, st));
  SPX_HIP(ctx, hipStreamSynchronize(st));
  return SPEXHIP_OK;
}

extern "C" int spexhip_copy_dev(spexhip_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (!bytes) return SPEXHIP_OK;
  SPX_CHECK(ctx, bytes % 16 == 0 && ((size_t)dst | (size_t)src) % 16 == 0,
            "copy: pointers and size must be multiples of 16 bytes");
  SPX_CHECK(ctx, spx_owns(ctx, dst, bytes) && spx_owns(ctx, src, bytes),
            "copy: ranges outside context allocations");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const long n16 = (long)(bytes / 16);
  const long grid = std::min<long>((n16 + 255) / 256, 8192);
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, ctx->stream, (const uint4 *)src,
                     (uint4 *)dst, n16);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}
