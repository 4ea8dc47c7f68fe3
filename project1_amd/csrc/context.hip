// Context, error reporting and budget-tracked device memory.
// The budget mirrors the reference's MEM accounting (checkmem.f:16-92):
// an allocation that would exceed it fails with SPEXHIP_E_NOMEM instead of
// aborting the process.
#include "common.hpp"

int spx_seterr(spexhip_ctx *ctx, int code, const char *fmt, ...) {
  if (!ctx) return code;
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->err = buf;
  return code;
}

// Allocations are rounded up to a size class (four per power of two) and a
// freed block is kept in the context's cache for the next request of its
// class, so a steady stream of pipelines of varying size does not call
// hipFree, which waits for every stream of the device.  The cache is
// returned to the device when an allocation fails and when the context ends.
static size_t size_class(size_t b) {
  if (b <= 4096) return 4096;
  const int e = 63 - __builtin_clzll(b);
  const size_t step = (size_t)1 << (e - 2);
  return (b + step - 1) / step * step;
}

size_t spx_alloc_bytes(size_t bytes) { return size_class(bytes); }

static void trim_cache(spexhip_ctx *ctx) {
  std::multimap<size_t, void *> blocks;
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    blocks.swap(ctx->cache);
    ctx->cached = 0;
  }
  for (auto &kv : blocks) (void)hipFree(kv.second);
}

int spx_alloc(spexhip_ctx *ctx, size_t bytes, void **p) {
  const size_t cls = size_class(bytes);
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->budget && ctx->used + cls > ctx->budget) {
      ctx->err = "HBM budget exceeded: used " + std::to_string(ctx->used) +
                 " + request " + std::to_string(cls) + " > budget " +
                 std::to_string(ctx->budget);
      return SPEXHIP_E_NOMEM;
    }
    ctx->used += cls;
    auto it = ctx->cache.find(cls);
    if (it != ctx->cache.end()) {
      *p = it->second;
      ctx->cache.erase(it);
      ctx->cached -= cls;
      ctx->allocs[*p] = {bytes, cls};
      return SPEXHIP_OK;
    }
  }
  hipError_t e = hipMalloc(p, cls);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    trim_cache(ctx);
    e = hipMalloc(p, cls);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();  // do not leave the failure for the next launch check
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->used -= cls;
    ctx->err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? SPEXHIP_E_NOMEM : SPEXHIP_E_DEVICE;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->allocs[*p] = {bytes, cls};
  return SPEXHIP_OK;
}

// The caller guarantees that no queued work still uses the block.
int spx_free(spexhip_ctx *ctx, void *p) {
  if (!p) return SPEXHIP_OK;
  std::lock_guard<std::mutex> g(ctx->mu);
  auto it = ctx->allocs.find(p);
  if (it == ctx->allocs.end()) {
    ctx->err = "spexhip_dev_free: pointer not owned by this context";
    return SPEXHIP_E_INVALID;
  }
  const size_t cls = it->second.cls;
  ctx->allocs.erase(it);
  ctx->used -= cls;
  ctx->cache.emplace(cls, p);
  ctx->cached += cls;
  return SPEXHIP_OK;
}

bool spx_owns(spexhip_ctx *ctx, const void *p, size_t bytes) {
  std::lock_guard<std::mutex> g(ctx->mu);
  const char *c = (const char *)p;
  for (auto &kv : ctx->allocs) {
    const char *b = (const char *)kv.first;
    if (c >= b && c + bytes <= b + kv.second.bytes) return true;
  }
  return false;
}

int spx_aux_stream(spexhip_ctx *ctx) {
  if (ctx->aux) return SPEXHIP_OK;
  SPX_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
  for (auto &e : ctx->aux_ev) SPX_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return SPEXHIP_OK;
}

int spx_scratch(spexhip_ctx *ctx, size_t bytes, void **p) {
  if (ctx->scratch_bytes >= bytes) {
    *p = ctx->scratch;
    return SPEXHIP_OK;
  }
  if (ctx->scratch) {
    SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int rc = spx_free(ctx, ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    if (rc) return rc;
  }
  int rc = spx_alloc(ctx, bytes, &ctx->scratch);
  if (rc) return rc;
  ctx->scratch_bytes = bytes;
  *p = ctx->scratch;
  return SPEXHIP_OK;
}

int spx_colscale_buf(spexhip_ctx *ctx, size_t bytes, void **p) {
  if (ctx->colscale_bytes >= bytes) {
    *p = ctx->colscale;
    return SPEXHIP_OK;
  }
  if (ctx->colscale) {
    SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int rc = spx_free(ctx, ctx->colscale);
    ctx->colscale = nullptr;
    ctx->colscale_bytes = 0;
    if (rc) return rc;
  }
  int rc = spx_alloc(ctx, bytes, &ctx->colscale);
  if (rc) return rc;
  ctx->colscale_bytes = bytes;
  *p = ctx->colscale;
  return SPEXHIP_OK;
}

extern "C" {

int spexhip_create(int device, size_t budget, spexhip_ctx **out) {
  if (!out) return SPEXHIP_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPEXHIP_E_DEVICE;
  if (device < 0 || device >= ndev) return SPEXHIP_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return SPEXHIP_E_DEVICE;
  spexhip_ctx *c = new spexhip_ctx();
  c->device = device;
  c->budget = budget;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SPEXHIP_E_DEVICE;
  }
  for (auto &e : c->phase_ev)
    if (hipEventCreate(&e) != hipSuccess) e = nullptr;
  if (spx_noalloc_setup(c) != SPEXHIP_OK) {  // `device` is current
    for (auto &e : c->phase_ev)
      if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return SPEXHIP_E_DEVICE;
  }
  *out = c;
  return SPEXHIP_OK;
}

int spexhip_destroy(spexhip_ctx *ctx) {
  if (!ctx) return SPEXHIP_OK;
  // servers still alive: stop their schedulers and drain their pipelines
  for (;;) {
    spexhip_server *sv = nullptr;
    {
      std::lock_guard<std::mutex> g(ctx->mu);
      if (ctx->servers.empty()) break;
      sv = ctx->servers.back();
    }
    spexhip_server_destroy(sv);  // unregisters itself
  }
  (void)hipSetDevice(ctx->device);
  // work on other streams (pipelines, batches) may still read context memory
  (void)hipDeviceSynchronize();
  for (auto &kv : ctx->allocs) (void)hipFree(kv.first);
  ctx->allocs.clear();
  trim_cache(ctx);
  for (auto &e : ctx->phase_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : ctx->aux_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SPEXHIP_OK;
}

int spexhip_last_error(spexhip_ctx *ctx, char *buf, size_t len) {
  if (!ctx || !buf || !len) return SPEXHIP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  snprintf(buf, len, "%s", ctx->err.c_str());
  return SPEXHIP_OK;
}

int spexhip_sync(spexhip_ctx *ctx) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SPEXHIP_OK;
}

int spexhip_mem_used(spexhip_ctx *ctx, size_t *used, size_t *budget) {
  if (!ctx) return SPEXHIP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (used) *used = ctx->used;
  if (budget) *budget = ctx->budget;
  return SPEXHIP_OK;
}

int spexhip_set_workspace_limit(spexhip_ctx *ctx, size_t bytes) {
  if (!ctx || bytes == 0) return SPEXHIP_E_INVALID;
  ctx->ozaki_ws_limit = bytes;
  return SPEXHIP_OK;
}

int spexhip_chi0_phase_ms(spexhip_ctx *ctx, float *prep_ms, float *rank_ms, int *groups) {
  if (!ctx || !prep_ms || !rank_ms) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, ctx->phase_groups > 0, "no Ozaki chi0 call recorded yet");
  SPX_HIP(ctx, hipEventSynchronize(ctx->phase_ev[2]));
  SPX_HIP(ctx, hipEventElapsedTime(prep_ms, ctx->phase_ev[0], ctx->phase_ev[1]));
  SPX_HIP(ctx, hipEventElapsedTime(rank_ms, ctx->phase_ev[1], ctx->phase_ev[2]));
  if (groups) *groups = ctx->phase_groups;
  return SPEXHIP_OK;
}

int spexhip_dev_alloc(spexhip_ctx *ctx, size_t bytes, void **dev) {
  if (!ctx || !dev) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  return spx_alloc(ctx, bytes, dev);
}

int spexhip_dev_free(spexhip_ctx *ctx, void *dev) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return spx_free(ctx, dev);
}

int spexhip_h2d(spexhip_ctx *ctx, void *dev, const void *host, size_t bytes) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (!bytes) return SPEXHIP_OK;
  SPX_CHECK(ctx, dev && host, "spexhip_h2d: null pointer");
  SPX_CHECK(ctx, spx_owns(ctx, dev, bytes), "spexhip_h2d: destination outside a context allocation");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SPEXHIP_OK;
}

int spexhip_d2h(spexhip_ctx *ctx, void *host, const void *dev, size_t bytes) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (!bytes) return SPEXHIP_OK;
  SPX_CHECK(ctx, dev && host, "spexhip_d2h: null pointer");
  SPX_CHECK(ctx, spx_owns(ctx, dev, bytes), "spexhip_d2h: source outside a context allocation");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SPEXHIP_OK;
}

int spexhip_memset(spexhip_ctx *ctx, void *dev, int value, size_t bytes) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (!bytes) return SPEXHIP_OK;
  SPX_CHECK(ctx, spx_owns(ctx, dev, bytes), "spexhip_memset: range outside a context allocation");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipMemsetAsync(dev, value, bytes, ctx->stream));
  return SPEXHIP_OK;
}

}  // extern "C"
