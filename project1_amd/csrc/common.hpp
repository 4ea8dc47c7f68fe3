// Shared host/device helpers of the spexhip library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/spexhip.h"

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

struct spexhip_ctx {
  int device = 0;
  size_t budget = 0;
  size_t used = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::mutex mu;  // guards err, used, allocs, cache
  struct Block {
    size_t bytes, cls;  // requested size (ownership checks), size class (cache)
  };
  std::unordered_map<void *, Block> allocs;
  std::multimap<size_t, void *> cache;  // freed blocks by size class
  size_t cached = 0;
  // scratch for host drop-in calls (grown on demand, owned by the context)
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // column scales of the split-FP32 chi0 form for calls on the context
  // stream (grown on demand; pipelines and batches carry their own)
  void *colscale = nullptr;
  size_t colscale_bytes = 0;
  // device address of the fixed column-scale array of
  // spexhip_chi0_noalloc_dev on this context's device (set at creation)
  void *noalloc_scales = nullptr;
  // largest slice workspace of the Ozaki form (frequencies are grouped to fit)
  size_t ozaki_ws_limit = (size_t)24 << 30;
  // the ALIGNBD input key: the tetrahedron-weight drop-ins stop before the
  // averaging over degenerate states and equivalent k points
  int alignbd = 0;
  // events bracketing the phases of the last Ozaki chi0 call (slice of the
  // first frequency group | rank-k updates and later groups)
  hipEvent_t phase_ev[3] = {nullptr, nullptr, nullptr};
  int phase_groups = 0;
  // second stream of the context and its events (created on first use):
  // the Ozaki form slices the next transition part on it while the rank-k
  // of the current part runs on the context stream
  hipStream_t aux = nullptr;
  hipEvent_t aux_ev[9] = {};
  // live request servers (each runs a scheduler thread that uses the
  // context); spexhip_destroy stops them before it frees anything
  std::vector<struct spexhip_server *> servers;
};

int spx_seterr(spexhip_ctx *ctx, int code, const char *fmt, ...);

#define SPX_HIP(ctx, call)                                                       \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return spx_seterr((ctx), SPEXHIP_E_DEVICE, "%s failed: %s (%s:%d)", #call, \
                        hipGetErrorString(e_), __FILE__, __LINE__);              \
  } while (0)

#define SPX_LAUNCHED(ctx) SPX_HIP(ctx, hipGetLastError())

#define SPX_CHECK(ctx, cond, ...)                                   \
  do {                                                              \
    if (!(cond)) return spx_seterr((ctx), SPEXHIP_E_INVALID, __VA_ARGS__); \
  } while (0)

// Allocation tracked against the context budget.
int spx_alloc(spexhip_ctx *ctx, size_t bytes, void **p);
// Bytes spx_alloc reserves for a request of `bytes` (its size class).
size_t spx_alloc_bytes(size_t bytes);
// Creates ctx->aux and ctx->aux_ev on first use (default priority).
int spx_aux_stream(spexhip_ctx *ctx);
int spx_free(spexhip_ctx *ctx, void *p);
// Context-owned scratch of at least `bytes` (contents undefined).
int spx_scratch(spexhip_ctx *ctx, size_t bytes, void **p);
// Context-owned buffer for split-FP32 column scales (context stream only).
int spx_colscale_buf(spexhip_ctx *ctx, size_t bytes, void **p);
// Is `p` (or p+bytes-1) inside an allocation of this context?
bool spx_owns(spexhip_ctx *ctx, const void *p, size_t bytes);
// Looks up the device address of the no-alloc column-scale array (chi0.hip)
// on the current device and stores it in ctx->noalloc_scales.
int spx_noalloc_setup(spexhip_ctx *ctx);

// Internal launchers shared between translation units.
int spx_chi0_ozaki(spexhip_ctx *ctx, hipStream_t st, long n, int nb, const void *M, long ldm,
                   int colmajor, int nw, const void *wght, double alpha, double beta, void *chi,
                   int ldchi, size_t ws_limit);
// scale_ws: nb*nw doubles for the split-FP32 column scales; nullptr takes
// the context's own buffer (then st must be the context stream).
int spx_chi0_launch(spexhip_ctx *ctx, hipStream_t st, long n, int nb,
                    const void *M, long ldm, int colmajor, int nw,
                    const void *wght, int precision, double alpha, double beta,
                    void *chi, int ldchi, void *scale_ws = nullptr);
// Bytes of scale_ws for the split-FP32 form.
size_t spx_split_ws_bytes(int nb, int nw);
