// Persistent grouped chi0 launch over many q-points of uneven size.
//
// Each problem p (one q-point: n_p transitions, basis size nb_p, nw_p
// frequencies) contributes nbt_p(nbt_p+1)/2 * nw_p tiles; the work list is
// ordered longest-first (largest n first, LPT scheduling) and consumed by a
// fixed grid of persistent workgroups through one device-scope atomic
// counter.  Every workgroup leaves the loop as soon as the counter passes
// the last item, so the grid always drains.  The per-tile arithmetic is
// chi0_tile.hpp (FP64 or split-FP32 MFMA), i.e. susceptibility.f:1183-1240
// per q.  Tile shapes (TM x TM outputs, KC-deep K slices) and the number of
// resident workgroups per CU are chosen by an autotuner that times each
// configuration on the actual batch.  In the split-FP32 form each problem's
// column scales (colscale.hpp) are computed right before the grouped launch.
#include <algorithm>
#include <numeric>
#include <vector>

#include "chi0_tile.hpp"
#include "colscale.hpp"
#include "internal.hpp"

namespace {

using chi0t::Lds;
using chi0t::Problem;

struct Cfg {
  int tm, kc, per_cu;
};
constexpr Cfg kCfgs[] = {{64, 16, 1}, {64, 16, 2}, {64, 32, 1}, {32, 16, 2},
                         {32, 16, 4}, {32, 32, 2}, {64, 8, 2},  {32, 8, 4}};
constexpr int kNCfg = sizeof(kCfgs) / sizeof(kCfgs[0]);

template <int TM, int KC, bool SPLIT>
__global__ __launch_bounds__(chi0t::NT) void chi0_grouped_kernel(
    const Problem *__restrict__ probs, const long *__restrict__ start, int nprob,
    unsigned int *__restrict__ counter, long total) {
  __shared__ Lds<TM, KC, SPLIT> s;
  __shared__ long item_s;
  for (;;) {
    if (threadIdx.x == 0) item_s = (long)atomicAdd(counter, 1u);
    __syncthreads();
    const long item = item_s;
    __syncthreads();  // every thread has read item_s before it is rewritten
    if (item >= total) break;
    int lo = 0, hi = nprob - 1;  // start[p] <= item < start[p+1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const Problem &P = probs[lo];
    const long local = item - start[lo];
    const int nbt = (P.nb + TM - 1) / TM;
    const long nt = (long)nbt * (nbt + 1) / 2;
    const int w = (int)(local / nt);
    int bi, bj;
    chi0t::tri_index((int)(local % nt), bi, bj);
    chi0t::tile<TM, KC, false, SPLIT>(P, bi, bj, w, s);
  }
}

__global__ void zero_counter_kernel(unsigned int *c) { c[0] = 0u; }

template <int TM, int KC>
void launch(hipStream_t st, int split, int grid, const Problem *probs, const long *start,
            int nprob, unsigned int *counter, long total) {
  if (split)
    hipLaunchKernelGGL((chi0_grouped_kernel<TM, KC, true>), dim3(grid), dim3(chi0t::NT), 0, st,
                       probs, start, nprob, counter, total);
  else
    hipLaunchKernelGGL((chi0_grouped_kernel<TM, KC, false>), dim3(grid), dim3(chi0t::NT), 0, st,
                       probs, start, nprob, counter, total);
}

int ncu_of(spexhip_ctx *ctx) {
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
    n = 256;
  return n;
}

}  // namespace

// Device table for a grouped launch: problems in LPT order + prefix sums +
// counter.  Built on the host, uploaded once (graph replays reuse it).
struct spexhip_group {
  int nprob = 0, cfg = 0, split = 0;
  long total = 0;
  // [Problem x nprob][long x (nprob+1)][uint counter][split: column scales]
  void *dev = nullptr;
  Problem *probs = nullptr;
  long *start = nullptr;
  unsigned int *counter = nullptr;
  std::vector<Problem> host;  // the problems as uploaded (scale passes)
};

static int group_build(spexhip_ctx *ctx, int nprob, const spexhip_chi0_problem *pp, int precision,
                       int cfg, spexhip_group *g) {
  SPX_CHECK(ctx, nprob > 0 && pp, "grouped chi0: no problems");
  SPX_CHECK(ctx, precision == SPEXHIP_PREC_FP64 || precision == SPEXHIP_PREC_SPLIT_FP32,
            "grouped chi0: precision must be FP64 or SPLIT_FP32");
  SPX_CHECK(ctx, cfg >= 0 && cfg < kNCfg, "grouped chi0: config %d out of range", cfg);
  const int tm = kCfgs[cfg].tm;
  std::vector<int> order(nprob);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pp[a].n > pp[b].n; });
  std::vector<Problem> probs(nprob);
  std::vector<long> start(nprob + 1, 0);
  for (int i = 0; i < nprob; i++) {
    const spexhip_chi0_problem &q = pp[order[i]];
    SPX_CHECK(ctx, q.n >= 0 && q.nb > 0 && q.nw > 0 && q.ldm >= q.nb && q.ldchi >= q.nb,
              "grouped chi0: bad problem %d", order[i]);
    SPX_CHECK(ctx, q.n == 0 || (spx_owns(ctx, q.M_dev, 16 * ((size_t)q.ldm * (q.n - 1) + q.nb)) &&
                                spx_owns(ctx, q.wght_dev, 8 * (size_t)q.n * q.nw)),
              "grouped chi0: problem %d inputs are not context allocations of the stated shape",
              order[i]);
    SPX_CHECK(ctx, spx_owns(ctx, q.chi_dev, 16 * (size_t)q.ldchi * q.nb * q.nw),
              "grouped chi0: problem %d output is not a context allocation of the stated shape",
              order[i]);
    probs[i] = Problem{q.n, q.nb, q.nw, (const double2 *)q.M_dev, q.ldm,
                       (const double *)q.wght_dev, (double2 *)q.chi_dev, (long)q.ldchi,
                       q.n == 0 ? 0.0 : q.alpha, q.beta, nullptr};
    const int nbt = (q.nb + tm - 1) / tm;
    start[i + 1] = start[i] + (long)nbt * (nbt + 1) / 2 * q.nw;
  }
  g->nprob = nprob;
  g->cfg = cfg;
  g->split = precision == SPEXHIP_PREC_SPLIT_FP32;
  g->total = start[nprob];
  const size_t pb = sizeof(Problem) * nprob, sb = sizeof(long) * (nprob + 1);
  size_t scb = 0;
  if (g->split)
    for (const auto &p : probs) scb += (spx_split_ws_bytes(p.nb, p.nw) + 255) & ~(size_t)255;
  int rc = spx_alloc(ctx, pb + sb + 64 + 256 + scb, &g->dev);
  if (rc) return rc;
  g->probs = (Problem *)g->dev;
  g->start = (long *)((char *)g->dev + pb);
  g->counter = (unsigned int *)((char *)g->dev + pb + sb);
  if (g->split) {
    char *sc = (char *)g->dev + ((pb + sb + 64 + 255) & ~(size_t)255);
    for (auto &p : probs) {
      p.scale = (const double *)sc;
      sc += (spx_split_ws_bytes(p.nb, p.nw) + 255) & ~(size_t)255;
    }
  }
  g->host = probs;
  SPX_HIP(ctx, hipMemcpy(g->probs, probs.data(), pb, hipMemcpyHostToDevice));
  SPX_HIP(ctx, hipMemcpy(g->start, start.data(), sb, hipMemcpyHostToDevice));
  return SPEXHIP_OK;
}

static int group_launch(spexhip_ctx *ctx, hipStream_t st, const spexhip_group *g) {
  const Cfg c = kCfgs[g->cfg];
  if (g->split)
    for (const Problem &p : g->host) {
      const int rc = spx_col_scales(ctx, st, p.n, p.nb, p.M, p.ldm, 0, p.nw, p.wght, 1,
                                    (double *)p.scale);
      if (rc) return rc;
    }
  hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, st, g->counter);
  SPX_LAUNCHED(ctx);
  const long grid = std::min<long>(g->total, (long)ncu_of(ctx) * c.per_cu);
  if (grid == 0) return SPEXHIP_OK;
  switch (g->cfg) {
    case 0: case 1: launch<64, 16>(st, g->split, (int)grid, g->probs, g->start, g->nprob, g->counter, g->total); break;
    case 2: launch<64, 32>(st, g->split, (int)grid, g->probs, g->start, g->nprob, g->counter, g->total); break;
    case 3: case 4: launch<32, 16>(st, g->split, (int)grid, g->probs, g->start, g->nprob, g->counter, g->total); break;
    case 5: launch<32, 32>(st, g->split, (int)grid, g->probs, g->start, g->nprob, g->counter, g->total); break;
    case 6: launch<64, 8>(st, g->split, (int)grid, g->probs, g->start, g->nprob, g->counter, g->total); break;
    default: launch<32, 8>(st, g->split, (int)grid, g->probs, g->start, g->nprob, g->counter, g->total); break;
  }
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spx_group_launch(spexhip_ctx *ctx, hipStream_t st, const spexhip_group *g) {
  return group_launch(ctx, st, g);
}

extern "C" {

int spexhip_chi0_configs(int *count, int *tm, int *kc, int *per_cu) {
  if (!count) return SPEXHIP_E_INVALID;
  for (int i = 0; i < kNCfg && i < *count; i++) {
    if (tm) tm[i] = kCfgs[i].tm;
    if (kc) kc[i] = kCfgs[i].kc;
    if (per_cu) per_cu[i] = kCfgs[i].per_cu;
  }
  *count = kNCfg;
  return SPEXHIP_OK;
}

int spexhip_group_create(spexhip_ctx *ctx, int nprob, const spexhip_chi0_problem *probs,
                         int precision, int config, spexhip_group **out) {
  if (!ctx || !out) return SPEXHIP_E_INVALID;
  *out = nullptr;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  spexhip_group *g = new spexhip_group();
  const int rc = group_build(ctx, nprob, probs, precision, config < 0 ? 0 : config, g);
  if (rc) {
    if (g->dev) spx_free(ctx, g->dev);
    delete g;
    return rc;
  }
  *out = g;
  return SPEXHIP_OK;
}

int spexhip_group_launch(spexhip_ctx *ctx, const spexhip_group *g) {
  if (!ctx || !g) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  return group_launch(ctx, ctx->stream, g);
}

int spexhip_group_destroy(spexhip_ctx *ctx, spexhip_group *g) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (!g) return SPEXHIP_OK;
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  spx_free(ctx, g->dev);
  delete g;
  return SPEXHIP_OK;
}

int spexhip_chi0_grouped_dev(spexhip_ctx *ctx, int nprob, const spexhip_chi0_problem *probs,
                             int precision, int config) {
  spexhip_group *g = nullptr;
  int rc = spexhip_group_create(ctx, nprob, probs, precision, config, &g);
  if (rc) return rc;
  rc = group_launch(ctx, ctx->stream, g);
  const int rc2 = spexhip_group_destroy(ctx, g);
  return rc ? rc : rc2;
}

// Time every configuration on the batch (reps launches each after one
// warm-up) and return the fastest.  The outputs are overwritten.
int spexhip_chi0_autotune(spexhip_ctx *ctx, int nprob, const spexhip_chi0_problem *probs,
                          int precision, int reps, int *best, float *ms) {
  if (!ctx || !best) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  hipEvent_t a, b;
  SPX_HIP(ctx, hipEventCreate(&a));
  SPX_HIP(ctx, hipEventCreate(&b));
  float bestms = 0;
  *best = 0;
  int rc = SPEXHIP_OK;
  for (int c = 0; c < kNCfg && !rc; c++) {
    spexhip_group *g = nullptr;
    rc = spexhip_group_create(ctx, nprob, probs, precision, c, &g);
    if (rc) break;
    rc = group_launch(ctx, ctx->stream, g);
    if (!rc) rc = hipEventRecord(a, ctx->stream) == hipSuccess ? 0 : SPEXHIP_E_DEVICE;
    for (int r = 0; r < std::max(reps, 1) && !rc; r++) rc = group_launch(ctx, ctx->stream, g);
    if (!rc) rc = hipEventRecord(b, ctx->stream) == hipSuccess ? 0 : SPEXHIP_E_DEVICE;
    float t = 0;
    if (!rc && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t, a, b) != hipSuccess))
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "autotune: event timing failed");
    t /= std::max(reps, 1);
    if (ms) ms[c] = t;
    if (!rc && (c == 0 || t < bestms)) {
      bestms = t;
      *best = c;
    }
    spexhip_group_destroy(ctx, g);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // extern "C"
