// Batched uneven q-points: per k packet, every q's products and weights run
// on their own streams, then ONE persistent grouped chi0 launch covers all
// q (tiles in longest-first order, chi0 += packet contribution); after the
// last packet every q's eps = 1 - v^1/2 chi0 v^1/2 and eps^-1 run on their
// streams again.  The fork/join structure is captured once into a HIP graph
// and replayed.  In the Ozaki form each q's chi0 += packet runs right after
// its products on its own stream (the int8 rank-k kernel is persistent and
// fills the device by itself); the q of one stream share one slice
// workspace.
//
// Reference flow per q: correlation.f:404-747 (susceptibility -> dielectric
// matrix -> Minverse); the k packets are the reference's (nk, no, nu)
// packets of susceptibility.f:614-713, here chosen so that the products of
// all q of one packet are resident together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace {

constexpr int kMaxStreams = 8;

struct BatchQ {
  spexhip_qdesc d{};
  std::vector<int> k1, k2, g2;
  int *kdev = nullptr, *oidx = nullptr;
  void *M = nullptr, *D = nullptr, *X = nullptr, *E = nullptr, *ws = nullptr;
  int s = 0;  // stream slot
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct Sizes {
  size_t M, D, X, ws, tab;
};

// Per-q buffers for packets of at most kp k-points.
Sizes sizes_of(const spexhip_qdesc &d, int kp) {
  const int kc = std::min(kp, d.nk);
  const long n = (long)kc * d.nocc * d.nunocc;
  Sizes s{};
  s.M = align256(16 * (size_t)n * d.nb);
  s.D = align256(8 * (size_t)n * d.nw);
  s.X = align256(16 * (size_t)d.nb * d.nb * d.nw);
  size_t ws = 16 * (size_t)d.nb * d.nb * d.nw;  // dense v^1/2 temporary
  ws = std::max(ws, spx_inverse_ws(d.nb, d.nw));
  ws = std::max(ws, 16 * (size_t)d.nb * (d.nb + 1) / 2 * d.nw);  // packing
  if (d.mt_plan) ws = std::max(ws, spx_wfprod_mt_ws(d.mt_plan, kc, d.nocc, d.nunocc));
  if (d.pw_plan) ws = std::max(ws, spx_wfprod_pw_ws_core(d.pw_plan, d.nocc, d.nunocc));
  s.ws = align256(ws);
  s.tab = align256(8 * (size_t)d.nk + 4 * (size_t)d.nk * (d.pw_plan ? d.pw_plan->ngptm : 0));
  return s;
}

}  // namespace

struct spexhip_batch {
  std::vector<BatchQ> q;
  int kpacket = 0, npacket = 0, precision = 0, config = 0, nstream = 0;
  float tune_ms[16] = {};
  std::vector<spexhip_group *> groups;  // one per packet
  hipStream_t main = nullptr;
  hipStream_t st[kMaxStreams] = {};
  hipEvent_t fork = nullptr, join[kMaxStreams] = {}, ev0 = nullptr, ev1 = nullptr;
  int *flags = nullptr;  // per q: [0, nq) eps not positive definite, [nq, 2nq) mixed signs
  void *ozws[kMaxStreams] = {};  // Ozaki slice workspace per stream
  size_t ozwsb[kMaxStreams] = {};
  std::vector<void *> owned;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int fork_all(spexhip_ctx *ctx, spexhip_batch *b) {
  SPX_HIP(ctx, hipEventRecord(b->fork, b->main));
  for (int s = 0; s < b->nstream; s++) SPX_HIP(ctx, hipStreamWaitEvent(b->st[s], b->fork, 0));
  return SPEXHIP_OK;
}

int join_all(spexhip_ctx *ctx, spexhip_batch *b) {
  for (int s = 0; s < b->nstream; s++) {
    SPX_HIP(ctx, hipEventRecord(b->join[s], b->st[s]));
    SPX_HIP(ctx, hipStreamWaitEvent(b->main, b->join[s], 0));
  }
  return SPEXHIP_OK;
}

int packet_products(spexhip_ctx *ctx, spexhip_batch *b, BatchQ &q, int k0, int cnt) {
  const spexhip_qdesc &d = q.d;
  hipStream_t st = b->st[q.s];
  const long ldm = d.nb;
  int rc = SPEXHIP_OK;
  if (d.mt_plan) {
    rc = spx_wfprod_mt(ctx, st, d.mt_plan, d.cmt_dev, d.ncent, d.nband, d.nkpt, d.q, cnt,
                       q.kdev + k0, q.kdev + d.nk + k0, d.b1lo, d.nocc, d.b2lo, d.nunocc, q.M,
                       ldm, 0, q.ws);
    if (rc) return rc;
  }
  if (d.pw_plan) {
    const int col0 = d.mt_plan ? d.mt_plan->nbasp : 0;
    rc = spx_wfprod_pw_async(ctx, st, d.pw_plan, d.cpw_dev, d.nband, cnt, q.k1.data() + k0,
                             q.k2.data() + k0, q.oidx + (size_t)k0 * d.pw_plan->ngptm, d.b1lo,
                             d.nocc, d.b2lo, d.nunocc, q.M, ldm, col0, q.ws);
    if (rc) return rc;
  }
  return spx_weights_launch(ctx, st, cnt, d.nocc, d.nunocc,
                            (const double *)d.e_occ_dev + (size_t)k0 * d.nocc,
                            (const double *)d.e_unocc_dev + (size_t)k0 * d.nunocc, d.nkpt_w,
                            d.nsym1, (const int *)d.nkpts_dev + k0, d.nw, d.omega_dev, q.D);
}

int enqueue(spexhip_ctx *ctx, spexhip_batch *b) {
  const bool ozaki = b->precision == SPEXHIP_PREC_OZAKI_INT8;
  const int nq = (int)b->q.size();
  int rc = spx_fill_zero(ctx, b->main, b->flags, 8 * (size_t)nq);
  if (!rc) rc = fork_all(ctx, b);
  for (auto &q : b->q)
    if (!rc) rc = spx_fill_zero(ctx, b->st[q.s], q.X, 16 * (size_t)q.d.nb * q.d.nb * q.d.nw);
  for (int p = 0; p < b->npacket && !rc; p++) {
    if (p > 0) rc = fork_all(ctx, b);
    for (int i = 0; i < nq && !rc; i++) {
      BatchQ &q = b->q[i];
      const int k0 = p * b->kpacket;
      const int cnt = std::min(b->kpacket, q.d.nk - k0);
      if (cnt <= 0) continue;
      rc = packet_products(ctx, b, q, k0, cnt);
      if (!rc && ozaki)
        rc = spx_chi0_ozaki_async(ctx, b->st[q.s], (long)cnt * q.d.nocc * q.d.nunocc, q.d.nb, q.M,
                                  q.d.nb, 0, q.d.nw, q.D, q.d.alpha, 1.0, q.X, q.d.nb,
                                  b->ozws[q.s], b->ozwsb[q.s], b->flags + nq + i);
    }
    if (!rc) rc = join_all(ctx, b);
    if (!rc && b->groups[p]) rc = spx_group_launch(ctx, b->main, b->groups[p]);
  }
  if (!rc) rc = fork_all(ctx, b);
  for (size_t i = 0; i < b->q.size() && !rc; i++) {
    BatchQ &q = b->q[i];
    const spexhip_qdesc &d = q.d;
    hipStream_t st = b->st[q.s];
    rc = spx_dielectric(ctx, st, d.nb, d.nw, q.X, d.nb, d.vkind, d.vhalf_dev, q.E, d.nb, q.ws);
    if (!rc) rc = spx_inverse(ctx, st, d.nb, d.nw, q.E, d.nb, q.ws, b->flags + i, nullptr);
  }
  if (!rc) rc = join_all(ctx, b);
  return rc;
}

void destroy(spexhip_ctx *ctx, spexhip_batch *b) {
  if (b->main) (void)hipStreamSynchronize(b->main);
  for (int s = 0; s < b->nstream; s++)
    if (b->st[s]) (void)hipStreamSynchronize(b->st[s]);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto *g : b->groups)
    if (g) spexhip_group_destroy(ctx, g);
  for (void *p : b->owned) spx_free(ctx, p);
  for (hipEvent_t e : {b->fork, b->ev0, b->ev1})
    if (e) (void)hipEventDestroy(e);
  for (int s = 0; s < kMaxStreams; s++) {
    if (b->join[s]) (void)hipEventDestroy(b->join[s]);
    if (b->st[s]) (void)hipStreamDestroy(b->st[s]);
  }
  if (b->main) (void)hipStreamDestroy(b->main);
  delete b;
}

// Problems of packet p (q with no k-points left in it are skipped).
std::vector<spexhip_chi0_problem> packet_problems(const spexhip_batch *b, int p) {
  std::vector<spexhip_chi0_problem> v;
  for (const auto &q : b->q) {
    const int cnt = std::min(b->kpacket, q.d.nk - p * b->kpacket);
    if (cnt <= 0) continue;
    spexhip_chi0_problem pr{};
    pr.n = (long)cnt * q.d.nocc * q.d.nunocc;
    pr.nb = q.d.nb;
    pr.nw = q.d.nw;
    pr.M_dev = q.M;
    pr.ldm = q.d.nb;
    pr.wght_dev = q.D;
    pr.chi_dev = q.X;
    pr.ldchi = q.d.nb;
    pr.alpha = q.d.alpha;
    pr.beta = 1.0;
    v.push_back(pr);
  }
  return v;
}

}  // namespace

extern "C" {

int spexhip_batch_footprint(int nq, const spexhip_qdesc *descs, int kpacket, size_t *bytes) {
  if (nq <= 0 || !descs || kpacket <= 0 || !bytes) return SPEXHIP_E_INVALID;
  size_t t = 256 + 8 * (size_t)nq;
  size_t ozs[kMaxStreams] = {};
  const int nstream = std::min(nq, kMaxStreams);
  for (int i = 0; i < nq; i++) {
    const spexhip_qdesc &d = descs[i];
    const Sizes s = sizes_of(d, kpacket);
    t += s.M + s.D + 2 * s.X + s.ws + s.tab;
    if (d.precision == SPEXHIP_PREC_OZAKI_INT8) {
      const long n = (long)std::min(kpacket, d.nk) * d.nocc * d.nunocc;
      ozs[i % nstream] = std::max(ozs[i % nstream],
                                  1024 + spx_ozaki_ws_bytes(n, d.nb, d.nw, d.ozaki_group > 0 ? std::min(d.ozaki_group, d.nw) : d.nw));
    }
  }
  for (int s = 0; s < nstream; s++) t += ozs[s];
  *bytes = t;
  return SPEXHIP_OK;
}

int spexhip_batch_create(spexhip_ctx *ctx, int nq, const spexhip_qdesc *descs, int kpacket,
                         int config, int tune_k, spexhip_batch **out) {
  if (!ctx || !out) return SPEXHIP_E_INVALID;
  *out = nullptr;
  SPX_CHECK(ctx, nq > 0 && descs && kpacket > 0, "batch: bad arguments");
  const int prec = descs[0].precision;
  SPX_CHECK(ctx, prec >= 0 && prec <= 2, "batch: unknown precision");
  const bool ozaki = prec == SPEXHIP_PREC_OZAKI_INT8;
  int cfgs = 16;
  SPX_CHECK(ctx, spexhip_chi0_configs(&cfgs, nullptr, nullptr, nullptr) == 0 && config < cfgs,
            "batch: config %d out of range", config);
  for (int i = 0; i < nq; i++) {
    const spexhip_qdesc &d = descs[i];
    SPX_CHECK(ctx, d.precision == prec, "batch: q %d has another precision", i);
    SPX_CHECK(ctx, !d.M_dev && (d.mt_plan || d.pw_plan), "batch: q %d needs product plans", i);
    SPX_CHECK(ctx, d.nk > 0 && d.nocc > 0 && d.nunocc > 0 && d.nw > 0 && d.nw <= 16,
              "batch: q %d has bad sizes", i);
    const int ncols = (d.mt_plan ? d.mt_plan->nbasp : 0) + (d.pw_plan ? d.pw_plan->ngptm : 0);
    SPX_CHECK(ctx, ncols == d.nb, "batch: q %d plans give %d columns but nb=%d", i, ncols, d.nb);
    SPX_CHECK(ctx, d.k1 && d.k2 && (!d.pw_plan || d.g2), "batch: q %d lacks k lists", i);
    SPX_CHECK(ctx, !d.mt_plan || (d.cmt_dev && d.ncent >= d.mt_plan->ncent),
              "batch: q %d bad MT inputs", i);
    SPX_CHECK(ctx, !d.pw_plan || d.cpw_dev, "batch: q %d lacks cpw", i);
    SPX_CHECK(ctx, d.e_occ_dev && d.e_unocc_dev && d.nkpts_dev && d.omega_dev && d.vhalf_dev,
              "batch: q %d lacks energies, k weights, frequencies or v^1/2", i);
    SPX_CHECK(ctx, d.vkind == 0 || d.vkind == 1, "batch: q %d vkind must be 0 or 1", i);
  }
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  spexhip_batch *b = new spexhip_batch();
  b->kpacket = kpacket;
  b->precision = prec;
  b->nstream = std::min(nq, kMaxStreams);
  b->q.resize(nq);
  int rc = SPEXHIP_OK;
  auto grab = [&](size_t bytes, void **p) {
    if (rc) return;
    rc = spx_alloc(ctx, bytes, p);
    if (!rc) b->owned.push_back(*p);
  };
  void *fl = nullptr;
  grab(256 + 8 * (size_t)nq, &fl);
  b->flags = (int *)fl;
  int maxnk = 0;
  for (int i = 0; i < nq && !rc; i++) {
    BatchQ &q = b->q[i];
    q.d = descs[i];
    q.s = i % b->nstream;
    q.k1.assign(q.d.k1, q.d.k1 + q.d.nk);
    q.k2.assign(q.d.k2, q.d.k2 + q.d.nk);
    if (q.d.pw_plan) q.g2.assign(q.d.g2, q.d.g2 + 3 * q.d.nk);
    q.d.k1 = q.k1.data();
    q.d.k2 = q.k2.data();
    q.d.g2 = q.g2.empty() ? nullptr : q.g2.data();
    maxnk = std::max(maxnk, q.d.nk);
    const Sizes s = sizes_of(q.d, kpacket);
    grab(s.M, &q.M);
    grab(s.D, &q.D);
    grab(s.X, &q.X);
    grab(s.X, &q.E);
    grab(s.ws, &q.ws);
    void *tab = nullptr;
    grab(s.tab, &tab);
    if (rc) break;
    q.kdev = (int *)tab;
    q.oidx = q.kdev + 2 * q.d.nk;
    std::vector<int> t(2 * (size_t)q.d.nk +
                       (q.d.pw_plan ? (size_t)q.d.nk * q.d.pw_plan->ngptm : 0));
    std::copy(q.k1.begin(), q.k1.end(), t.begin());
    std::copy(q.k2.begin(), q.k2.end(), t.begin() + q.d.nk);
    if (q.d.pw_plan) rc = spx_pw_oidx(ctx, q.d.pw_plan, q.d.nk, q.g2.data(), t.data() + 2 * q.d.nk);
    if (!rc && hipMemcpy(q.kdev, t.data(), 4 * t.size(), hipMemcpyHostToDevice) != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "batch: table upload failed");
  }
  b->npacket = (maxnk + kpacket - 1) / kpacket;
  hipError_t e = hipSuccess;
  if (!rc) {
    e = hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking);
    for (int s = 0; s < b->nstream && e == hipSuccess; s++) {
      e = hipStreamCreateWithFlags(&b->st[s], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[s], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "batch: stream/event creation failed");
  }
  // Ozaki: one slice workspace per stream, sized for its largest packet
  for (int s = 0; s < b->nstream && !rc && ozaki; s++) {
    size_t need = 0;
    for (const auto &q : b->q)
      if (q.s == s) {
        const long n = (long)std::min(kpacket, q.d.nk) * q.d.nocc * q.d.nunocc;
        need = std::max(need, spx_ozaki_ws_bytes(n, q.d.nb, q.d.nw,
                                                 q.d.ozaki_group > 0 ? std::min(q.d.ozaki_group, q.d.nw) : q.d.nw));
      }
    void *raw = nullptr;
    grab(need + 1024, &raw);  // the slice kernels want a 1 KB aligned workspace
    b->ozws[s] = (void *)(((uintptr_t)raw + 1023) & ~(uintptr_t)1023);
    b->ozwsb[s] = need;
  }
  // tile configuration (grouped kernel): given, or timed on packet 0 with
  // every q's K scaled down by one common factor (so the longest-first
  // balance is preserved)
  if (!rc && ozaki) {
    b->config = -1;
  } else if (!rc && config < 0) {
    std::vector<spexhip_chi0_problem> pr = packet_problems(b, 0);
    long nmax = 1;
    for (auto &x : pr) nmax = std::max(nmax, x.n);
    const double f = tune_k > 0 ? std::min(1.0, (double)tune_k / nmax) : 1.0;
    for (auto &x : pr) x.n = std::max(1L, (long)(x.n * f));
    rc = spexhip_chi0_autotune(ctx, (int)pr.size(), pr.data(), prec, 1, &b->config, b->tune_ms);
  } else {
    b->config = config;
  }
  for (int p = 0; p < b->npacket && !rc; p++) {
    std::vector<spexhip_chi0_problem> pr = packet_problems(b, p);
    spexhip_group *g = nullptr;
    if (!pr.empty() && !ozaki)
      rc = spexhip_group_create(ctx, (int)pr.size(), pr.data(), prec, b->config, &g);
    b->groups.push_back(g);
  }
  if (rc) {
    destroy(ctx, b);
    return rc;
  }
  *out = b;
  return SPEXHIP_OK;
}

int spexhip_batch_config(spexhip_ctx *ctx, const spexhip_batch *b, int *config, int *npacket,
                         float *tune_ms) {
  if (!ctx || !b) return SPEXHIP_E_INVALID;
  if (config) *config = b->config;
  if (npacket) *npacket = b->npacket;
  if (tune_ms) {
    int n = 16;
    spexhip_chi0_configs(&n, nullptr, nullptr, nullptr);
    std::copy(b->tune_ms, b->tune_ms + n, tune_ms);
  }
  return SPEXHIP_OK;
}

int spexhip_batch_launch(spexhip_ctx *ctx, spexhip_batch *b, int use_graph) {
  if (!ctx || !b) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipEventRecord(b->ev0, b->main));
  if (use_graph) {
    if (!b->exec) {
      SPX_HIP(ctx, hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue(ctx, b);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->main, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess)
        return spx_seterr(ctx, SPEXHIP_E_DEVICE, "batch: capture failed: %s", hipGetErrorString(e));
      b->graph = g;
      SPX_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    SPX_HIP(ctx, hipGraphLaunch(b->exec, b->main));
  } else {
    const int rc = enqueue(ctx, b);
    if (rc) return rc;
  }
  SPX_HIP(ctx, hipEventRecord(b->ev1, b->main));
  return SPEXHIP_OK;
}

int spexhip_batch_wait(spexhip_ctx *ctx, spexhip_batch *b, float *ms) {
  if (!ctx || !b) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipEventSynchronize(b->ev1));
  if (ms) SPX_HIP(ctx, hipEventElapsedTime(ms, b->ev0, b->ev1));
  const size_t nq = b->q.size();
  std::vector<int> f(2 * nq);
  SPX_HIP(ctx, hipMemcpy(f.data(), b->flags, 4 * f.size(), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < nq; i++)
    if (f[nq + i]) return spx_seterr(ctx, SPEXHIP_E_INVALID, "batch: chi0 weights of q %zu have mixed signs (Ozaki form)", i);
  for (size_t i = 0; i < nq; i++)
    if (f[i]) return spx_seterr(ctx, SPEXHIP_E_NOTPD, "batch: eps of q %zu not positive definite", i);
  return SPEXHIP_OK;
}

int spexhip_batch_result(spexhip_ctx *ctx, spexhip_batch *b, int iq, double *chi_packed,
                         double *einv_packed) {
  if (!ctx || !b) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, iq >= 0 && iq < (int)b->q.size(), "batch: q index %d out of range", iq);
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipEventSynchronize(b->ev1));
  BatchQ &q = b->q[iq];
  const int nb = q.d.nb, nw = q.d.nw;
  const size_t bytes = 16 * (size_t)nb * (nb + 1) / 2 * nw;
  hipStream_t st = b->st[q.s];
  for (int k = 0; k < 2; k++) {
    double *dst = k ? einv_packed : chi_packed;
    if (!dst) continue;
    int rc = spx_pack_upper(ctx, st, nb, nw, k ? q.E : q.X, nb, q.ws);
    if (rc) return rc;
    SPX_HIP(ctx, hipMemcpyAsync(dst, q.ws, bytes, hipMemcpyDeviceToHost, st));
    SPX_HIP(ctx, hipStreamSynchronize(st));
  }
  return SPEXHIP_OK;
}

int spexhip_batch_destroy(spexhip_ctx *ctx, spexhip_batch *b) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (b) destroy(ctx, b);
  return SPEXHIP_OK;
}

}  // extern "C"
