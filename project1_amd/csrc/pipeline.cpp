// Per-q pipeline: products -> weights -> chi0 -> eps -> eps^-1 on one
// stream, with every buffer and table allocated at creation so that a run
// only enqueues kernels (capturable into a HIP graph and replayed).
//
// Reference flow: correlation.f:404-747 for one q (susceptibility ->
// dielectric matrix -> Minverse), with the products of
// susceptibility.f:756-832 and the HILBERT NONE weights :949-975.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.hpp"

struct spexhip_qpipe {
  spexhip_qdesc d{};
  std::vector<int> k1, k2, g2;
  long n = 0;
  int kp = 0, npacket = 1;  // k-points per packet, packets
  hipStream_t st = nullptr;
  void *M = nullptr;
  bool ownM = false;
  long ldm = 0;
  void *D = nullptr, *X = nullptr, *E = nullptr, *ws = nullptr;
  size_t wsb = 0;
  int *kdev = nullptr, *oidx = nullptr, *flags = nullptr;  // flags[0]: not PD, [1]: mixed signs
  std::vector<void *> owned;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  size_t footprint = 0;
};

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct Sizes {
  size_t M, D, X, ws, tab;
};

int packet_k(const spexhip_qdesc &d) {
  return d.kpacket > 0 && d.kpacket < d.nk ? d.kpacket : (d.nk > 0 ? d.nk : 1);
}

// Buffers are sized for one packet of packet_k(d) k-points.
Sizes sizes_of(const spexhip_qdesc &d) {
  Sizes s{};
  const int kp = packet_k(d);
  const long n = (long)kp * d.nocc * d.nunocc;
  s.M = d.M_dev ? 0 : align256(16 * (size_t)n * d.nb);
  s.D = align256(8 * (size_t)n * d.nw);
  s.X = align256(16 * (size_t)d.nb * d.nb * d.nw);
  size_t ws = 16 * (size_t)d.nb * d.nb * d.nw;          // dense v^1/2 temporary
  ws = std::max(ws, spx_inverse_ws(d.nb, d.nw));
  ws = std::max(ws, 16 * (size_t)d.nb * (d.nb + 1) / 2 * d.nw);  // packing
  ws = std::max(ws, spx_split_ws_bytes(d.nb, d.nw));  // split-FP32 column scales
  if (d.precision == SPEXHIP_PREC_OZAKI_INT8)
    ws = std::max(ws, spx_ozaki_ws_bytes(n, d.nb, d.nw, std::min(d.nw, d.ozaki_group > 0 ? d.ozaki_group : d.nw)));
  if (d.mt_plan) ws = std::max(ws, spx_wfprod_mt_ws(d.mt_plan, kp, d.nocc, d.nunocc));
  if (d.pw_plan) ws = std::max(ws, spx_wfprod_pw_ws_core(d.pw_plan, d.nocc, d.nunocc));
  s.ws = align256(ws);
  s.tab = align256(8 * (size_t)d.nk + 4 * (size_t)d.nk * (d.pw_plan ? d.pw_plan->ngptm : 0) + 64);
  return s;
}

// Device bytes create() reserves: each buffer at its allocator size class.
size_t footprint_of(const Sizes &s) {
  size_t b = 0;
  for (size_t x : {s.M, s.D, s.X, s.X, s.ws, s.tab})
    if (x) b += spx_alloc_bytes(x);
  return b;
}

// The k range runs in packets (susceptibility.f:614-713): per packet the
// products, weights and chi0 += packet; eps and eps^-1 after the last one.
int enqueue(spexhip_ctx *ctx, spexhip_qpipe *q) {
  const spexhip_qdesc &d = q->d;
  hipStream_t st = q->st;
  int rc = spx_fill_zero(ctx, st, q->flags, 64);
  if (rc) return rc;
  for (int p = 0; p < q->npacket; p++) {
    const int k0 = p * q->kp;
    const int cnt = std::max(0, std::min(q->kp, d.nk - k0));
    const long np = (long)cnt * d.nocc * d.nunocc;
    void *Mp = d.M_dev ? (void *)((char *)q->M + 16 * (size_t)k0 * d.nocc * d.nunocc * q->ldm)
                       : q->M;
    if (cnt > 0) {
      if (d.mt_plan) {
        rc = spx_wfprod_mt(ctx, st, d.mt_plan, d.cmt_dev, d.ncent, d.nband, d.nkpt, d.q, cnt,
                           q->kdev + k0, q->kdev + d.nk + k0, d.b1lo, d.nocc, d.b2lo, d.nunocc,
                           Mp, q->ldm, 0, q->ws);
        if (rc) return rc;
      }
      if (d.pw_plan) {
        const int col0 = d.mt_plan ? d.mt_plan->nbasp : 0;
        rc = spx_wfprod_pw_async(ctx, st, d.pw_plan, d.cpw_dev, d.nband, cnt, q->k1.data() + k0,
                                 q->k2.data() + k0, q->oidx + (size_t)k0 * d.pw_plan->ngptm,
                                 d.b1lo, d.nocc, d.b2lo, d.nunocc, Mp, q->ldm, col0, q->ws);
        if (rc) return rc;
      }
      rc = spx_weights_launch(ctx, st, cnt, d.nocc, d.nunocc,
                              (const double *)d.e_occ_dev + (size_t)k0 * d.nocc,
                              (const double *)d.e_unocc_dev + (size_t)k0 * d.nunocc, d.nkpt_w,
                              d.nsym1, (const int *)d.nkpts_dev + k0, d.nw, d.omega_dev, q->D);
      if (rc) return rc;
    }
    const double beta = p > 0 ? 1.0 : 0.0;
    if (d.precision == SPEXHIP_PREC_OZAKI_INT8)
      rc = spx_chi0_ozaki_async(ctx, st, np, d.nb, Mp, q->ldm, 0, d.nw, q->D, d.alpha, beta, q->X,
                                d.nb, q->ws, q->wsb, q->flags + 1);
    else
      rc = spx_chi0_launch(ctx, st, np, d.nb, Mp, q->ldm, 0, d.nw, q->D, d.precision, d.alpha,
                           beta, q->X, d.nb, q->ws);
    if (rc) return rc;
  }
  rc = spx_dielectric(ctx, st, d.nb, d.nw, q->X, d.nb, d.vkind, d.vhalf_dev, q->E, d.nb, q->ws);
  if (rc) return rc;
  return spx_inverse(ctx, st, d.nb, d.nw, q->E, d.nb, q->ws, q->flags, nullptr);
}

void destroy(spexhip_ctx *ctx, spexhip_qpipe *q) {
  if (q->st) (void)hipStreamSynchronize(q->st);
  if (q->exec) (void)hipGraphExecDestroy(q->exec);
  if (q->graph) (void)hipGraphDestroy(q->graph);
  for (void *p : q->owned) spx_free(ctx, p);
  if (q->ev0) (void)hipEventDestroy(q->ev0);
  if (q->ev1) (void)hipEventDestroy(q->ev1);
  if (q->st) (void)hipStreamDestroy(q->st);
  delete q;
}

}  // namespace

extern "C" {

int spexhip_qpipe_footprint(const spexhip_qdesc *d, size_t *bytes) {
  if (!d || !bytes) return SPEXHIP_E_INVALID;
  *bytes = footprint_of(sizes_of(*d));
  return SPEXHIP_OK;
}

int spexhip_qpipe_create(spexhip_ctx *ctx, const spexhip_qdesc *desc, spexhip_qpipe **out) {
  if (!ctx || !desc || !out) return SPEXHIP_E_INVALID;
  *out = nullptr;
  const spexhip_qdesc &d = *desc;
  SPX_CHECK(ctx, d.nk >= 0 && d.nocc >= 0 && d.nunocc >= 0 && d.nb > 0 && d.nw > 0 && d.nw <= 16,
            "qpipe: bad sizes");
  SPX_CHECK(ctx, d.M_dev || d.mt_plan || d.pw_plan, "qpipe: neither products nor plans given");
  SPX_CHECK(ctx, !d.M_dev || d.ldm >= d.nb, "qpipe: ldm < nb");
  SPX_CHECK(ctx, d.e_occ_dev && d.e_unocc_dev && d.nkpts_dev && d.omega_dev && d.vhalf_dev,
            "qpipe: missing energies, k weights, frequencies or v^1/2");
  SPX_CHECK(ctx, d.vkind == 0 || d.vkind == 1, "qpipe: vkind must be 0 or 1");
  SPX_CHECK(ctx, d.precision >= 0 && d.precision <= 2, "qpipe: unknown precision");
  if (!d.M_dev) {
    const int ncols = (d.mt_plan ? d.mt_plan->nbasp : 0) + (d.pw_plan ? d.pw_plan->ngptm : 0);
    SPX_CHECK(ctx, ncols == d.nb, "qpipe: plans give %d product columns but nb=%d", ncols, d.nb);
    SPX_CHECK(ctx, d.k1 && d.k2 && (!d.pw_plan || d.g2), "qpipe: missing k lists");
    SPX_CHECK(ctx, !d.mt_plan || (d.cmt_dev && d.ncent >= d.mt_plan->ncent), "qpipe: bad MT inputs");
    SPX_CHECK(ctx, !d.pw_plan || d.cpw_dev, "qpipe: missing cpw");
  }
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  spexhip_qpipe *q = new spexhip_qpipe();
  q->d = d;
  q->n = (long)d.nk * d.nocc * d.nunocc;
  q->kp = packet_k(d);
  q->npacket = d.nk > 0 ? (d.nk + q->kp - 1) / q->kp : 1;
  if (!d.M_dev) {
    q->k1.assign(d.k1, d.k1 + d.nk);
    q->k2.assign(d.k2, d.k2 + d.nk);
    if (d.pw_plan) q->g2.assign(d.g2, d.g2 + 3 * d.nk);
  }
  const Sizes s = sizes_of(d);
  q->footprint = footprint_of(s);
  int rc = SPEXHIP_OK;
  auto grab = [&](size_t bytes, void **p) {
    if (rc) return;
    rc = spx_alloc(ctx, bytes, p);
    if (!rc) q->owned.push_back(*p);
  };
  if (d.M_dev) {
    q->M = const_cast<void *>(d.M_dev);
    q->ldm = d.ldm;
  } else {
    grab(s.M, &q->M);
    q->ownM = true;
    q->ldm = d.nb;
  }
  grab(s.D, &q->D);
  grab(s.X, &q->X);
  grab(s.X, &q->E);
  grab(s.ws, &q->ws);
  q->wsb = s.ws;
  void *tab = nullptr;
  grab(s.tab, &tab);
  if (rc) {
    destroy(ctx, q);
    return rc;
  }
  q->flags = (int *)tab;
  q->kdev = q->flags + 16;
  q->oidx = q->kdev + 2 * d.nk;
  hipError_t e = hipStreamCreateWithFlags(&q->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&q->ev0);
  if (e == hipSuccess) e = hipEventCreate(&q->ev1);
  if (e != hipSuccess) {
    destroy(ctx, q);
    return spx_seterr(ctx, SPEXHIP_E_DEVICE, "qpipe: stream/event creation failed");
  }
  if (!d.M_dev) {
    std::vector<int> t(2 * (size_t)d.nk + (d.pw_plan ? (size_t)d.nk * d.pw_plan->ngptm : 0));
    for (int i = 0; i < d.nk; i++) {
      t[i] = q->k1[i];
      t[d.nk + i] = q->k2[i];
    }
    if (d.pw_plan) {
      rc = spx_pw_oidx(ctx, d.pw_plan, d.nk, q->g2.data(), t.data() + 2 * d.nk);
      if (rc) {
        destroy(ctx, q);
        return rc;
      }
    }
    e = hipMemcpy(q->kdev, t.data(), 4 * t.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      destroy(ctx, q);
      return spx_seterr(ctx, SPEXHIP_E_DEVICE, "qpipe: table upload failed");
    }
  }
  *out = q;
  return SPEXHIP_OK;
}

int spexhip_qpipe_launch(spexhip_ctx *ctx, spexhip_qpipe *q, int use_graph) {
  if (!ctx || !q) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipEventRecord(q->ev0, q->st));
  if (use_graph) {
    if (!q->exec) {
      SPX_HIP(ctx, hipStreamBeginCapture(q->st, hipStreamCaptureModeThreadLocal));
      int rc = enqueue(ctx, q);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(q->st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess)
        return spx_seterr(ctx, SPEXHIP_E_DEVICE, "qpipe: capture failed: %s", hipGetErrorString(e));
      q->graph = g;
      SPX_HIP(ctx, hipGraphInstantiate(&q->exec, q->graph, nullptr, nullptr, 0));
    }
    SPX_HIP(ctx, hipGraphLaunch(q->exec, q->st));
  } else {
    int rc = enqueue(ctx, q);
    if (rc) return rc;
  }
  SPX_HIP(ctx, hipEventRecord(q->ev1, q->st));
  return SPEXHIP_OK;
}

int spexhip_qpipe_query(spexhip_ctx *ctx, spexhip_qpipe *q) {
  if (!ctx || !q) return SPEXHIP_E_INVALID;
  const hipError_t e = hipEventQuery(q->ev1);
  if (e == hipErrorNotReady) return SPEXHIP_E_AGAIN;
  if (e != hipSuccess) return spx_seterr(ctx, SPEXHIP_E_DEVICE, "qpipe: %s", hipGetErrorString(e));
  return SPEXHIP_OK;
}

int spexhip_qpipe_wait(spexhip_ctx *ctx, spexhip_qpipe *q, float *ms) {
  if (!ctx || !q) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipEventSynchronize(q->ev1));
  if (ms) SPX_HIP(ctx, hipEventElapsedTime(ms, q->ev0, q->ev1));
  int f[2] = {0, 0};
  SPX_HIP(ctx, hipMemcpy(f, q->flags, sizeof f, hipMemcpyDeviceToHost));
  if (f[1]) return spx_seterr(ctx, SPEXHIP_E_INVALID, "qpipe: chi0 weights of mixed sign (Ozaki form)");
  if (f[0]) return spx_seterr(ctx, SPEXHIP_E_NOTPD, "qpipe: dielectric matrix not positive definite");
  return SPEXHIP_OK;
}

int spexhip_qpipe_result(spexhip_ctx *ctx, spexhip_qpipe *q, double *chi_packed,
                         double *einv_packed) {
  if (!ctx || !q) return SPEXHIP_E_INVALID;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipEventSynchronize(q->ev1));
  const int nb = q->d.nb, nw = q->d.nw;
  const size_t bytes = 16 * (size_t)nb * (nb + 1) / 2 * nw;
  for (int k = 0; k < 2; k++) {
    double *dst = k ? einv_packed : chi_packed;
    if (!dst) continue;
    int rc = spx_pack_upper(ctx, q->st, nb, nw, k ? q->E : q->X, nb, q->ws);
    if (rc) return rc;
    SPX_HIP(ctx, hipMemcpyAsync(dst, q->ws, bytes, hipMemcpyDeviceToHost, q->st));
    SPX_HIP(ctx, hipStreamSynchronize(q->st));
  }
  return SPEXHIP_OK;
}

int spexhip_qpipe_buffers(spexhip_ctx *ctx, spexhip_qpipe *q, void **M_dev, void **chi_dev,
                          void **einv_dev) {
  if (!ctx || !q) return SPEXHIP_E_INVALID;
  if (M_dev) *M_dev = q->M;
  if (chi_dev) *chi_dev = q->X;
  if (einv_dev) *einv_dev = q->E;
  return SPEXHIP_OK;
}

int spexhip_qpipe_destroy(spexhip_ctx *ctx, spexhip_qpipe *q) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (q) destroy(ctx, q);
  return SPEXHIP_OK;
}

}  // extern "C"
