// Continuous-batching request loop for chi0 / eps^-1 jobs.
//
// No reference counterpart: SPEX runs one q at a time per MPI rank (its
// MPIKPT task queue, correlation.f:316-402 / Mutil.f:494-548, hands q-points
// to idle ranks).  Here one process serves a stream of q-point requests on
// one GPU:
//   * submit() copies the request and queues it (never blocks on the GPU);
//   * a scheduler thread admits queued requests in arrival order whenever a
//     slot is free and their HBM footprint (spexhip_qpipe_footprint) fits the
//     server budget -- the MEM test of susceptibility.f:674-713, applied per
//     request -- creates their pipeline (own stream, all buffers) and
//     launches it, so admitted requests run concurrently on the device;
//   * completions are detected with stream events; the slot is freed at once
//     and the memory when the client releases the request.
// Every request computes exactly what it would compute alone: pipelines
// share no buffers, and each kernel's reduction order does not depend on
// what else runs.
#include <algorithm>
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
  long id = 0;
  spexhip_qdesc d{};
  std::vector<int> k1, k2, g2;
  size_t footprint = 0;
  int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
  int status = SPEXHIP_OK;
  std::string err;
  spexhip_qpipe *qp = nullptr;
  Clock::time_point t_submit, t_start, t_done;
  float device_ms = 0;
};

}  // namespace

struct spexhip_server {
  spexhip_ctx *ctx = nullptr;
  int max_inflight = 1;
  size_t budget = 0, used = 0;
  int backfill = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<long> queue;
  std::vector<long> running;
  std::map<long, std::unique_ptr<Req>> reqs;
  long next_id = 1;
  std::atomic<bool> stop{false};
  std::thread th;
  long completed = 0, admitted = 0;
  size_t peak_used = 0;
  int peak_inflight = 0;
};

static void scheduler(spexhip_server *s) {
  (void)hipSetDevice(s->ctx->device);
  while (!s->stop.load()) {
    bool idle = true;
    std::unique_lock<std::mutex> lk(s->mu);
    // completions
    for (size_t i = 0; i < s->running.size();) {
      Req *r = s->reqs[s->running[i]].get();
      const int q = spexhip_qpipe_query(s->ctx, r->qp);
      if (q == SPEXHIP_E_AGAIN) {
        i++;
        continue;
      }
      float ms = 0;
      const int rc = spexhip_qpipe_wait(s->ctx, r->qp, &ms);
      r->t_done = Clock::now();
      r->device_ms = ms;
      r->status = rc;
      r->state = rc ? 3 : 2;
      if (rc) {
        char buf[512];
        spexhip_last_error(s->ctx, buf, sizeof buf);
        r->err = buf;
      }
      s->running.erase(s->running.begin() + i);
      s->completed++;
      idle = false;
      s->cv.notify_all();
    }
    // admission in arrival order (optionally backfilling smaller requests)
    for (auto it = s->queue.begin(); it != s->queue.end();) {
      if ((int)s->running.size() >= s->max_inflight) break;
      Req *r = s->reqs[*it].get();
      if (s->used + r->footprint > s->budget) {
        if (!s->backfill) break;
        ++it;
        continue;
      }
      const long id = *it;
      it = s->queue.erase(it);
      s->used += r->footprint;
      s->peak_used = std::max(s->peak_used, s->used);
      lk.unlock();  // pipeline creation allocates and uploads; other clients may submit
      int rc = spexhip_qpipe_create(s->ctx, &r->d, &r->qp);
      if (!rc) rc = spexhip_qpipe_launch(s->ctx, r->qp, 0);
      lk.lock();
      r->t_start = Clock::now();
      if (rc) {
        char buf[512];
        spexhip_last_error(s->ctx, buf, sizeof buf);
        r->err = buf;
        r->status = rc;
        r->state = 3;
        r->t_done = r->t_start;
        if (r->qp) {
          spexhip_qpipe_destroy(s->ctx, r->qp);
          r->qp = nullptr;
        }
        s->used -= r->footprint;
        s->cv.notify_all();
      } else {
        r->state = 1;
        s->running.push_back(id);
        s->admitted++;
        s->peak_inflight = std::max(s->peak_inflight, (int)s->running.size());
      }
      idle = false;
      it = s->queue.begin();  // re-scan from the head after the unlock window
    }
    if (idle) s->cv.wait_for(lk, std::chrono::microseconds(100));
  }
}

extern "C" {

int spexhip_server_create(spexhip_ctx *ctx, int max_inflight, size_t hbm_budget, int backfill,
                          spexhip_server **out) {
  if (!ctx || !out || max_inflight < 1) return SPEXHIP_E_INVALID;
  spexhip_server *s = new spexhip_server();
  s->ctx = ctx;
  s->max_inflight = max_inflight;
  s->budget = hbm_budget;
  s->backfill = backfill;
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->servers.push_back(s);
  }
  s->th = std::thread(scheduler, s);
  *out = s;
  return SPEXHIP_OK;
}

int spexhip_server_submit(spexhip_server *s, const spexhip_qdesc *d, long *id) {
  if (!s || !d || !id) return SPEXHIP_E_INVALID;
  auto r = std::make_unique<Req>();
  r->d = *d;
  if (!d->M_dev) {
    if (d->k1) r->k1.assign(d->k1, d->k1 + d->nk);
    if (d->k2) r->k2.assign(d->k2, d->k2 + d->nk);
    if (d->g2) r->g2.assign(d->g2, d->g2 + 3 * d->nk);
    r->d.k1 = r->k1.empty() ? nullptr : r->k1.data();
    r->d.k2 = r->k2.empty() ? nullptr : r->k2.data();
    r->d.g2 = r->g2.empty() ? nullptr : r->g2.data();
  }
  spexhip_qpipe_footprint(&r->d, &r->footprint);
  if (r->footprint > s->budget)
    return spx_seterr(s->ctx, SPEXHIP_E_NOMEM,
                      "server: request footprint %zu exceeds the server budget %zu", r->footprint,
                      s->budget);
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> g(s->mu);
  r->id = s->next_id++;
  *id = r->id;
  s->queue.push_back(r->id);
  s->reqs[r->id] = std::move(r);
  s->cv.notify_all();
  return SPEXHIP_OK;
}

int spexhip_server_poll(spexhip_server *s, long id, int *state) {
  if (!s || !state) return SPEXHIP_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return SPEXHIP_E_INVALID;
  *state = it->second->state;
  return SPEXHIP_OK;
}

// Block until the request finished.  latency_ms: submit -> completion;
// queue_ms: submit -> launch; device_ms: the pipeline's own device time.
int spexhip_server_wait(spexhip_server *s, long id, double *latency_ms, double *queue_ms,
                        double *device_ms) {
  if (!s) return SPEXHIP_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return SPEXHIP_E_INVALID;
  Req *r = it->second.get();
  s->cv.wait(lk, [&] { return r->state >= 2; });
  using ms = std::chrono::duration<double, std::milli>;
  if (latency_ms) *latency_ms = ms(r->t_done - r->t_submit).count();
  if (queue_ms) *queue_ms = ms(r->t_start - r->t_submit).count();
  if (device_ms) *device_ms = r->device_ms;
  if (r->status) return spx_seterr(s->ctx, r->status, "server request %ld: %s", id, r->err.c_str());
  return SPEXHIP_OK;
}

int spexhip_server_result(spexhip_server *s, long id, double *chi_packed, double *einv_packed) {
  if (!s) return SPEXHIP_E_INVALID;
  spexhip_qpipe *qp = nullptr;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->state != 2) return SPEXHIP_E_INVALID;
    qp = it->second->qp;
  }
  return spexhip_qpipe_result(s->ctx, qp, chi_packed, einv_packed);
}

// Free the request's device memory (its HBM returns to the admission budget).
int spexhip_server_release(spexhip_server *s, long id) {
  if (!s) return SPEXHIP_E_INVALID;
  std::unique_ptr<Req> r;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->state < 2) return SPEXHIP_E_INVALID;
    r = std::move(it->second);
    s->reqs.erase(it);
  }
  if (r->qp) spexhip_qpipe_destroy(s->ctx, r->qp);
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (r->qp) s->used -= r->footprint;  // creation failures were returned at admission
    s->cv.notify_all();
  }
  return SPEXHIP_OK;
}

int spexhip_server_stats(spexhip_server *s, long *admitted, long *completed, size_t *peak_used,
                         int *peak_inflight) {
  if (!s) return SPEXHIP_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  if (admitted) *admitted = s->admitted;
  if (completed) *completed = s->completed;
  if (peak_used) *peak_used = s->peak_used;
  if (peak_inflight) *peak_inflight = s->peak_inflight;
  return SPEXHIP_OK;
}

int spexhip_server_destroy(spexhip_server *s) {
  if (!s) return SPEXHIP_OK;
  {
    std::lock_guard<std::mutex> g(s->ctx->mu);
    auto &v = s->ctx->servers;
    v.erase(std::remove(v.begin(), v.end(), s), v.end());
  }
  s->stop.store(true);
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();
  for (auto &kv : s->reqs)
    if (kv.second->qp) spexhip_qpipe_destroy(s->ctx, kv.second->qp);
  delete s;
  return SPEXHIP_OK;
}

}  // extern "C"
