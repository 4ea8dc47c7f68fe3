// C ABI of the wave-function-product kernels (plans, device and host
// drop-in entry points).
#include <hip/hip_runtime.h>

#include <vector>

#include "plan.hpp"

int spx_mt_plan_build(spexhip_ctx *ctx, int ntype, const spexhip_mt_type *types,
                      const double *cent, int maxlmindx, spexhip_plan *p);
int spx_pw_plan_build(spexhip_ctx *ctx, const spexhip_pw_desc *d, spexhip_plan *p);
int spx_wfprod_mt(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cmt,
                  int ncent, int nband, int nkpt, const double *q, int nk, const int *k1_dev,
                  const int *k2_dev, int b1lo, int nocc, int b2lo, int nunocc, void *M, long ldm,
                  int col0, void *ws);
size_t spx_wfprod_mt_ws(const spexhip_plan *p, int nk, int nocc, int nunocc);
int spx_wfprod_pw(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cpw,
                  int nband, int nk, const int *k1, const int *k2, const int *g2, int b1lo,
                  int nocc, int b2lo, int nunocc, void *M, long ldm, int col0, void *ws);
size_t spx_wfprod_pw_ws(const spexhip_plan *p, int nk, int nocc, int nunocc);
int spx_transpose_d2h(spexhip_ctx *ctx, hipStream_t st, const void *M, long rows, int cols,
                      long ldm, double *host_colmajor);

static void plan_free(spexhip_ctx *ctx, spexhip_plan *p) {
  for (void *d : p->owned) spx_free(ctx, d);
  delete p;
}

static int check_packet(spexhip_ctx *ctx, int nk, const int *k1, const int *k2, int nkpt, int b1lo,
                        int nocc, int b2lo, int nunocc, int nband) {
  SPX_CHECK(ctx, nk >= 0 && nocc >= 0 && nunocc >= 0, "wfprod: negative packet size");
  SPX_CHECK(ctx, b1lo >= 0 && b2lo >= 0 && b1lo + nocc <= nband && b2lo + nunocc <= nband,
            "wfprod: band range outside 0..nband-1");
  for (int i = 0; i < nk; i++)
    SPX_CHECK(ctx, k1[i] >= 0 && k1[i] < nkpt && k2[i] >= 0 && k2[i] < nkpt,
              "wfprod: k index %d out of range", i);
  return SPEXHIP_OK;
}

extern "C" {

int spexhip_mt_plan_create(spexhip_ctx *ctx, int ntype, const spexhip_mt_type *types,
                           const double *cent, int maxlmindx, spexhip_plan **plan) {
  if (!ctx || !plan) return SPEXHIP_E_INVALID;
  *plan = nullptr;
  SPX_CHECK(ctx, ntype > 0 && types && cent && maxlmindx > 0, "mt plan: bad arguments");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  spexhip_plan *p = new spexhip_plan();
  int rc = spx_mt_plan_build(ctx, ntype, types, cent, maxlmindx, p);
  if (rc) {
    plan_free(ctx, p);
    return rc;
  }
  *plan = p;
  return SPEXHIP_OK;
}

int spexhip_pw_plan_create(spexhip_ctx *ctx, const spexhip_pw_desc *desc, spexhip_plan **plan) {
  if (!ctx || !plan) return SPEXHIP_E_INVALID;
  *plan = nullptr;
  SPX_CHECK(ctx, desc, "pw plan: null descriptor");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  spexhip_plan *p = new spexhip_plan();
  int rc = spx_pw_plan_build(ctx, desc, p);
  if (rc) {
    plan_free(ctx, p);
    return rc;
  }
  *plan = p;
  return SPEXHIP_OK;
}

int spexhip_plan_destroy(spexhip_ctx *ctx, spexhip_plan *plan) {
  if (!ctx) return SPEXHIP_E_INVALID;
  if (!plan) return SPEXHIP_OK;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  plan_free(ctx, plan);
  return SPEXHIP_OK;
}

int spexhip_plan_ncols(spexhip_ctx *ctx, const spexhip_plan *plan, int *ncols) {
  if (!ctx || !plan || !ncols) return SPEXHIP_E_INVALID;
  *ncols = plan->kind == 1 ? plan->nbasp : plan->ngptm;
  return SPEXHIP_OK;
}

int spexhip_pw_plan_grid(spexhip_ctx *ctx, const spexhip_plan *plan, int *nfft) {
  if (!ctx || !plan || !nfft) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plan->kind == 2, "pw_plan_grid: not an interstitial plan");
  for (int d = 0; d < 3; d++) nfft[d] = plan->nfft[d];
  return SPEXHIP_OK;
}

int spexhip_wfprod2_mt_dev(spexhip_ctx *ctx, const spexhip_plan *plan, const void *cmt_dev,
                           int ncent, int nband, int nkpt, const double *q, int nk, const int *k1,
                           const int *k2, int b1lo, int nocc, int b2lo, int nunocc, void *M_dev,
                           long ldm, int col0) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plan && plan->kind == 1, "wfprod_mt: not a muffin-tin plan");
  SPX_CHECK(ctx, ncent >= plan->ncent, "wfprod_mt: ncent=%d but the plan has %d centres", ncent, plan->ncent);
  SPX_CHECK(ctx, q && k1 && k2, "wfprod_mt: null argument");
  int rc = check_packet(ctx, nk, k1, k2, nkpt, b1lo, nocc, b2lo, nunocc, nband);
  if (rc) return rc;
  const long n = (long)nk * nocc * nunocc;
  if (n == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, col0 >= 0 && ldm >= col0 + plan->nbasp, "wfprod_mt: ldm too small for the MT columns");
  SPX_CHECK(ctx, spx_owns(ctx, M_dev, 16 * ((size_t)ldm * (n - 1) + col0 + plan->nbasp)),
            "wfprod_mt: M is not a context allocation of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, cmt_dev, 16 * (size_t)plan->maxlmindx * ncent * nband * nkpt),
            "wfprod_mt: cmt is not a context allocation of maxlmindx*ncent*nband*nkpt complex");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t wsb = spx_wfprod_mt_ws(plan, nk, nocc, nunocc);
  void *ws;
  rc = spx_scratch(ctx, wsb + 8 * (size_t)nk + 256, &ws);
  if (rc) return rc;
  int *kd = (int *)((char *)ws + wsb);
  std::vector<int> kk(2 * nk);
  for (int i = 0; i < nk; i++) {
    kk[i] = k1[i];
    kk[nk + i] = k2[i];
  }
  SPX_HIP(ctx, hipMemcpyAsync(kd, kk.data(), 8 * (size_t)nk, hipMemcpyHostToDevice, ctx->stream));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return spx_wfprod_mt(ctx, ctx->stream, plan, cmt_dev, ncent, nband, nkpt, q, nk, kd, kd + nk,
                       b1lo, nocc, b2lo, nunocc, M_dev, ldm, col0, ws);
}

int spexhip_wfprod2_pw_dev(spexhip_ctx *ctx, const spexhip_plan *plan, const void *cpw_dev,
                           int nband, int nk, const int *k1, const int *k2, const int *g2,
                           int b1lo, int nocc, int b2lo, int nunocc, void *M_dev, long ldm,
                           int col0) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plan && plan->kind == 2, "wfprod_pw: not a plane-wave plan");
  SPX_CHECK(ctx, k1 && k2 && g2, "wfprod_pw: null argument");
  int rc = check_packet(ctx, nk, k1, k2, plan->nkpt, b1lo, nocc, b2lo, nunocc, nband);
  if (rc) return rc;
  const long n = (long)nk * nocc * nunocc;
  if (n == 0 || plan->ngptm == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, col0 >= 0 && ldm >= col0 + plan->ngptm, "wfprod_pw: ldm too small for the IPW columns");
  SPX_CHECK(ctx, spx_owns(ctx, M_dev, 16 * ((size_t)ldm * (n - 1) + col0 + plan->ngptm)),
            "wfprod_pw: M is not a context allocation of the stated shape");
  SPX_CHECK(ctx, spx_owns(ctx, cpw_dev, 16 * (size_t)plan->maxgpt * nband * plan->nkpt),
            "wfprod_pw: cpw is not a context allocation of maxgpt*nband*nkpt complex");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *ws;
  rc = spx_scratch(ctx, spx_wfprod_pw_ws(plan, nk, nocc, nunocc), &ws);
  if (rc) return rc;
  return spx_wfprod_pw(ctx, ctx->stream, plan, cpw_dev, nband, nk, k1, k2, g2, b1lo, nocc, b2lo,
                       nunocc, M_dev, ldm, col0, ws);
}

int spexhip_wavefproducts2_mt(spexhip_ctx *ctx, const spexhip_plan *plan, const double *cmt,
                              int ncent, int nband, int nkpt, const double *q, int nk,
                              const int *k1, const int *k2, int b1lo, int nocc, int b2lo,
                              int nunocc, double *cprod) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plan && plan->kind == 1 && cmt && cprod, "wavefproducts2_mt: bad arguments");
  const long n = (long)nk * nocc * nunocc;
  if (n == 0) return SPEXHIP_OK;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t cb = 16 * (size_t)plan->maxlmindx * ncent * nband * nkpt;
  void *dC = nullptr, *dM = nullptr;
  int rc = spx_alloc(ctx, cb, &dC);
  if (!rc) rc = spx_alloc(ctx, 16 * (size_t)n * plan->nbasp, &dM);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(dC, cmt, cb, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "upload: %s", hipGetErrorString(e));
  }
  if (!rc)
    rc = spexhip_wfprod2_mt_dev(ctx, plan, dC, ncent, nband, nkpt, q, nk, k1, k2, b1lo, nocc, b2lo,
                                nunocc, dM, plan->nbasp, 0);
  if (!rc) rc = spx_transpose_d2h(ctx, ctx->stream, dM, n, plan->nbasp, plan->nbasp, cprod);
  (void)hipStreamSynchronize(ctx->stream);
  spx_free(ctx, dC);
  spx_free(ctx, dM);
  return rc;
}

int spexhip_wavefproducts2_pw(spexhip_ctx *ctx, const spexhip_plan *plan, const double *cpw,
                              int nband, int nk, const int *k1, const int *k2, const int *g2,
                              int b1lo, int nocc, int b2lo, int nunocc, double *cprod) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plan && plan->kind == 2 && cpw && cprod, "wavefproducts2_pw: bad arguments");
  const long n = (long)nk * nocc * nunocc;
  if (n == 0 || plan->ngptm == 0) return SPEXHIP_OK;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t cb = 16 * (size_t)plan->maxgpt * nband * plan->nkpt;
  void *dC = nullptr, *dM = nullptr;
  int rc = spx_alloc(ctx, cb, &dC);
  if (!rc) rc = spx_alloc(ctx, 16 * (size_t)n * plan->ngptm, &dM);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(dC, cpw, cb, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "upload: %s", hipGetErrorString(e));
  }
  if (!rc)
    rc = spexhip_wfprod2_pw_dev(ctx, plan, dC, nband, nk, k1, k2, g2, b1lo, nocc, b2lo, nunocc, dM,
                                plan->ngptm, 0);
  if (!rc) rc = spx_transpose_d2h(ctx, ctx->stream, dM, n, plan->ngptm, plan->ngptm, cprod);
  (void)hipStreamSynchronize(ctx->stream);
  spx_free(ctx, dC);
  spx_free(ctx, dM);
  return rc;
}

}  // extern "C"
