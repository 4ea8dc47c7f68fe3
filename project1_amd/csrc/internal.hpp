// Internal launchers shared between translation units.  Every function here
// enqueues work on the given stream only (no host synchronisation, no
// memset nodes), so the per-q pipeline can be captured into a HIP graph.
#pragma once
#include "plan.hpp"

// weights (HILBERT NONE closed form)
int spx_weights_launch(spexhip_ctx *ctx, hipStream_t st, int nk, int nocc, int nunocc,
                       const void *e_occ, const void *e_unocc, int nkpt, int nsym1,
                       const void *nkpts, int nw, const void *omega, void *wght);
// grouped chi0 (persistent, many problems)
int spx_group_launch(spexhip_ctx *ctx, hipStream_t st, const spexhip_group *g);
// chi0 Ozaki form
size_t spx_ozaki_ws_bytes(long n, int nb, int nw, int nwg);
int spx_chi0_ozaki_async(spexhip_ctx *ctx, hipStream_t st, long n, int nb, const void *M,
                         long ldm, int colmajor, int nw, const void *wght, double alpha,
                         double beta, void *chi, int ldchi, void *ws, size_t ws_bytes,
                         int *err_dev);
// products
int spx_wfprod_mt(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cmt,
                  int ncent, int nband, int nkpt, const double *q, int nk, const int *k1_dev,
                  const int *k2_dev, int b1lo, int nocc, int b2lo, int nunocc, void *M, long ldm,
                  int col0, void *ws);
size_t spx_wfprod_mt_ws(const spexhip_plan *p, int nk, int nocc, int nunocc);
int spx_pw_oidx(spexhip_ctx *ctx, const spexhip_plan *p, int nk, const int *g2, int *oidx);
size_t spx_wfprod_pw_ws_core(const spexhip_plan *p, int nocc, int nunocc);
int spx_wfprod_pw_async(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cpw,
                        int nband, int nk, const int *k1, const int *k2, const int *oidx_dev,
                        int b1lo, int nocc, int b2lo, int nunocc, void *M, long ldm, int col0,
                        void *ws);
// dielectric step and inverse
int spx_dielectric(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, void *X, long ldx, int vkind,
                   const void *vh, void *E, long lde, void *tmp);
size_t spx_inverse_ws(int nb, int nw);
// general != 0: complex pivots (a general matrix, no pivoting); logdet_parts
// only for the HPD form
int spx_inverse(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, void *X, long ldx, void *ws,
                int *flag, double *logdet_parts, int general = 0);
int spx_screen(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, const void *Einv, long lde,
               int vkind, const void *vh, void *W, long ldw, void *tmp);
// packing and fills
int spx_pack_upper(spexhip_ctx *ctx, hipStream_t st, int nb, int nw, const void *X, long ldx,
                   void *P);
int spx_fill_zero(spexhip_ctx *ctx, hipStream_t st, void *p, size_t bytes);
