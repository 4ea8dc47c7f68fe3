// Muffin-tin part of the wave-function products, wavefproducts.f:95-431.
//
// Reference (SUBTYPE 2, noinv, no SOC): per atom ic, k and occupied b1,
//   mat(lm1+n1, llm+nn) += sum_n cmt1(lm+n) integral(ln+n, ln1+n1, lln+nn)
//                          * gaunt(ll,l1,l,mm,m1,m)               (:294-339)
//   cprod(b2,b1,k,llm0+I) += cexp * conjg(cmt2(:,b2)^H mat(:,I))     (:359)
// with cexp = exp(-2 pi i q.R_ic) (:252).  Since the Gaunt/radial factor
// is a fixed real tensor T(lmi, lmo, I) per atom type, the product is
//   cprod(b2, I) = cexp * sum_lmo cmt2(lmo,b2) * Y_{k,b1}(lmo, I),
//   Y_{k,b1}(lmo, I) = sum_lmi conj(cmt1(lmi)) T(lmi, lmo, I),
// i.e. two FP64 MFMA GEMMs: stage A (all (k,b1) at once against T) and a
// batched stage B (one 147x400x147-class complex GEMM per (k,b1)).  T is
// built once per plan on the host (Gaunt coefficients by Racah's formula
// and the caller's radial-integral tensor).
#include <cmath>
#include <vector>

#include "gemm_f64.hpp"
#include "plan.hpp"

using namespace spxg;

namespace {

// Racah formula for the Wigner 3j symbol (the textbook closed form; the
// reference evaluates the same formula, numerics.f:666-700).
double lfac(int n) { return std::lgamma((double)n + 1.0); }

double wigner3j(int l1, int l2, int l3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0.0;
  if (std::abs(m1) > l1 || std::abs(m2) > l2 || std::abs(m3) > l3) return 0.0;
  if (l3 < std::abs(l1 - l2) || l3 > l1 + l2) return 0.0;
  const int f1 = l3 - l2 + m1, f2 = l3 - l1 - m2, f3 = l1 + l2 - l3, f4 = l1 - m1, f5 = l2 + m2;
  const int tmin = std::max(0, std::max(-f1, -f2)), tmax = std::min(f3, std::min(f4, f5));
  // sqrt of the triangle and m factorials, in log form for stability
  const double lpre =
      0.5 * (lfac(l1 + l2 - l3) + lfac(l1 - l2 + l3) + lfac(-l1 + l2 + l3) - lfac(l1 + l2 + l3 + 1) +
             lfac(l1 + m1) + lfac(l1 - m1) + lfac(l2 + m2) + lfac(l2 - m2) + lfac(l3 + m3) +
             lfac(l3 - m3));
  double s = 0.0;
  for (int t = tmin; t <= tmax; t++) {
    const double term =
        std::exp(lpre - (lfac(t) + lfac(f1 + t) + lfac(f2 + t) + lfac(f3 - t) + lfac(f4 - t) + lfac(f5 - t)));
    s += (t & 1) ? -term : term;
  }
  const int e = l1 - l2 - m3;
  return ((e % 2 + 2) % 2) ? -s : s;
}

// integral of conj(Y_l1m1) Y_l2m2 conj(Y_l3m3)
double gaunt(int l1, int l2, int l3, int m1, int m2, int m3) {
  if (m3 != m2 - m1) return 0.0;
  if (std::abs(m1) > l1 || std::abs(m2) > l2 || std::abs(m3) > l3) return 0.0;
  if ((l1 + l2 + l3) % 2) return 0.0;
  if (l3 < std::abs(l1 - l2) || l3 > l1 + l2) return 0.0;
  const double sgn = ((m1 + m3) % 2 + 2) % 2 ? -1.0 : 1.0;
  return sgn * std::sqrt((2.0 * l1 + 1) * (2.0 * l2 + 1) * (2.0 * l3 + 1) / (4.0 * M_PI)) *
         wigner3j(l1, l2, l3, -m1, m2, -m3) * wigner3j(l1, l2, l3, 0, 0, 0);
}

__global__ void gather_cmt_kernel(const double2 *__restrict__ cmt, int mlx, int ncent, int nband,
                                  int ic, int maxlm, const int *__restrict__ kidx, int bandlo,
                                  int nbnd, int nk, int conj_, double2 *__restrict__ out) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = (long)nk * nbnd * maxlm;
  if (g >= tot) return;
  const int lm = (int)(g % maxlm);
  const long r = g / maxlm;
  const int b = (int)(r % nbnd), ik = (int)(r / nbnd);
  double2 v = cmt[lm + (long)mlx * (ic + (long)ncent * ((bandlo + b) + (long)nband * kidx[ik]))];
  if (conj_) v.y = -v.y;
  out[g] = v;
}

}  // namespace

int spx_mt_plan_build(spexhip_ctx *ctx, int ntype, const spexhip_mt_type *types,
                      const double *cent, int maxlmindx, spexhip_plan *p) {
  p->kind = 1;
  p->maxlmindx = maxlmindx;
  p->nbasp = 0;
  int ncent = 0;
  for (int it = 0; it < ntype; it++) {
    const spexhip_mt_type &T = types[it];
    SPX_CHECK(ctx, T.lcut >= 0 && T.lcut < 32 && T.lcutm >= 0 && T.lcutm < 32 && T.neq > 0 &&
                       T.nindx && T.nindxm && T.integral,
              "mt plan: bad type %d", it);
    PlanMtType d;
    d.neq = T.neq;
    d.maxlm = 0;
    d.maxllm = 0;
    for (int l = 0; l <= T.lcut; l++) d.maxlm += (2 * l + 1) * T.nindx[l];
    for (int l = 0; l <= T.lcutm; l++) d.maxllm += (2 * l + 1) * T.nindxm[l];
    SPX_CHECK(ctx, d.maxlm <= maxlmindx, "mt plan: type %d needs %d > maxlmindx=%d", it, d.maxlm, maxlmindx);
    const long LD = (long)d.maxlm * d.maxllm;
    std::vector<double2> Tc((size_t)LD * d.maxlm, make_double2(0.0, 0.0));
    // lmstart(m,l), wavefproducts.f:225-232
    std::vector<int> lmstart((T.lcut + 1) * (2 * T.lcut + 1));
    {
      int lm = 0;
      for (int l = 0; l <= T.lcut; l++)
        for (int m = -l; m <= l; m++) {
          lmstart[l * (2 * T.lcut + 1) + m + T.lcut] = lm;
          lm += T.nindx[l];
        }
    }
    auto integ = [&](int a, int b, int c) -> double {
      return T.integral[a + (long)T.d1 * (b + (long)T.d2 * c)];
    };
    // index bookkeeping of :294-339
    int llm = 0, lln = 0;
    for (int ll = 0; ll <= T.lcutm; ll++) {
      for (int mm = -ll; mm <= ll; mm++) {
        int lm = 0, ln = 0;
        for (int l = 0; l <= T.lcut; l++) {
          const int nx = T.nindx[l];
          for (int m = -l; m <= l; m++) {
            const int m1 = m + mm;
            int ln1 = 0;
            for (int l1 = std::abs(l - ll); l1 <= std::min(l + ll, T.lcut); l1 += 2) {
              const int nx1 = T.nindx[l1];
              if (l1 >= std::abs(m1)) {
                const double gnt = gaunt(ll, l1, l, mm, m1, m);
                if (gnt != 0.0) {
                  const int lm1 = lmstart[l1 * (2 * T.lcut + 1) + m1 + T.lcut];
                  for (int nn = 0; nn < T.nindxm[ll]; nn++)
                    for (int n1 = 0; n1 < nx1; n1++)
                      for (int nq = 0; nq < nx; nq++) {
                        SPX_CHECK(ctx, ln + nq < T.d1 && ln1 + n1 < T.d2 && lln + nn < T.d3,
                                  "mt plan: integral tensor of type %d too small", it);
                        const long lmoI = (llm + nn) + (long)d.maxllm * (lm1 + n1);
                        Tc[lmoI + LD * (lm + nq)].x += integ(ln + nq, ln1 + n1, lln + nn) * gnt;
                      }
                }
              }
              ln1 += nx1;
            }
            lm += nx;
          }
          ln += nx;
        }
        llm += T.nindxm[ll];
      }
      lln += T.nindxm[ll];
    }
    int rc = spx_alloc(ctx, sizeof(double2) * Tc.size(), (void **)&d.T);
    if (rc) return rc;
    p->owned.push_back(d.T);
    SPX_HIP(ctx, hipMemcpy(d.T, Tc.data(), sizeof(double2) * Tc.size(), hipMemcpyHostToDevice));
    d.llm0 = p->nbasp;
    d.ic0 = ncent;
    p->nbasp += d.neq * d.maxllm;
    ncent += d.neq;
    p->mt.push_back(d);
  }
  p->ncent = ncent;
  p->cent.assign(cent, cent + 3 * ncent);
  return SPEXHIP_OK;
}

// Device MT products into columns col0 .. col0+nbasp-1 of M (t*ldm + I).
int spx_wfprod_mt(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cmt,
                  int ncent, int nband, int nkpt, const double *q, int nk, const int *k1_dev,
                  const int *k2_dev, int b1lo, int nocc, int b2lo, int nunocc, void *M, long ldm,
                  int col0, void *ws) {
  const int nkb = nk * nocc;
  int maxlm = 0;
  long maxLD = 0;
  for (auto &d : p->mt) {
    maxlm = std::max(maxlm, d.maxlm);
    maxLD = std::max(maxLD, (long)d.maxlm * d.maxllm);
  }
  double2 *G1 = (double2 *)ws;
  double2 *G2 = G1 + (long)nkb * maxlm;
  double2 *Y = G2 + (long)nk * nunocc * maxlm;
  for (auto &d : p->mt) {
    const long LD = (long)d.maxlm * d.maxllm;
    for (int e = 0; e < d.neq; e++) {
      const int ic = d.ic0 + e;
      const double ph = q[0] * p->cent[3 * ic] + q[1] * p->cent[3 * ic + 1] + q[2] * p->cent[3 * ic + 2];
      const double2 cexp = make_double2(std::cos(2 * M_PI * ph), -std::sin(2 * M_PI * ph));
      long tot = (long)nkb * d.maxlm;
      hipLaunchKernelGGL(gather_cmt_kernel, dim3((tot + 255) / 256), dim3(256), 0, st,
                         (const double2 *)cmt, p->maxlmindx, ncent, nband, ic, d.maxlm, k1_dev,
                         b1lo, nocc, nk, 1, G1);
      SPX_LAUNCHED(ctx);
      tot = (long)nk * nunocc * d.maxlm;
      hipLaunchKernelGGL(gather_cmt_kernel, dim3((tot + 255) / 256), dim3(256), 0, st,
                         (const double2 *)cmt, p->maxlmindx, ncent, nband, ic, d.maxlm, k2_dev,
                         b2lo, nunocc, nk, 0, G2);
      SPX_LAUNCHED(ctx);
      // stage A: Y(lmoI, kb) = Tc(lmoI, lmi) * G1(lmi, kb)
      ZgemmArgs a{(int)LD, nkb, d.maxlm, make_double2(1.0, 0.0), d.T, LD, 0, 1, G1, d.maxlm, 0, 1,
                  0.0, Y, LD, 0, 1, 0.0};
      int rc = spx_zgemm_ex(ctx, st, a, 1);
      if (rc) return rc;
      // stage B: M(I, t0+b2) = cexp * Y_kb(I, lmo) * G2_k(lmo, b2)
      ZgemmArgs b{d.maxllm, nunocc, d.maxlm, cexp,
                  Y, d.maxllm, LD, 1,
                  G2, d.maxlm, (long)nunocc * d.maxlm, nocc,
                  0.0, (double2 *)M + col0 + d.llm0 + (long)e * d.maxllm, ldm, (long)nunocc * ldm, 1,
                  0.0};
      rc = spx_zgemm_ex(ctx, st, b, nkb);
      if (rc) return rc;
    }
  }
  (void)maxLD;
  return SPEXHIP_OK;
}

size_t spx_wfprod_mt_ws(const spexhip_plan *p, int nk, int nocc, int nunocc) {
  int maxlm = 0;
  long maxLD = 0;
  for (auto &d : p->mt) {
    maxlm = std::max(maxlm, d.maxlm);
    maxLD = std::max(maxLD, (long)d.maxlm * d.maxllm);
  }
  return sizeof(double2) * ((size_t)nk * nocc * maxlm + (size_t)nk * nunocc * maxlm +
                            (size_t)maxLD * nk * nocc);
}
