// Gaussian Brillouin-zone weights of the Hilbert mesh: the wtetra of the
// HILBERT pair sum when the GAUSS key is set (susceptibility.f:941-942).
//
// Reference: gauss4_init (gauss.f:120-186).  Occupation numbers are smeared
// with a Gaussian of width sigma1, f(e) = (1 + erf((efermi - e) / (sqrt2
// sigma1))) / 2, the transition energy d = e_unocc(k+q) - e_occ(k) with a
// Gaussian of width sigma2, and bin i of the exponential mesh, [w(i) -
// df(i), w(i) + df(i)], takes the integral of that Gaussian over the bin
// divided by the bin width:
//   w(t, i) += f_occ (1 - f_unocc) g2fac factor (erf(gfac (w + df - d)) -
//              erf(gfac (w - df - d))) / (2 df),
// summed over the mesh points k of each parent row in ascending order and
// divided by nkpt at the end.  Bins at w = 0 stay 0.  One thread per output
// weight (row, b1, b2) and bin; it walks the parent's members itself, so
// the sums need no atomics and keep the reference's order.
#include "common.hpp"
#include "internal.hpp"

#include <cmath>
#include <vector>

namespace {

struct GaussBin {
  double f1, f2, df2;  // bin edges and 2 df; df2 = 0 marks a bin at w = 0
};

__global__ void gauss_weights_kernel(int nrow, int nocc, int nunocc, const int *__restrict__ off,
                                     const int *__restrict__ mem, const double *__restrict__ eo,
                                     const double *__restrict__ eu, double efermi, double s1,
                                     double gfac, double g2fac, double factor, double nkpt,
                                     int nfrqs, const GaussBin *__restrict__ bins,
                                     double *__restrict__ out, long ldo) {
  const long per = (long)nocc * nunocc;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (t >= (long)nrow * per || i >= nfrqs) return;
  const int p = (int)(t / per);
  const int b1 = (int)((t - (long)p * per) / nunocc);
  const int b2 = (int)(t - (long)p * per - (long)b1 * nunocc);
  const GaussBin bin = bins[i];
  double w = 0.0;
  if (bin.df2 != 0.0) {
    const double sq2s1 = sqrt(2.0) * s1;
    for (int m = off[p]; m < off[p + 1]; m++) {
      const long k = mem[m];
      const double e1 = eo[k * nocc + b1], e2 = eu[k * nunocc + b2];
      const double r1 = (1.0 + erf((efermi - e1) / sq2s1)) / 2;
      const double r2 = (1.0 + erf((e2 - efermi) / sq2s1)) / 2;
      const double dene = e2 - e1;
      w = w + r1 * r2 * g2fac * factor * (erf(gfac * (bin.f2 - dene)) - erf(gfac * (bin.f1 - dene))) /
                  bin.df2;
    }
  }
  out[t + ldo * i] = w / nkpt;
}

}  // namespace

extern "C" {

int spexhip_gauss_weights_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc,
                              const void *e_occ_dev, const void *e_unocc_dev, double efermi,
                              double sigma1, double sigma2, int nfrqs, const double *frqs,
                              const int *kpt1p, int nkpt1, void *wtetra_dev, long ldwt) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nk > 0 && nocc > 0 && nunocc > 0 && frqs, "gauss_weights: bad arguments");
  SPX_CHECK(ctx, sigma1 > 0 && sigma2 > 0, "gauss_weights: the GAUSS widths must be positive");
  // the mesh factor a of gauss.f:137-147 needs three points
  SPX_CHECK(ctx, nfrqs >= 3, "gauss_weights: the Hilbert mesh needs at least 3 points");
  const int nrow = kpt1p ? nkpt1 : nk;
  SPX_CHECK(ctx, nrow > 0, "gauss_weights: nkpt1 must be positive with kpt1p");
  const long nr = (long)nrow * nocc * nunocc;
  SPX_CHECK(ctx, ldwt >= nr, "gauss_weights: ldwt < rows*nocc*nunocc");
  SPX_CHECK(ctx, spx_owns(ctx, e_occ_dev, 8 * (size_t)nk * nocc) &&
                     spx_owns(ctx, e_unocc_dev, 8 * (size_t)nk * nunocc) &&
                     spx_owns(ctx, wtetra_dev, 8 * ((size_t)ldwt * (nfrqs - 1) + nr)),
            "gauss_weights: arrays are not context allocations of the stated shapes");
  // mesh: ascending and exponential (the same test as gauss.f:137-147)
  double a = 0.0, acc = 0.0;
  for (int i = 1; i < nfrqs; i++) {
    SPX_CHECK(ctx, frqs[i] > frqs[i - 1], "gauss_weights: frequency mesh not ascending");
    if (i != nfrqs - 1) {
      acc += (frqs[i] - frqs[i - 1]) / (frqs[i + 1] - frqs[i - 1]);
      const double r = acc / i;
      SPX_CHECK(ctx, i == 1 || std::fabs(r - a) <= 1e-10, "gauss_weights: mesh not exponential");
      a = r;
    }
  }
  std::vector<GaussBin> bins(nfrqs);
  for (int i = 0; i < nfrqs; i++) {
    if (frqs[i] == 0) {
      bins[i] = {0.0, 0.0, 0.0};
      continue;
    }
    const double df = i != nfrqs - 1 ? a * (frqs[i + 1] - frqs[i])
                                     : (1 - a) * (frqs[i] - frqs[i - 1]);
    bins[i] = {frqs[i] - df, frqs[i] + df, 2 * df};
  }
  // members of each row, ascending (the reference's loop over ikpt1)
  std::vector<int> off(nrow + 1, 0), mem;
  for (int k = 0; k < nk; k++) {
    const int p = kpt1p ? kpt1p[k] : k;
    SPX_CHECK(ctx, p >= -1 && p < nrow, "gauss_weights: kpt1p out of range");
    if (p >= 0) off[p + 1]++;
  }
  for (int p = 0; p < nrow; p++) off[p + 1] += off[p];
  mem.resize(off[nrow]);
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int k = 0; k < nk; k++) {
    const int p = kpt1p ? kpt1p[k] : k;
    if (p >= 0) mem[fill[p]++] = k;
  }
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *ws;
  const size_t bb = sizeof(GaussBin) * bins.size();
  int rc = spx_scratch(ctx, bb + 4 * (off.size() + mem.size()) + 16, &ws);
  if (rc) return rc;
  GaussBin *dbins = (GaussBin *)ws;
  int *doff = (int *)((char *)ws + bb), *dmem = doff + off.size();
  SPX_HIP(ctx, hipMemcpyAsync(dbins, bins.data(), bb, hipMemcpyHostToDevice, ctx->stream));
  SPX_HIP(ctx, hipMemcpyAsync(doff, off.data(), 4 * off.size(), hipMemcpyHostToDevice, ctx->stream));
  if (!mem.empty())
    SPX_HIP(ctx, hipMemcpyAsync(dmem, mem.data(), 4 * mem.size(), hipMemcpyHostToDevice, ctx->stream));
  const double pi = 3.14159265358979323846;
  const double gfac = 1.0 / (std::sqrt(2.0) * sigma2);
  const double g2fac = 1.0 / (std::sqrt(2 * pi) * sigma2);
  const double factor = std::sqrt(0.5 * pi) * sigma2;
  hipLaunchKernelGGL(gauss_weights_kernel, dim3((unsigned)((nr + 255) / 256), (unsigned)nfrqs),
                     dim3(256), 0, ctx->stream, nrow, nocc, nunocc, doff, dmem,
                     (const double *)e_occ_dev, (const double *)e_unocc_dev, efermi, sigma1, gfac,
                     g2fac, factor, (double)nk, nfrqs, dbins, (double *)wtetra_dev, ldwt);
  SPX_LAUNCHED(ctx);
  // the host tables are read by the copies above
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SPEXHIP_OK;
}

int spexhip_gauss_weights(spexhip_ctx *ctx, int nk, int nocc, int nunocc, const double *e_occ,
                          const double *e_unocc, double efermi, double sigma1, double sigma2,
                          int nfrqs, const double *frqs, const int *kpt1p, int nkpt1,
                          double *wtetra) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nk > 0 && nocc > 0 && nunocc > 0 && nfrqs > 0 && e_occ && e_unocc && wtetra,
            "gauss_weights: bad arguments");
  const int nrow = kpt1p ? nkpt1 : nk;
  SPX_CHECK(ctx, nrow > 0, "gauss_weights: nkpt1 must be positive with kpt1p");
  const long nr = (long)nrow * nocc * nunocc;
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *dEo = nullptr, *dEu = nullptr, *dW = nullptr;
  int rc = spx_alloc(ctx, 8 * (size_t)nk * nocc, &dEo);
  if (!rc) rc = spx_alloc(ctx, 8 * (size_t)nk * nunocc, &dEu);
  if (!rc) rc = spx_alloc(ctx, 8 * (size_t)nr * nfrqs, &dW);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(dEo, e_occ, 8 * (size_t)nk * nocc, hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dEu, e_unocc, 8 * (size_t)nk * nunocc, hipMemcpyHostToDevice,
                         ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "gauss_weights: upload failed: %s",
                      hipGetErrorString(e));
  }
  if (!rc)
    rc = spexhip_gauss_weights_dev(ctx, nk, nocc, nunocc, dEo, dEu, efermi, sigma1, sigma2, nfrqs,
                                   frqs, kpt1p, nkpt1, dW, nr);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(wtetra, dW, 8 * (size_t)nr * nfrqs, hipMemcpyDeviceToHost,
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "gauss_weights: download failed: %s",
                      hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(ctx->stream);
  for (void *p : {dEo, dEu, dW})
    if (p) spx_free(ctx, p);
  return rc;
}

}  // extern "C"
