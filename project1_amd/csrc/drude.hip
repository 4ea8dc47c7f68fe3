// Drude term of the head at q = 0 for metals.
//
// * Fermi-surface weights, replacing tetrahedron4_init (tetrahedron.f:
//   708-838) for a full cubic k mesh: per band and tetrahedron the
//   cross-section e = efermi is a triangle or a quadrangle of two triangles
//   (:756-784) whose area times the corner weights goes to the cube corners
//   (:786-791); the sum is divided by 3*6*nkpt and averaged over degenerate
//   states (:826-838).  The reference scatters from cubes to corners; here
//   one thread per (k, band) gathers from the 8 (cube, corner position)
//   pairs whose corner is k, the same terms.
// * Plasma frequency, replacing susceptibility.f:1404-1441 and :1693-1700:
//   plasma^2 = 4 pi/(3 vol) sum_{n,k} wintgr4(n,k) sum_j (Re of the mean of
//   momentum(i,i,k,j) over the degenerate group of n)^2, times the spin
//   factor; one workgroup, a fixed reduction order.
// * Head update, replacing correlation.f:847-862: head(j,j,w) -= drude(w)
//   with drude = -plasma^2/w^2 on the imaginary axis (correlation.f:
//   1147-1157) and drudec = plasma^2/w^2 at complex frequencies (:1159-1169).
#include <cmath>
#include <vector>

#include "tetra_mesh.hpp"

namespace {

// Corner weight of original corner `pos` of one tetrahedron cut by ef
// (tetrahedron.f:753-784): atria * wtria summed over the triangles.  The
// corners are ordered by a stable insertion sort (rorderp, util.f:37-56).
__device__ double cross_section_weight(const double et[4], double ef, int pos) {
  int pnt[4];
  for (int i = 0; i < 4; i++) {
    pnt[i] = i;
    for (int j = 0; j < i; j++)
      if (et[pnt[j]] > et[i]) {
        for (int k = i; k > j; k--) pnt[k] = pnt[k - 1];
        pnt[j] = i;
        break;
      }
  }
  const double e1 = et[pnt[0]], e2 = et[pnt[1]], e3 = et[pnt[2]], e4 = et[pnt[3]];
  int r = 0;  // sorted position of corner pos
  for (int j = 0; j < 4; j++)
    if (pnt[j] == pos) r = j;
  if (ef > e3) {
    const double f[3] = {(ef - e1) / (e4 - e1), (ef - e2) / (e4 - e2), (ef - e3) / (e4 - e3)};
    const double a = 3 * (ef - e4) * (ef - e4) / ((e4 - e1) * (e4 - e2) * (e4 - e3));
    const double w = r < 3 ? 1 - f[r] : f[0] + f[1] + f[2];
    return a * w;
  }
  if (ef > e2) {
    const double f1 = (ef - e1) / (e3 - e1), f2 = (ef - e1) / (e4 - e1);
    const double f3 = (ef - e2) / (e3 - e2), f4 = (ef - e2) / (e4 - e2);
    const double rd = (e4 - e2) * (e3 - e1);
    const double a1 = 3 * f3 * (e3 - ef) / rd, a2 = 3 * f2 * (e4 - ef) / rd;
    switch (r) {
      case 0: return a1 * (1 - f1) + a2 * (2 - f1 - f2);
      case 1: return a1 * (2 - f3 - f4) + a2 * (1 - f4);
      case 2: return a1 * (f1 + f3) + a2 * f1;
      default: return a1 * f4 + a2 * (f2 + f4);
    }
  }
  // the reference's area here is 3 f1 f2/(e3 - e1) (tetrahedron.f:779)
  const double f[3] = {(ef - e1) / (e2 - e1), (ef - e1) / (e3 - e1), (ef - e1) / (e4 - e1)};
  const double a = 3 * f[0] * f[1] / (e3 - e1);
  return a * (r == 0 ? 3 - f[0] - f[1] - f[2] : f[r - 1]);
}

template <int DEC>
__global__ __launch_bounds__(256) void fermi_weights_kernel(int n1, int n2, int n3, int nband,
                                                            const double *__restrict__ ene,
                                                            double ef, double *__restrict__ w) {
  const long nk = (long)n1 * n2 * n3;
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nk * nband) return;
  const int n = (int)(g % nband);
  const long k = g / nband;
  const int i1 = (int)(k / ((long)n2 * n3)), i2 = (int)((k / n3) % n2), i3 = (int)(k % n3);
  double acc = 0.0;
  for (int c = 0; c < 8; c++) {  // k is corner c of the cube at k - offset(c)
    const int b1 = (i1 - (c & 1) + n1) % n1, b2 = (i2 - ((c >> 1) & 1) + n2) % n2,
              b3 = (i3 - (c >> 2) + n3) % n3;
    double ec[8];
    for (int q = 0; q < 8; q++) {
      const long kq = ((long)((b1 + (q & 1)) % n1) * n2 + (b2 + ((q >> 1) & 1)) % n2) * n3 +
                      (b3 + (q >> 2)) % n3;
      ec[q] = ene[kq * nband + n];
    }
    for (int it = 0; it < 6; it++) {
      const int pos = corner_pos<DEC>(it, c);
      if (pos < 0) continue;
      double et[4];
      bool above = true, below = true;
      for (int j = 0; j < 4; j++) {
        et[j] = ec[kDecomp[DEC][it][j]];
        above = above && et[j] >= ef;
        below = below && et[j] <= ef;
      }
      if (above || below) continue;
      acc += cross_section_weight(et, ef, pos);
    }
  }
  w[g] = acc / 3 / 6 / (double)nk;
}

// Degenerate states: a band n and the following ones within 1e-8 of e(n)
// share their mean weight (tetrahedron.f:826-838).  One thread per k.
__global__ void fermi_degenerate_kernel(long nk, int nband, const double *__restrict__ ene,
                                        double *__restrict__ w) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  const double *e = ene + k * nband;
  double *wk = w + k * nband;
  int n = 0;
  while (n < nband - 1) {
    int n1 = n;
    while (fabs(e[n1 + 1] - e[n]) < 1e-8) {
      n1++;
      if (n1 == nband - 1) break;
    }
    double s = 0.0;
    for (int i = n; i <= n1; i++) s += wk[i];
    s /= (n1 - n + 1);
    for (int i = n; i <= n1; i++) wk[i] = s;
    n = n1 + 1;
  }
}

// sum_{k,n} w(k,n) * integrand(k,n) in one 256-thread workgroup: thread k
// (strided) walks the degenerate groups of k (SPEX's deg, 0-based: at the
// first band of a group, its last band), then a tree reduction in LDS.
__global__ __launch_bounds__(256) void plasma_sum_kernel(long nk, int nband,
                                                         const double *__restrict__ w,
                                                         const double2 *__restrict__ P, long ldp,
                                                         const int *__restrict__ deg,
                                                         double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (long k = threadIdx.x; k < nk; k += 256) {
    int i1 = 0;
    while (i1 < nband) {
      const int i2 = deg[k * nband + i1];
      if (i2 >= nband) break;  // the group reaches past the bands given
      double re[3] = {0.0, 0.0, 0.0};
      for (int j = 0; j < 3; j++) {
        double a = 0.0;
        for (int i = i1; i <= i2; i++) a += P[(k * nband + i) * nband + i + ldp * j].x;
        re[j] = a / (i2 - i1 + 1);
      }
      const double integrand = re[0] * re[0] + re[1] * re[1] + re[2] * re[2];
      for (int i = i1; i <= i2; i++) s += w[k * nband + i] * integrand;
      i1 = i2 + 1;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// head(j,j,w) -= d(w) on the 3 x 3 x nw complex head (k1 + 3 k2 + 9 w).
__global__ void head_diag_sub_kernel(int nw, const double2 *__restrict__ d,
                                     double2 *__restrict__ head) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 3 * nw) return;
  const int w = g / 3, j = g % 3;
  double2 &h = head[j + 3 * j + 9 * w];
  h.x -= d[w].x;
  h.y -= d[w].y;
}

}  // namespace

extern "C" {

int spexhip_fermi_weights_dev(spexhip_ctx *ctx, const int *nkpt3, const double *rlat, int nband,
                              const void *ene, double efermi, void *w) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nkpt3 && rlat && nband > 0 && nkpt3[0] > 0 && nkpt3[1] > 0 && nkpt3[2] > 0,
            "fermi_weights: bad arguments");
  const long nk = (long)nkpt3[0] * nkpt3[1] * nkpt3[2];
  SPX_CHECK(ctx, spx_owns(ctx, ene, 8 * (size_t)nk * nband) && spx_owns(ctx, w, 8 * (size_t)nk * nband),
            "fermi_weights: ene and w must be context allocations of nk x nband doubles");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const long tot = nk * nband;
  const dim3 grid((unsigned)((tot + 255) / 256)), block(256);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, block, 0, ctx->stream, nkpt3[0], nkpt3[1], nkpt3[2], nband,
                       (const double *)ene, efermi, (double *)w);
  };
  switch (tetra_choice(nkpt3, rlat)) {
    case 0: go(fermi_weights_kernel<0>); break;
    case 1: go(fermi_weights_kernel<1>); break;
    case 2: go(fermi_weights_kernel<2>); break;
    default: go(fermi_weights_kernel<3>); break;
  }
  SPX_LAUNCHED(ctx);
  hipLaunchKernelGGL(fermi_degenerate_kernel, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0,
                     ctx->stream, nk, nband, (const double *)ene, (double *)w);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_plasma_sum_dev(spexhip_ctx *ctx, int nk, int nband, const void *w, const void *P,
                           long ldp, const int *deg, double vol, double *plasma2) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nk > 0 && nband > 0 && deg && plasma2 && vol > 0,
            "plasma_frequency: bad arguments");
  const long n = (long)nk * nband * nband;
  SPX_CHECK(ctx, ldp >= n, "plasma_frequency: ldp < nk*nband^2");
  SPX_CHECK(ctx, spx_owns(ctx, w, 8 * (size_t)nk * nband) &&
                     spx_owns(ctx, P, 16 * ((size_t)ldp * 2 + n)),
            "plasma_frequency: w or P is not a context allocation of the stated shape");
  // the groups as the sum walks them (susceptibility.f:1412-1420): at the
  // first band of a group its last band; a group reaching past the bands
  // given ends the walk for that k point
  for (long k = 0; k < nk; k++)
    for (int i1 = 0; i1 < nband;) {
      const int i2 = deg[k * nband + i1];
      if (i2 >= nband) break;
      SPX_CHECK(ctx, i2 >= i1, "plasma_frequency: deg not consistent at k=%ld band %d", k, i1);
      i1 = i2 + 1;
    }
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t db = 4 * (size_t)nk * nband;
  void *ws;
  int rc = spx_scratch(ctx, 256 + db, &ws);
  if (rc) return rc;
  double *sum = (double *)ws;
  int *ddeg = (int *)((char *)ws + 256);
  SPX_HIP(ctx, hipMemcpyAsync(ddeg, deg, db, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(plasma_sum_kernel, dim3(1), dim3(256), 0, ctx->stream, (long)nk, nband,
                     (const double *)w, (const double2 *)P, ldp, (const int *)ddeg, sum);
  SPX_LAUNCHED(ctx);
  double s = 0.0;
  SPX_HIP(ctx, hipMemcpyAsync(&s, sum, 8, hipMemcpyDeviceToHost, ctx->stream));
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *plasma2 = *plasma2 + 4 * M_PI / 3 / vol * s;  // susceptibility.f:1433
  return SPEXHIP_OK;
}

int spexhip_plasma_finish(spexhip_ctx *ctx, double plasma2, double spin_factor, double *plasma) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plasma && spin_factor > 0, "plasma_frequency: bad arguments");
  const double p2 = plasma2 * spin_factor;  // :1439
  SPX_CHECK(ctx, p2 >= -1e-10, "plasma_frequency: Plasma frequency undefined (%g)", p2);
  double p = std::sqrt(p2 > 0.0 ? p2 : 0.0);
  if (p * 27.2113831243217 < 1e-6) p = 0.0;  // below 1e-6 eV: zero (susceptibility.f:1696)
  *plasma = p;
  return SPEXHIP_OK;
}

int spexhip_plasma_frequency_dev(spexhip_ctx *ctx, int nk, int nband, const void *w,
                                 const void *P, long ldp, const int *deg, double vol,
                                 double spin_factor, double *plasma) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plasma && spin_factor > 0, "plasma_frequency: bad arguments");
  double p2 = 0.0;
  const int rc = spexhip_plasma_sum_dev(ctx, nk, nband, w, P, ldp, deg, vol, &p2);
  if (rc) return rc;
  return spexhip_plasma_finish(ctx, p2, spin_factor, plasma);
}

int spexhip_gamma_drude_dev(spexhip_ctx *ctx, int nw, const double *freq, double plasma,
                            void *head, int nwc, const double *freqc, void *headc) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nw >= 0 && nwc >= 0 && (nw == 0 || (freq && head)) && (nwc == 0 || (freqc && headc)),
            "gamma_drude: bad arguments");
  if (!(plasma > 0.0)) return SPEXHIP_OK;  // drude = 0
  SPX_CHECK(ctx, nw == 0 || spx_owns(ctx, head, 16 * 9 * (size_t)nw), "gamma_drude: head is not 3x3xnw");
  SPX_CHECK(ctx, nwc == 0 || spx_owns(ctx, headc, 16 * 9 * (size_t)nwc), "gamma_drude: headc is not 3x3xnwc");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<double2> d((size_t)nw + nwc);
  const double p2 = plasma * plasma;
  for (int i = 0; i < nw; i++)  // drude(frq) = -plasma^2/frq^2, infinity at 0
    d[i] = make_double2(freq[i] == 0.0 ? INFINITY : -p2 / (freq[i] * freq[i]), 0.0);
  for (int i = 0; i < nwc; i++) {  // drudec(frq) = plasma^2/frq^2 (complex frq)
    const double a = freqc[2 * i], b = freqc[2 * i + 1];
    if (a == 0.0 && b == 0.0) {
      d[nw + i] = make_double2(INFINITY, 0.0);
    } else {  // plasma^2 / (a + ib)^2 = plasma^2 (a^2 - b^2 - 2iab) / (a^2 + b^2)^2
      const double m = a * a + b * b;
      d[nw + i] = make_double2(p2 * (a * a - b * b) / (m * m), -p2 * 2 * a * b / (m * m));
    }
  }
  void *ws;
  int rc = spx_scratch(ctx, 16 * d.size(), &ws);
  if (rc) return rc;
  SPX_HIP(ctx, hipMemcpyAsync(ws, d.data(), 16 * d.size(), hipMemcpyHostToDevice, ctx->stream));
  if (nw > 0) {
    hipLaunchKernelGGL(head_diag_sub_kernel, dim3((3 * nw + 63) / 64), dim3(64), 0, ctx->stream, nw,
                       (const double2 *)ws, (double2 *)head);
    SPX_LAUNCHED(ctx);
  }
  if (nwc > 0) {
    hipLaunchKernelGGL(head_diag_sub_kernel, dim3((3 * nwc + 63) / 64), dim3(64), 0, ctx->stream,
                       nwc, (const double2 *)ws + nw, (double2 *)headc);
    SPX_LAUNCHED(ctx);
  }
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // d is a host temporary
  return SPEXHIP_OK;
}

}  // extern "C"
