// Momentum matrix of the k.p expansion at q = 0, replacing momentum_matrix
// (dwavefproducts.f:89-284) for one spin, k points in full, no SOC:
//
//   momentum(b2, b1, k, j) = <phi_b1,k | -i d/dx_j | phi_b2,k>
//
// Muffin tin (:129-254).  For one atom the reference builds, per occupied
// band, left vectors cvec1/2/3 = A_x cmt1 (A_x real: the gradient-formula
// coefficients F, G of :131-142 times the radial integrals qmat1/qmat2),
// dots them with cmt2, divides by i and turns the spherical components
// (c1, c2, c3) into Cartesian ones ((c1 - c3)/sqrt2, -i (c1 + c3)/sqrt2,
// c2).  Folding the division and the transform into the operators gives
//   momentum_j = sum_l conj((B_j cmt1)(l)) cmt2(l),
//   B_1 = i/sqrt2 (A_1 - A_3),  B_2 = -1/sqrt2 (A_1 + A_3),  B_3 = i A_2,
// built once per plan on the host.  The device applies B_j to the occupied
// coefficients of every k at once (apply_b_kernel) and contracts the result
// with the unoccupied coefficients (conj_contract_kernel).
//
// Plane waves (:256-281).  vec1 = olapvec(cpw1) (overlap.f:298-315),
// vec1(G_i) = sum_j conj(cstep(G_j - G_i)) cpw1(G_j), then
// momentum_j += sum_G conj(vec1(G)) (k+G)_j cpw2(G): olapvec_kernel forms
// vec1 scaled by the three Cartesian components of k+G, the same
// contraction kernel sums over G.
#include <cmath>
#include <vector>

#include "gemm_f64.hpp"
#include "plan.hpp"

namespace {

constexpr int NT = 256;

// Z[j][k](l, b1) = sum_l' B_j(l, l') C1_k(l', b1) for the nocc occupied bands
// of every packet k (blockIdx.z = 3*ik + j).  B_j: nn x nn column-major;
// C1_k(l', b1) = cmt(l', ic, b1lo + b1, k1[ik]).  Z column (b1 + nocc*j) of
// packet ik at Z + ldz*(b1 + nocc*j) + zstride*ik.
__global__ __launch_bounds__(NT) void apply_b_kernel(int nn, int nocc, const double2 *__restrict__ B,
                                                     const double2 *__restrict__ cmt, long ldcol,
                                                     long kstride, const int *__restrict__ k1,
                                                     double2 *__restrict__ Z, long ldz,
                                                     long zstride) {
  const int l = blockIdx.x * 64 + (threadIdx.x & 63);
  const int b1 = blockIdx.y * (NT / 64) + (threadIdx.x >> 6);
  const int ik = blockIdx.z / 3, j = blockIdx.z % 3;
  if (l >= nn || b1 >= nocc) return;
  const double2 *Bj = B + (long)j * nn * nn;
  const double2 *c1 = cmt + kstride * k1[ik] + ldcol * b1;
  double sr = 0.0, si = 0.0;
  for (int lp = 0; lp < nn; lp++) {
    const double2 b = Bj[l + (long)nn * lp];
    const double2 c = c1[lp];
    sr += b.x * c.x - b.y * c.y;
    si += b.x * c.y + b.y * c.x;
  }
  Z[zstride * ik + ldz * (b1 + (long)nocc * j) + l] = make_double2(sr, si);
}

// vec1 of every basis function G_i of packet k (blockIdx.z), scaled by the
// Cartesian components of k + G_i:
//   Z[j](i, b1) = (rlat (k + G_i))_j sum_jj conj(cstep(G_jj - G_i)) cpw(G_jj, b1lo+b1, k).
// One thread per G_i, eight occupied bands per block row (blockIdx.y); the
// G vectors and coefficients of 64 basis functions at a time in LDS.
constexpr int OB = 8;
__global__ __launch_bounds__(NT) void olapvec_kernel(
    const int *__restrict__ gpt, int maxgpt, const int *__restrict__ ngpt,
    const double *__restrict__ kpt, const double2 *__restrict__ cpw, int nband, int b1lo,
    int nocc, const double2 *__restrict__ cstep, int c1, int c2, int c3, double r0, double r1,
    double r2, double r3, double r4, double r5, double r6, double r7, double r8,
    const int *__restrict__ k1, double2 *__restrict__ Z, long ldz, long zstride) {
  __shared__ int Gs[64][3];
  __shared__ double2 Cs[64][OB];
  const int ik = blockIdx.z, k = k1[ik];
  const int ng = ngpt[k];
  const int i = blockIdx.x * NT + threadIdx.x;
  const int bb = blockIdx.y * OB;
  const int *G = gpt + 3L * maxgpt * k;
  const double2 *C = cpw + (long)maxgpt * ((long)nband * k + b1lo);
  int gi[3] = {0, 0, 0};
  if (i < ng)
    for (int d = 0; d < 3; d++) gi[d] = G[3 * i + d];
  double2 acc[OB];
#pragma unroll
  for (int q = 0; q < OB; q++) acc[q] = make_double2(0.0, 0.0);
  const int w0 = 2 * c1 + 1, w1 = 2 * c2 + 1;
  for (int j0 = 0; j0 < ng; j0 += 64) {
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 3; e += NT) {
      const int jj = j0 + e / 3;
      Gs[e / 3][e % 3] = jj < ng ? G[3 * jj + e % 3] : 0;
    }
    for (int e = threadIdx.x; e < 64 * OB; e += NT) {
      const int jj = j0 + e / OB, b = bb + e % OB;
      Cs[e / OB][e % OB] = (jj < ng && b < nocc) ? C[(long)maxgpt * b + jj] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (i < ng) {
      const int len = ng - j0 < 64 ? ng - j0 : 64;
      for (int jj = 0; jj < len; jj++) {
        const int dx = Gs[jj][0] - gi[0], dy = Gs[jj][1] - gi[1], dz = Gs[jj][2] - gi[2];
        const double2 s = cstep[((long)(dz + c3) * w1 + (dy + c2)) * w0 + (dx + c1)];
#pragma unroll
        for (int q = 0; q < OB; q++) {  // conj(s) * c
          const double2 c = Cs[jj][q];
          acc[q].x += s.x * c.x + s.y * c.y;
          acc[q].y += s.x * c.y - s.y * c.x;
        }
      }
    }
  }
  if (i >= ng) return;
  const double *kp = kpt + 3L * k;
  const double v0 = kp[0] + gi[0], v1 = kp[1] + gi[1], v2 = kp[2] + gi[2];
  const double qg[3] = {r0 * v0 + r3 * v1 + r6 * v2, r1 * v0 + r4 * v1 + r7 * v2,
                        r2 * v0 + r5 * v1 + r8 * v2};
  double2 *Zk = Z + zstride * ik;
#pragma unroll
  for (int q = 0; q < OB; q++) {
    const int b = bb + q;
    if (b >= nocc) break;
#pragma unroll
    for (int j = 0; j < 3; j++)
      Zk[ldz * (b + (long)nocc * j) + i] = make_double2(acc[q].x * qg[j], acc[q].y * qg[j]);
  }
}

// P[t0(ik) + b2 + nunocc*b1 + ldp*j] += sum_{l < len(ik)} conj(Z_ik(l, c)) C2_ik(l, b2)
// with c = b1 + nocc*j.  C2_ik(l, b2) = C + kstride*k1[ik] + ld2*b2 + l, len
// either fixed (lenv == nullptr) or lenv[k1[ik]].  Tile: 64 unoccupied
// bands x 32 columns per 256-thread block, 16 l per LDS stage.
__global__ __launch_bounds__(NT) void conj_contract_kernel(
    int len0, const int *__restrict__ lenv, int nunocc, int nocc, const double2 *__restrict__ C,
    long ld2, long kstride, const int *__restrict__ k1, const double2 *__restrict__ Z, long ldz,
    long zstride, double2 *__restrict__ P, long ldp) {
  __shared__ double2 Cs[16][65];
  __shared__ double2 Zs[16][32];
  const int ik = blockIdx.z, k = k1[ik];
  const int len = lenv ? lenv[k] : len0;
  const int b0 = blockIdx.x * 64, c0 = blockIdx.y * 32;
  const int tid = threadIdx.x, bl = tid & 63, cg = tid >> 6;
  const int ncol = 3 * nocc;
  const double2 *C2 = C + kstride * k;
  const double2 *Zk = Z + zstride * ik;
  double2 acc[8];
#pragma unroll
  for (int q = 0; q < 8; q++) acc[q] = make_double2(0.0, 0.0);
  for (int l0 = 0; l0 < len; l0 += 16) {
    __syncthreads();
    for (int e = tid; e < 16 * 64; e += NT) {
      const int l = e % 16, b = e / 16;
      Cs[l][b] = (l0 + l < len && b0 + b < nunocc) ? C2[ld2 * (b0 + b) + l0 + l]
                                                  : make_double2(0.0, 0.0);
    }
    for (int e = tid; e < 16 * 32; e += NT) {
      const int l = e % 16, c = e / 16;
      Zs[l][c] = (l0 + l < len && c0 + c < ncol) ? Zk[ldz * (c0 + c) + l0 + l]
                                                : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int l = 0; l < 16; l++) {
      const double2 cv = Cs[l][bl];
#pragma unroll
      for (int q = 0; q < 8; q++) {  // conj(z) * c
        const double2 z = Zs[l][cg * 8 + q];
        acc[q].x += z.x * cv.x + z.y * cv.y;
        acc[q].y += z.x * cv.y - z.y * cv.x;
      }
    }
  }
  const int b2 = b0 + bl;
  if (b2 >= nunocc) return;
  const long t0 = (long)ik * nocc * nunocc;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int c = c0 + cg * 8 + q;
    if (c >= ncol) break;
    const int b1 = c % nocc, j = c / nocc;
    double2 &p = P[t0 + b2 + (long)nunocc * b1 + ldp * j];
    p.x += acc[q].x;
    p.y += acc[q].y;
  }
}

__global__ void zero_cols_kernel(double2 *__restrict__ P, long n, long ldp) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 3 * n) return;
  P[(g / n) * ldp + g % n] = make_double2(0.0, 0.0);
}

// F and G coefficients of the gradient formula (dwavefproducts.f:131-142),
// [lm][0..2] for m offsets -1, 0, 1 of the reference.
void fg_coeffs(int lmax, std::vector<double> &f, std::vector<double> &g) {
  const int n = (lmax + 1) * (lmax + 1);
  f.assign(3 * n, 0.0);
  g.assign(3 * n, 0.0);
  int lm = 0;
  for (int l = 0; l <= lmax; l++)
    for (int m = -l; m <= l; m++, lm++) {
      f[3 * lm + 0] = -std::sqrt(1.0 * (l + m + 1) * (l + m + 2) / (2.0 * (2 * l + 1) * (2 * l + 3)));
      f[3 * lm + 1] = std::sqrt(1.0 * (l - m + 1) * (l + m + 1) / ((2.0 * l + 1) * (2 * l + 3)));
      f[3 * lm + 2] = -std::sqrt(1.0 * (l - m + 1) * (l - m + 2) / (2.0 * (2 * l + 1) * (2 * l + 3)));
      if (l > 0) {
        g[3 * lm + 0] = std::sqrt(1.0 * (l - m) * (l - m - 1) / (2.0 * (2 * l - 1) * (2 * l + 1)));
        g[3 * lm + 1] = std::sqrt(1.0 * (l - m) * (l + m) / ((2.0 * l - 1) * (2 * l + 1)));
        g[3 * lm + 2] = std::sqrt(1.0 * (l + m) * (l + m - 1) / (2.0 * (2 * l - 1) * (2 * l + 1)));
      }
    }
}

// The operators A_x (x = 0, 1, 2 for cvec1, cvec2, cvec3) of one atom type
// as the loops of dwavefproducts.f:198-240 apply them:
// A_x(lm0 + n0', lm1 + n1') += coefficient * qmat(n1', n0', l).
void build_a(const spexhip_mom_type &T, const std::vector<double> &f, const std::vector<double> &g,
             int nn, std::vector<double> &A) {
  A.assign(3 * (size_t)nn * nn, 0.0);
  const int *nx = T.nindx;
  const int mi = T.maxindx;
  auto q2 = [&](int n1, int n0, int l) { return T.qmat2[n1 + (long)mi * (n0 + (long)mi * l)]; };
  auto q1 = [&](int n1, int n0, int l) { return T.qmat1[n1 + (long)mi * (n0 + (long)mi * l)]; };
  auto add = [&](int x, int row, int col, double v) { A[(size_t)x * nn * nn + row + (size_t)nn * col] += v; };
  const int dm[3] = {1, 0, -1};
  int lm_0 = 0, lm_1 = nx[0], lm = 0;
  for (int l = 0; l < T.lcut; l++) {  // band-1 functions of l+1 (qmat2)
    const int n0 = nx[l], n1 = nx[l + 1];
    for (int m = -l; m <= l; m++, lm++) {
      const int lm0 = lm_0 + (m + l) * n0;
      for (int x = 0; x < 3; x++) {
        const int lm1 = lm_1 + (m + dm[x] + l + 1) * n1;
        for (int a = 0; a < n0; a++)
          for (int b = 0; b < n1; b++) add(x, lm0 + a, lm1 + b, f[3 * lm + x] * q2(b, a, l));
      }
    }
    lm_0 += (2 * l + 1) * n0;
    lm_1 += (2 * l + 3) * n1;
  }
  lm_0 = nx[0];
  lm_1 = 0;
  lm = 1;
  for (int l = 1; l <= T.lcut; l++) {  // band-1 functions of l-1 (qmat1)
    const int n0 = nx[l], n1 = nx[l - 1];
    for (int m = -l; m <= l; m++, lm++) {
      const int lm0 = lm_0 + (m + l) * n0;
      for (int x = 0; x < 3; x++) {
        if (std::abs(m + dm[x]) > l - 1) continue;
        const int lm1 = lm_1 + (m + dm[x] + l - 1) * n1;
        for (int a = 0; a < n0; a++)
          for (int b = 0; b < n1; b++) add(x, lm0 + a, lm1 + b, g[3 * lm + x] * q1(b, a, l));
      }
    }
    lm_0 += (2 * l + 1) * n0;
    lm_1 += (2 * l - 1) * n1;
  }
}

int momentum_plan_build(spexhip_ctx *ctx, const spexhip_mom_desc *d, spexhip_plan *p) {
  p->kind = 3;
  SPX_CHECK(ctx, d->ntype >= 0 && (d->ntype == 0 || d->types) && d->maxlmindx >= 0 &&
                     d->rlat && d->nkpt > 0 && d->maxgpt > 0 && d->gpt && d->ngpt && d->kpt &&
                     d->cstep,
            "momentum plan: bad descriptor");
  for (int i = 0; i < 3; i++)
    SPX_CHECK(ctx, d->cext[i] >= 0 && d->cext[i] <= 256, "momentum plan: cext out of range");
  int lmax = 0;
  for (int it = 0; it < d->ntype; it++) {
    const spexhip_mom_type &T = d->types[it];
    SPX_CHECK(ctx, T.lcut >= 0 && T.lcut <= 12 && T.nindx && T.neq > 0 && T.qmat1 && T.qmat2,
              "momentum plan: bad atom type %d", it);
    for (int l = 0; l <= T.lcut; l++)
      SPX_CHECK(ctx, T.nindx[l] >= 0 && T.nindx[l] <= T.maxindx, "momentum plan: nindx > maxindx");
    lmax = std::max(lmax, T.lcut);
  }
  std::vector<double> f, g, A;
  fg_coeffs(lmax, f, g);
  const double s2 = 1.0 / std::sqrt(2.0);
  int ncent = 0;
  for (int it = 0; it < d->ntype; it++) {
    const spexhip_mom_type &T = d->types[it];
    int nn = 0;
    for (int l = 0; l <= T.lcut; l++) nn += (2 * l + 1) * T.nindx[l];
    SPX_CHECK(ctx, nn <= d->maxlmindx, "momentum plan: type %d needs %d > maxlmindx coefficients",
              it, nn);
    build_a(T, f, g, nn, A);
    const size_t nn2 = (size_t)nn * nn;
    std::vector<double2> B(3 * nn2);
    for (size_t e = 0; e < nn2; e++) {
      const double a1 = A[e], a2 = A[nn2 + e], a3 = A[2 * nn2 + e];
      B[e] = make_double2(0.0, s2 * (a1 - a3));      // i/sqrt2 (A1 - A3)
      B[nn2 + e] = make_double2(-s2 * (a1 + a3), 0.0);  // -1/sqrt2 (A1 + A3)
      B[2 * nn2 + e] = make_double2(0.0, a2);        // i A2
    }
    spexhip_plan::MomType mt;
    mt.neq = T.neq;
    mt.nn = nn;
    if (nn > 0) {
      int rc = spx_alloc(ctx, 16 * B.size(), (void **)&mt.B);
      if (rc) return rc;
      p->owned.push_back(mt.B);
      SPX_HIP(ctx, hipMemcpy(mt.B, B.data(), 16 * B.size(), hipMemcpyHostToDevice));
    }
    p->mom.push_back(mt);
    ncent += T.neq;
  }
  p->mom_ncent = ncent;
  p->maxlmindx = d->maxlmindx;
  // plane waves: every G_j - G_i of one k's basis must lie inside cstep
  p->nkpt = d->nkpt;
  p->maxgpt = d->maxgpt;
  p->ngpt.assign(d->ngpt, d->ngpt + d->nkpt);
  for (int k = 0; k < d->nkpt; k++) {
    SPX_CHECK(ctx, d->ngpt[k] >= 0 && d->ngpt[k] <= d->maxgpt, "momentum plan: ngpt[%d] out of range", k);
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int i = 0; i < d->ngpt[k]; i++)
      for (int x = 0; x < 3; x++) {
        const int v = d->gpt[3 * ((long)k * d->maxgpt + i) + x];
        lo[x] = std::min(lo[x], v);
        hi[x] = std::max(hi[x], v);
      }
    for (int x = 0; x < 3; x++)
      SPX_CHECK(ctx, d->ngpt[k] == 0 || hi[x] - lo[x] <= d->cext[x],
                "momentum plan: G differences of k %d reach %d > cext[%d] = %d", k, hi[x] - lo[x],
                x, d->cext[x]);
  }
  const size_t gb = sizeof(int) * 3 * (size_t)d->nkpt * d->maxgpt;
  int rc = spx_alloc(ctx, gb, (void **)&p->mom_gpt);
  if (rc) return rc;
  p->owned.push_back(p->mom_gpt);
  SPX_HIP(ctx, hipMemcpy(p->mom_gpt, d->gpt, gb, hipMemcpyHostToDevice));
  const size_t ncs = (size_t)(2 * d->cext[0] + 1) * (2 * d->cext[1] + 1) * (2 * d->cext[2] + 1);
  rc = spx_alloc(ctx, 16 * ncs, (void **)&p->mom_cstep);
  if (rc) return rc;
  p->owned.push_back(p->mom_cstep);
  SPX_HIP(ctx, hipMemcpy(p->mom_cstep, d->cstep, 16 * ncs, hipMemcpyHostToDevice));
  for (int i = 0; i < 3; i++) p->cext[i] = d->cext[i];
  for (int i = 0; i < 9; i++) p->rlat[i] = d->rlat[i];
  rc = spx_alloc(ctx, sizeof(int) * d->nkpt, (void **)&p->mom_ngpt);
  if (rc) return rc;
  p->owned.push_back(p->mom_ngpt);
  SPX_HIP(ctx, hipMemcpy(p->mom_ngpt, d->ngpt, sizeof(int) * d->nkpt, hipMemcpyHostToDevice));
  rc = spx_alloc(ctx, 8 * 3 * (size_t)d->nkpt, (void **)&p->mom_kpt);
  if (rc) return rc;
  p->owned.push_back(p->mom_kpt);
  SPX_HIP(ctx, hipMemcpy(p->mom_kpt, d->kpt, 8 * 3 * (size_t)d->nkpt, hipMemcpyHostToDevice));
  return SPEXHIP_OK;
}

int momentum_launch(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cmt,
                    int ncent, const void *cpw, int nband, int nk, const int *k1, int b1lo,
                    int nocc, int b2lo, int nunocc, void *P, long ldp, int accumulate) {
  const long n = (long)nk * nocc * nunocc;
  if (!accumulate && n > 0) {
    hipLaunchKernelGGL(zero_cols_kernel, dim3((unsigned)((3 * n + 255) / 256)), dim3(256), 0, st,
                       (double2 *)P, n, ldp);
    SPX_LAUNCHED(ctx);
  }
  if (n == 0) return SPEXHIP_OK;
  int nnmax = 0;
  for (auto &t : p->mom) nnmax = std::max(nnmax, t.nn);
  const long zmt = (long)nnmax * nocc * 3, zpw = (long)p->maxgpt * nocc * 3;
  const size_t kbytes = (4 * (size_t)nk + 255) / 256 * 256;
  const size_t bytes = kbytes + 16 * (size_t)nk * (size_t)(zmt + zpw);
  void *ws = nullptr;
  int rc = spx_scratch(ctx, bytes, &ws);
  if (rc) return rc;
  int *k1d = (int *)ws;
  double2 *Zm = (double2 *)((char *)ws + kbytes), *Zp = Zm + (long)nk * zmt;
  SPX_HIP(ctx, hipMemcpyAsync(k1d, k1, 4 * (size_t)nk, hipMemcpyHostToDevice, st));
  // muffin tin, atom by atom
  const long ldcol = (long)p->maxlmindx * ncent;  // between bands
  const long kstr = ldcol * nband;                // between k points
  int ic = 0;
  for (auto &t : p->mom)
    for (int e = 0; e < t.neq; e++, ic++) {
      if (t.nn == 0) continue;
      const double2 *base = (const double2 *)cmt + (long)p->maxlmindx * ic;
      hipLaunchKernelGGL(apply_b_kernel, dim3((t.nn + 63) / 64, (nocc + 3) / 4, 3 * nk), dim3(NT),
                         0, st, t.nn, nocc, t.B, base + ldcol * b1lo, ldcol, kstr, k1d, Zm,
                         (long)t.nn, zmt);
      SPX_LAUNCHED(ctx);
      hipLaunchKernelGGL(conj_contract_kernel, dim3((nunocc + 63) / 64, (3 * nocc + 31) / 32, nk),
                         dim3(NT), 0, st, t.nn, (const int *)nullptr, nunocc, nocc,
                         base + ldcol * b2lo, ldcol, kstr, k1d, Zm, (long)t.nn, zmt,
                         (double2 *)P, ldp);
      SPX_LAUNCHED(ctx);
    }
  // plane waves
  const double *r = p->rlat;
  hipLaunchKernelGGL(olapvec_kernel, dim3((p->maxgpt + NT - 1) / NT, (nocc + OB - 1) / OB, nk),
                     dim3(NT), 0, st, p->mom_gpt, p->maxgpt, p->mom_ngpt, p->mom_kpt,
                     (const double2 *)cpw, nband, b1lo, nocc, p->mom_cstep, p->cext[0], p->cext[1],
                     p->cext[2], r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], k1d, Zp,
                     (long)p->maxgpt, zpw);
  SPX_LAUNCHED(ctx);
  hipLaunchKernelGGL(conj_contract_kernel, dim3((nunocc + 63) / 64, (3 * nocc + 31) / 32, nk),
                     dim3(NT), 0, st, 0, p->mom_ngpt, nunocc, nocc,
                     (const double2 *)cpw + (long)p->maxgpt * b2lo, (long)p->maxgpt,
                     (long)p->maxgpt * nband, k1d, Zp, (long)p->maxgpt, zpw, (double2 *)P, ldp);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

}  // namespace

extern "C" {

int spexhip_momentum_plan_create(spexhip_ctx *ctx, const spexhip_mom_desc *desc,
                                 spexhip_plan **plan) {
  if (!ctx || !plan) return SPEXHIP_E_INVALID;
  *plan = nullptr;
  SPX_CHECK(ctx, desc, "momentum plan: null descriptor");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  spexhip_plan *p = new spexhip_plan();
  int rc = momentum_plan_build(ctx, desc, p);
  if (rc) {
    for (void *d : p->owned) spx_free(ctx, d);
    delete p;
    return rc;
  }
  *plan = p;
  return SPEXHIP_OK;
}

int spexhip_momentum_dev(spexhip_ctx *ctx, const spexhip_plan *plan, const void *cmt,
                         int ncent, const void *cpw, int nband, int nk, const int *k1,
                         int b1lo, int nocc, int b2lo, int nunocc, void *P, long ldp,
                         int accumulate) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plan && plan->kind == 3, "momentum: not a momentum plan");
  SPX_CHECK(ctx, nk >= 0 && nocc >= 0 && nunocc >= 0 && nband > 0 && (nk == 0 || k1),
            "momentum: bad sizes");
  SPX_CHECK(ctx, ncent == plan->mom_ncent, "momentum: ncent %d != the plan's %d centres", ncent,
            plan->mom_ncent);
  SPX_CHECK(ctx, b1lo >= 0 && b2lo >= 0 && b1lo + nocc <= nband && b2lo + nunocc <= nband,
            "momentum: band range outside 0..nband-1");
  for (int i = 0; i < nk; i++)
    SPX_CHECK(ctx, k1[i] >= 0 && k1[i] < plan->nkpt, "momentum: k index %d out of range", i);
  const long n = (long)nk * nocc * nunocc;
  SPX_CHECK(ctx, ldp >= n, "momentum: ldp %ld < n %ld", ldp, n);
  if (n > 0) {
    SPX_CHECK(ctx, spx_owns(ctx, P, 16 * ((size_t)ldp * 2 + n)), "momentum: P is not a context allocation of n x 3");
    if (plan->mom_ncent > 0)
      SPX_CHECK(ctx, spx_owns(ctx, cmt, 16 * (size_t)plan->maxlmindx * ncent * nband * plan->nkpt),
                "momentum: cmt is not a context allocation of (maxlmindx, ncent, nband, nkpt)");
    SPX_CHECK(ctx, spx_owns(ctx, cpw, 16 * (size_t)plan->maxgpt * nband * plan->nkpt),
              "momentum: cpw is not a context allocation of (maxgpt, nband, nkpt)");
  }
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  int rc = momentum_launch(ctx, ctx->stream, plan, cmt, ncent, cpw, nband, nk, k1, b1lo, nocc,
                           b2lo, nunocc, P, ldp, accumulate);
  if (rc) return rc;
  SPX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // k1 was copied from the caller's array
  return SPEXHIP_OK;
}

int spexhip_momentum_matrix(spexhip_ctx *ctx, const spexhip_plan *plan, const double *cmt,
                            int ncent, const double *cpw, int nband, int nk, const int *k1,
                            int b1lo, int nocc, int b2lo, int nunocc, double *momentum,
                            int accumulate) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, plan && plan->kind == 3 && cpw && momentum && (cmt || plan->mom_ncent == 0),
            "momentum_matrix: bad arguments");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const long n = (long)nk * nocc * nunocc;
  const size_t cb = 16 * (size_t)plan->maxlmindx * ncent * nband * plan->nkpt;
  const size_t pb = 16 * (size_t)plan->maxgpt * nband * plan->nkpt;
  void *dC = nullptr, *dW = nullptr, *dP = nullptr;
  int rc = cb ? spx_alloc(ctx, cb, &dC) : SPEXHIP_OK;
  if (!rc) rc = spx_alloc(ctx, pb, &dW);
  if (!rc) rc = spx_alloc(ctx, 16 * 3 * (size_t)(n > 0 ? n : 1), &dP);
  if (!rc) {
    hipError_t e = cb ? hipMemcpyAsync(dC, cmt, cb, hipMemcpyHostToDevice, ctx->stream) : hipSuccess;
    if (e == hipSuccess) e = hipMemcpyAsync(dW, cpw, pb, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && accumulate && n > 0)
      e = hipMemcpyAsync(dP, momentum, 16 * 3 * n, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "momentum_matrix: upload failed: %s", hipGetErrorString(e));
  }
  if (!rc)
    rc = spexhip_momentum_dev(ctx, plan, dC, ncent, dW, nband, nk, k1, b1lo, nocc, b2lo, nunocc,
                              dP, n > 0 ? n : 1, accumulate);
  if (!rc && n > 0) {
    hipError_t e = hipMemcpy(momentum, dP, 16 * 3 * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "momentum_matrix: download failed: %s", hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (dC) spx_free(ctx, dC);
  if (dW) spx_free(ctx, dW);
  if (dP) spx_free(ctx, dP);
  return rc;
}

}  // extern "C"
