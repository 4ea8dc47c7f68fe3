// Interstitial (IPW) part of the wave-function products by FFT convolution,
// replacing the FFTW path of wavefproducts.f:576-758 (plans fft.f:90-93).
//
// Reference (SUBTYPE 2, noinv, fullpw optional):
//   phi1(r) = conjg(FFT_bwd(c1)(r) / sqrt(vol)) [* theta(r)]    (:624-633)
//   psi2(r) = FFT_bwd(c2)(r) / nnfft                             (:658-665)
//   cprod(b2,b1,k,ig) = FFT_fwd(phi1*psi2)(imati.(Gm_ig - g_k) mod nfft)
//                                                                (:670-684)
// with FFT_bwd = sum_G c(G) e^{+2 pi i G.r/N}, FFT_fwd = sum_r p(r)
// e^{-2 pi i G.r/N}, both unnormalised (FFTW BACKWARD / FORWARD).
//
// Every transform here is a pruned separable DFT evaluated as FP64 complex
// MFMA products on v_mfma_f64_16x16x4f64.  Wave-function coefficients live
// in a 16^3 box of Fourier components (-8..7 per axis), so the backward
// transform starts from that box and the forward transform only produces
// the box (then gathers the mixed-basis IPWs):
//   pw_inverse  (one workgroup per state, waves take z planes):
//     U(gx,gy) = sum_gz e^{+i..gz z} C(gz,gy,gx)           (VALU, from LDS)
//     V(gx,y)  = U x E2         (MFMA, K = 16 gy)
//     psi(y,x) = V^T x E1       (MFMA, K = 16 gx) -> grid, scaled
//   pw_pair_xy (one wave per pair (b1,b2), workgroup = 4 b1 sharing b2):
//     R^T(y,gx) = (phi1*psi2)(y,x) x F1(x,gx)   (MFMA, K = N1)
//     Q(gy,gx)  = F2(gy,y) x R^T(y,gx)          (MFMA, K = N2)  -> slab
//   pw_pair_z  (one workgroup per pair):
//     cprod(t, ig) = sum_z e^{-i..gz z} Q_z(gy,gx) at the IPW's box point.
// The accumulator layout of one MFMA (column j = lane%16, rows lane/16+4r)
// is exactly the operand layout the next MFMA needs, so no transposes go
// through LDS between passes.
#include <algorithm>
#include <cmath>
#include <vector>

#include "plan.hpp"

namespace {

constexpr int NT = 256;

__device__ inline double2 twid(long m, int N, double sign) {
  long r = m % N;
  if (r < 0) r += N;
  double s, c;
  sincospi(2.0 * (double)r / (double)N, &s, &c);
  return make_double2(c, sign * s);
}

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// acc += a * b (complex, one 16x16x4 MFMA step)
__device__ inline void cmfma(d4_t &ar, d4_t &ai, double2 a, double2 b) {
  ar = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, ar, 0, 0, 0);
  ar = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, ar, 0, 0, 0);
  ai = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, ai, 0, 0, 0);
  ai = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, ai, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Backward transform of states: s < nA -> (kA, bandA+s) as phi1, else
// (kB, bandB+s-nA) as psi2.  grids[s] is N3 x N2 x N1 (x fastest).
template <int NB>
__global__ __launch_bounds__(NT) void pw_inverse_kernel(
    const double2 *__restrict__ cpw, int maxgpt, int nband, const int *__restrict__ gidx,
    const int *__restrict__ ngpt, int kA, int bandA, int nA, int kB, int bandB, int N1, int N2,
    int N3, double scaleA, double scaleB, const double2 *__restrict__ step,
    double2 *__restrict__ grids) {
  __shared__ double Cr[PW_BOX * PW_BOX * PW_BOX], Ci[PW_BOX * PW_BOX * PW_BOX];
  __shared__ double2 tz[64][PW_BOX];  // N3 <= 64
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  const int s = blockIdx.x;
  const bool occ = s < nA;
  const int kq = occ ? kA : kB;
  const int band = occ ? bandA + s : bandB + (s - nA);
  for (int e = tid; e < PW_BOX * PW_BOX * PW_BOX; e += NT) {
    Cr[e] = 0.0;
    Ci[e] = 0.0;
  }
  for (int e = tid; e < N3 * PW_BOX; e += NT) {
    const int z = e / PW_BOX, g = e % PW_BOX;
    tz[z][g] = twid((long)(g - PW_OFF) * z, N3, +1.0);
  }
  __syncthreads();
  const int ng = ngpt[kq];
  const double2 *c = cpw + (long)maxgpt * (band + (long)nband * kq);
  const int *gi = gidx + (long)maxgpt * kq;
  for (int ig = tid; ig < ng; ig += NT) {
    const int b = gi[ig];
    Cr[b] = c[ig].x;
    Ci[b] = c[ig].y;
  }
  __syncthreads();

  // constant MFMA operands: E2(gy, y) and E1(gx, x), lane layout B[k][j]
  double2 E2[NB][4], E1[NB][4];
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int g = 4 * kk + lk - PW_OFF, j = 16 * b + lr;
      E2[b][kk] = j < N2 ? twid((long)g * j, N2, +1.0) : make_double2(0.0, 0.0);
      E1[b][kk] = j < N1 ? twid((long)g * j, N1, +1.0) : make_double2(0.0, 0.0);
    }

  const long NN = (long)N1 * N2 * N3;
  double2 *out = grids + (long)s * NN;
  const double scale = occ ? scaleA : scaleB;
  for (int z = wv; z < N3; z += NT / 64) {
    double2 U[4];
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int gy = 4 * kk + lk;
      double ur = 0.0, ui = 0.0;
#pragma unroll 4
      for (int gz = 0; gz < PW_BOX; gz++) {
        const double2 t = tz[z][gz];
        const int b = gz * PW_BOX * PW_BOX + gy * PW_BOX + lr;
        const double cr = Cr[b], ci = Ci[b];
        ur += t.x * cr - t.y * ci;
        ui += t.x * ci + t.y * cr;
      }
      U[kk] = make_double2(ur, ui);
    }
    // y and x passes; for grids of up to 32 points per axis with three real
    // MFMAs per complex product (Gauss): t1 = ar.br, t2 = ai.bi,
    // t3 = (ar + ai)(br + bi); re = t1 - t2, im = t3 - t1 - t2.  Larger grids
    // keep four (the third accumulators would cost the second wave per SIMD).
    constexpr bool M3 = NB <= 2;
#pragma unroll
    for (int mb = 0; mb < NB; mb++) {
      if (16 * mb >= N2) continue;
      d4_t vr = {0, 0, 0, 0}, vi = {0, 0, 0, 0}, vs = {0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < 4; kk++) {
        const double2 u = U[kk], e = E2[mb][kk];
        if (M3) {
          vr = __builtin_amdgcn_mfma_f64_16x16x4f64(u.x, e.x, vr, 0, 0, 0);
          vi = __builtin_amdgcn_mfma_f64_16x16x4f64(u.y, e.y, vi, 0, 0, 0);
          vs = __builtin_amdgcn_mfma_f64_16x16x4f64(u.x + u.y, e.x + e.y, vs, 0, 0, 0);
        } else {
          cmfma(vr, vi, u, e);
        }
      }
      const d4_t yr = M3 ? vr - vi : vr, yi = M3 ? vs - vr - vi : vi;
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        if (16 * nb >= N1) continue;
        d4_t tr = {0, 0, 0, 0}, ti = {0, 0, 0, 0}, ts = {0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
          const double2 e = E1[nb][kk];
          if (M3) {
            tr = __builtin_amdgcn_mfma_f64_16x16x4f64(yr[kk], e.x, tr, 0, 0, 0);
            ti = __builtin_amdgcn_mfma_f64_16x16x4f64(yi[kk], e.y, ti, 0, 0, 0);
            ts = __builtin_amdgcn_mfma_f64_16x16x4f64(yr[kk] + yi[kk], e.x + e.y, ts, 0, 0, 0);
          } else {
            cmfma(tr, ti, make_double2(yr[kk], yi[kk]), e);
          }
        }
        const d4_t orr = M3 ? tr - ti : tr, oi = M3 ? ts - tr - ti : ti;
        const int x = 16 * nb + lr;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int y = 16 * mb + lk + 4 * r;
          if (x < N1 && y < N2) {
            const long g = x + (long)N1 * (y + (long)N2 * z);
            double2 v = make_double2(orr[r], oi[r]);
            if (occ) {
              v = make_double2(v.x * scale, -v.y * scale);
              if (step) v = cmul(v, step[g]);  // fullpw: conj(psi1)/sqrt(vol) * theta
            } else {
              v.x *= scale;
              v.y *= scale;
            }
            out[g] = v;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Forward passes x and y of the products of pairs (b1, b2).
// grids: [0, nocc) phi1 of the occupied states, [nocc, nocc+nunocc) psi2.
// slab[(b1*nunocc + b2)][z][gy][gx], 16 x 16 per plane.
template <int NB>
__global__ __launch_bounds__(NT) void pw_pair_xy_kernel(const double2 *__restrict__ grids,
                                                        int nocc, int nunocc, int N1, int N2,
                                                        int N3, double2 *__restrict__ slab) {
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  const int b2 = blockIdx.x;
  const int b1 = blockIdx.y * (NT / 64) + wv;
  if (b1 >= nocc) return;  // whole wave; this kernel has no workgroup barrier
  const long NN = (long)N1 * N2 * N3;
  const double2 *p1 = grids + (long)b1 * NN;
  const double2 *p2 = grids + (long)(nocc + b2) * NN;
  double2 *S = slab + ((long)b1 * nunocc + b2) * N3 * PW_BOX * PW_BOX;
  constexpr int KX = 4 * NB;  // x steps (N1 <= 16*NB)
  double2 F1[KX], F2[KX];
#pragma unroll
  for (int kk = 0; kk < KX; kk++) {
    const int x = 4 * kk + lk;  // B[k = x][j = gx]
    F1[kk] = x < N1 ? twid((long)(lr - PW_OFF) * x, N1, -1.0) : make_double2(0.0, 0.0);
    const int y = 4 * kk + lk;  // A[i = gy][k = y]
    F2[kk] = y < N2 ? twid((long)(lr - PW_OFF) * y, N2, -1.0) : make_double2(0.0, 0.0);
  }
  for (int z = 0; z < N3; z++) {
    d4_t xr[NB], xi[NB];
#pragma unroll
    for (int mb = 0; mb < NB; mb++) {
      xr[mb] = d4_t{0, 0, 0, 0};
      xi[mb] = d4_t{0, 0, 0, 0};
      if (16 * mb >= N2) continue;
      const int y = 16 * mb + lr;
#pragma unroll
      for (int kk = 0; kk < KX; kk++) {
        if (4 * kk < N1) {  // wave-uniform guard (keeps the loop unrollable)
          const int x = 4 * kk + lk;
          double2 a = make_double2(0.0, 0.0);
          if (x < N1 && y < N2) {
            const long g = x + (long)N1 * (y + (long)N2 * z);
            a = cmul(p1[g], p2[g]);
          }
          cmfma(xr[mb], xi[mb], a, F1[kk]);
        }
      }
    }
    d4_t qr = {0, 0, 0, 0}, qi = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < KX; kk++)
      if (4 * kk < N2) cmfma(qr, qi, F2[kk], make_double2(xr[kk / 4][kk % 4], xi[kk / 4][kk % 4]));
    double2 *Sz = S + (long)z * PW_BOX * PW_BOX;
#pragma unroll
    for (int r = 0; r < 4; r++) Sz[(lk + 4 * r) * PW_BOX + lr] = make_double2(qr[r], qi[r]);
  }
}

// Fused forward transform of the products of pairs (b1, b2): passes x and y
// on MFMA as in pw_pair_xy_kernel, then per plane the 16x16 result goes
// through a per-wave LDS tile into the z pass, accumulated directly at the
// mixed-basis IPWs each lane owns (ig = lane + 64 j), with twiddles
// e^{-2 pi i gz z/N3} from an LDS table.  No slab round trip through HBM.
// Requires ngptm <= 64*PW_FOUT.
constexpr int PW_FOUT = 16;
template <int NB>
__global__ __launch_bounds__(NT) void pw_pair_fused_kernel(
    const double2 *__restrict__ grids, int nocc, int nunocc, int N1, int N2, int N3,
    const int *__restrict__ oidx, int ngptm, double2 *__restrict__ M, long ldm, long t0, int col0) {
  constexpr int KX = 4 * NB;
  __shared__ double2 tz[64][PW_BOX];               // e^{-2 pi i (g-8) z/N3}
  __shared__ double2 Qs[NT / 64][PW_BOX * PW_BOX];  // per-wave plane result
  // MFMA operands of the x and y passes, per lane: e^{-2 pi i (lr-8) x/N}
  // with x = 4kk + lk (kept in LDS rather than registers, for occupancy)
  __shared__ double2 F1s[KX][64], F2s[KX][64];
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  for (int e = tid; e < N3 * PW_BOX; e += NT) {
    const int z = e / PW_BOX, g = e % PW_BOX;
    tz[z][g] = twid((long)(g - PW_OFF) * z, N3, -1.0);
  }
  for (int e = tid; e < KX * 64; e += NT) {
    const int kk = e >> 6, ln = e & 63, x = 4 * kk + (ln >> 4), g = (ln & 15) - PW_OFF;
    F1s[kk][ln] = x < N1 ? twid((long)g * x, N1, -1.0) : make_double2(0.0, 0.0);
    F2s[kk][ln] = x < N2 ? twid((long)g * x, N2, -1.0) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int b2 = blockIdx.x;
  const int b1 = blockIdx.y * (NT / 64) + wv;
  if (b1 >= nocc) return;  // whole wave; no workgroup barrier after this point
  const long NN = (long)N1 * N2 * N3;
  const double2 *p1 = grids + (long)b1 * NN;
  const double2 *p2 = grids + (long)(nocc + b2) * NN;
  int ob[PW_FOUT];
  double2 acc[PW_FOUT];
#pragma unroll
  for (int j = 0; j < PW_FOUT; j++) {
    const int ig = l + 64 * j;
    ob[j] = ig < ngptm ? oidx[ig] : -1;
    acc[j] = make_double2(0.0, 0.0);
  }
  double2 *Q = Qs[wv];
  for (int z = 0; z < N3; z++) {
    d4_t xr[NB], xi[NB];
#pragma unroll
    for (int mb = 0; mb < NB; mb++) {
      xr[mb] = d4_t{0, 0, 0, 0};
      xi[mb] = d4_t{0, 0, 0, 0};
      if (16 * mb >= N2) continue;
      const int y = 16 * mb + lr;
#pragma unroll
      for (int kk = 0; kk < KX; kk++) {
        if (4 * kk < N1) {
          const int x = 4 * kk + lk;
          double2 a = make_double2(0.0, 0.0);
          if (x < N1 && y < N2) {
            const long g = x + (long)N1 * (y + (long)N2 * z);
            a = cmul(p1[g], p2[g]);
          }
          cmfma(xr[mb], xi[mb], a, F1s[kk][l]);
        }
      }
    }
    d4_t qr = {0, 0, 0, 0}, qi = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < KX; kk++)
      if (4 * kk < N2) cmfma(qr, qi, F2s[kk][l], make_double2(xr[kk / 4][kk % 4], xi[kk / 4][kk % 4]));
    // Q(gy, gx): lane holds gx = lr, gy = lk + 4r
#pragma unroll
    for (int r = 0; r < 4; r++) Q[(lk + 4 * r) * PW_BOX + lr] = make_double2(qr[r], qi[r]);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
    for (int j = 0; j < PW_FOUT; j++) {
      if (ob[j] >= 0) {
        const double2 t = tz[z][ob[j] >> 8];
        const double2 v = Q[ob[j] & 255];
        acc[j].x += t.x * v.x - t.y * v.y;
        acc[j].y += t.x * v.y + t.y * v.x;
      }
    }
    __builtin_amdgcn_wave_barrier();  // Q is rewritten by the next plane
  }
  double2 *out = M + (t0 + (long)b1 * nunocc + b2) * ldm + col0;
#pragma unroll
  for (int j = 0; j < PW_FOUT; j++) {
    const int ig = l + 64 * j;
    if (ig < ngptm) out[ig] = acc[j];
  }
}

constexpr int gcd16(int n) { return (n % 16 == 0) ? 16 : (n % 8 == 0) ? 8 : (n % 4 == 0) ? 4 : (n % 2 == 0) ? 2 : 1; }

// Planes packed per x pass of pw_pair_joint_kernel: all 16/gcd(N2,16)
// planes that fill whole row blocks when that is at most five blocks, else
// two planes with the last block partly padded (28: 56 of 64 rows, 40: 80
// in 5).  36 takes one plane in three blocks (36 of 48 rows): then the
// three-product complex MFMA fits the registers, 14 % fewer MFMAs per plane
// than two planes in five blocks with four products.
constexpr int joint_planes(int n2) {
  return n2 == 36 ? 1 : (16 / gcd16(n2)) * n2 / 16 <= 5 ? 16 / gcd16(n2) : 2;
}
// Waves per SIMD: two (256 registers).  The x pass keeps the operands of
// its next step in flight while the current step's MFMAs run, which needs
// more registers than three waves per SIMD leave; on the 24^3 and 36^3
// grids this beats three waves without prefetch by 6 % and 32 %.
constexpr int joint_waves(int) { return 2; }

// pw_pair_fused_kernel for compile-time in-plane sizes N1 x N2: the x pass
// covers G consecutive z planes at once, with the G*N2 rows (y, z) packed
// into RB = ceil(G*N2/16) MFMA row blocks (a 24-point grid: 2 planes = 3
// blocks instead of 2 x 2 half-empty ones; rows past G*N2 read as zero).
// Row R of the packed x-pass result is (y, zz) = (R % N2, R / N2); the y
// pass of plane zz takes K-step kq (y = 4kq..4kq+3) from row block
// (N2*zz + 4kq)/16, register ((N2*zz + 4kq) % 16)/4 -- the accumulator
// layout (rows lane/16 + 4r) is again the operand layout.
template <int N1, int N2>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(joint_waves(N2), joint_waves(N2)))) void pw_pair_joint_kernel(
    const double2 *__restrict__ grids, int nocc, int nunocc, int N3, const int *__restrict__ oidx,
    int ngptm, double2 *__restrict__ M, long ldm, long t0, int col0) {
  static_assert(N1 % 4 == 0 && N2 % 4 == 0 && N1 >= PW_BOX && N2 >= PW_BOX, "grid");
  constexpr int G = joint_planes(N2), RB = (G * N2 + 15) / 16, KX = N1 / 4, KY = N2 / 4;
  constexpr int GI = G % 2 == 0 ? 2 : 1;  // planes per y-pass batch
  constexpr bool M3 = RB <= 3;            // three-product complex MFMA where registers allow
  // z twiddles, (N3 + G) x PW_BOX, zero past N3 (dynamic: sized to the grid
  // instead of a 64-plane table)
  extern __shared__ double2 tz[];
  __shared__ double2 Qs[NT / 64][GI * PW_BOX * PW_BOX];
  __shared__ double2 F1s[KX][64], F2s[KY][64];
  // output box indices in LDS rather than 16 registers per lane: with the
  // register cap of three waves per SIMD the kernel then fits without spills
  __shared__ int obs[64 * PW_FOUT];
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  // Workgroup -> (unoccupied state b2, group of four occupied states): the
  // ngrp workgroups of one b2 get ids that are equal mod 8, so they run on
  // one XCD (ids go round-robin over the 8) and read psi2 through one L2.
  const int ngrp = (nocc + 3) / 4;
  const int xcd = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int b2 = (r / ngrp) * 8 + xcd, grp = r % ngrp;
  if (b2 >= nunocc) return;  // whole workgroup, before any barrier
  for (int e = tid; e < (N3 + G) * PW_BOX; e += NT) {
    const int z = e / PW_BOX, g = e % PW_BOX;
    tz[e] = z < N3 ? twid((long)(g - PW_OFF) * z, N3, -1.0) : make_double2(0.0, 0.0);
  }
  for (int e = tid; e < 64 * PW_FOUT; e += NT) obs[e] = e < ngptm ? oidx[e] : 0;
  for (int e = tid; e < KX * 64; e += NT) {
    const int kk = e >> 6, ln = e & 63;
    F1s[kk][ln] = twid((long)((ln & 15) - PW_OFF) * (4 * kk + (ln >> 4)), N1, -1.0);
  }
  for (int e = tid; e < KY * 64; e += NT) {
    const int kk = e >> 6, ln = e & 63;
    F2s[kk][ln] = twid((long)((ln & 15) - PW_OFF) * (4 * kk + (ln >> 4)), N2, -1.0);
  }
  __syncthreads();
  const int b1 = grp * (NT / 64) + wv;
  if (b1 >= nocc) return;  // whole wave; no workgroup barrier after this point
  const long NN = (long)N1 * N2 * N3;
  const double2 *p1 = grids + (long)b1 * NN;
  const double2 *p2 = grids + (long)(nocc + b2) * NN;
  double2 acc[PW_FOUT];
#pragma unroll
  for (int j = 0; j < PW_FOUT; j++) acc[j] = make_double2(0.0, 0.0);
  double2 *Q = Qs[wv];
  // Row offsets of this lane's x-pass rows (clamped to row 0 where the row
  // lies past the packed planes or past N3, then masked to zero), so every
  // load is unconditional and the loads of step kk+1 are in flight while
  // step kk's MFMAs run.
  // (32-bit byte offsets from the wave-uniform state bases: the loads take
  // the scalar-base form and need no 64-bit address registers)
  unsigned rowoff[RB];
  bool rowok[RB];
  auto rows_of = [&](int z0) __attribute__((always_inline)) {
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      const int R = 16 * rb + lr, y = R % N2, z = z0 + R / N2;
      rowok[rb] = R < G * N2 && z < N3;
      rowoff[rb] = 16u * (unsigned)(rowok[rb] ? N1 * (y + N2 * z) + lk : lk);
    }
  };
  const char *q1 = (const char *)p1, *q2 = (const char *)p2;
  double2 c1[RB], c2[RB];
  auto load_step = [&](int kk, double2 *d1, double2 *d2) __attribute__((always_inline)) {
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      d1[rb] = *(const double2 *)(q1 + (rowoff[rb] + 64u * kk));
      d2[rb] = *(const double2 *)(q2 + (rowoff[rb] + 64u * kk));
    }
  };
  rows_of(0);
  load_step(0, c1, c2);
  for (int z0 = 0; z0 < N3; z0 += G) {
    // M3: three real MFMAs per complex product (Gauss): t1 = ar.fr,
    // t2 = ai.fi, t3 = (ar + ai)(fr + fi); re = t1 - t2, im = t3 - t1 - t2
    d4_t xr[RB], xi[RB], xs[M3 ? RB : 1];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      xr[rb] = d4_t{0, 0, 0, 0};
      xi[rb] = d4_t{0, 0, 0, 0};
      if (M3) xs[M3 ? rb : 0] = d4_t{0, 0, 0, 0};
    }
#pragma unroll
    for (int kk = 0; kk < KX; kk++) {
      const double2 f = F1s[kk][l];
      const double fsum = f.x + f.y;
      double2 a[RB];
#pragma unroll
      for (int rb = 0; rb < RB; rb++)
        a[rb] = rowok[rb] ? cmul(c1[rb], c2[rb]) : make_double2(0.0, 0.0);
      if (kk + 1 < KX) {
        load_step(kk + 1, c1, c2);
      } else if (z0 + G < N3) {  // first step of the next plane group
        rows_of(z0 + G);
        load_step(0, c1, c2);
      }
#pragma unroll
      for (int rb = 0; rb < RB; rb++) {
        if (M3) {
          xr[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb].x, f.x, xr[rb], 0, 0, 0);
          xi[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb].y, f.y, xi[rb], 0, 0, 0);
          xs[M3 ? rb : 0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb].x + a[rb].y, fsum,
                                                                  xs[M3 ? rb : 0], 0, 0, 0);
        } else {
          cmfma(xr[rb], xi[rb], a[rb], f);
        }
      }
    }
    if (M3) {
#pragma unroll
      for (int rb = 0; rb < RB; rb++) {
        const d4_t t1 = xr[rb], t2 = xi[rb];
        xr[rb] = t1 - t2;
        xi[rb] = xs[M3 ? rb : 0] - t1 - t2;
      }
    }
    // y pass of GI planes at a time, their MFMA chains interleaved; then the
    // z pass of those planes in one sweep over the outputs
#pragma unroll
    for (int zb = 0; zb < G; zb += GI) {
      d4_t qr[GI], qi[GI], qs[M3 ? GI : 1];
#pragma unroll
      for (int zi = 0; zi < GI; zi++) {
        qr[zi] = d4_t{0, 0, 0, 0};
        qi[zi] = d4_t{0, 0, 0, 0};
        if (M3) qs[M3 ? zi : 0] = d4_t{0, 0, 0, 0};
      }
#pragma unroll
      for (int kq = 0; kq < KY; kq++) {
        const double2 f2 = F2s[kq][l];
        const double f2sum = f2.x + f2.y;
#pragma unroll
        for (int zi = 0; zi < GI; zi++) {
          const int R0 = N2 * (zb + zi) + 4 * kq, rb = R0 / 16, r = (R0 % 16) / 4;
          const double br = xr[rb][r], bi = xi[rb][r];
          if (M3) {
            qr[zi] = __builtin_amdgcn_mfma_f64_16x16x4f64(f2.x, br, qr[zi], 0, 0, 0);
            qi[zi] = __builtin_amdgcn_mfma_f64_16x16x4f64(f2.y, bi, qi[zi], 0, 0, 0);
            qs[M3 ? zi : 0] = __builtin_amdgcn_mfma_f64_16x16x4f64(f2sum, br + bi,
                                                                    qs[M3 ? zi : 0], 0, 0, 0);
          } else {
            cmfma(qr[zi], qi[zi], f2, make_double2(br, bi));
          }
        }
      }
      if (M3) {
#pragma unroll
        for (int zi = 0; zi < GI; zi++) {
          const d4_t u1 = qr[zi], u2 = qi[zi];
          qr[zi] = u1 - u2;
          qi[zi] = qs[M3 ? zi : 0] - u1 - u2;
        }
      }
      // Q(gy, gx) of plane zi: lane holds gx = lr, gy = lk + 4r
#pragma unroll
      for (int zi = 0; zi < GI; zi++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          Q[zi * PW_BOX * PW_BOX + (lk + 4 * r) * PW_BOX + lr] = make_double2(qr[zi][r], qi[zi][r]);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      // Branch-free: slots past ngptm hold index 0 and accumulate values
      // that are never stored, so the LDS reads of all slots can be in
      // flight together.
      // Planes past N3 hold Q = 0 (their x-pass rows are masked) and the
      // twiddle table is zero-padded past N3, so they add zero.
      // Slots are taken four at a time, their LDS reads issued together.
      const double2 *tzp = tz + (z0 + zb) * PW_BOX;
#pragma unroll
      for (int j0 = 0; j0 < PW_FOUT; j0 += 4) {
        double2 T[4][GI], V[4][GI];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int o = obs[l + 64 * (j0 + j)];
          const int gz = o >> 8, c = o & 255;
#pragma unroll
          for (int zi = 0; zi < GI; zi++) {
            T[j][zi] = tzp[zi * PW_BOX + gz];
            V[j][zi] = Q[zi * PW_BOX * PW_BOX + c];
          }
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int zi = 0; zi < GI; zi++) {
            const double2 t = T[j][zi], v = V[j][zi];
            double2 &a = acc[j0 + j];
            a.x = fma(t.x, v.x, a.x);
            a.x = fma(-t.y, v.y, a.x);
            a.y = fma(t.x, v.y, a.y);
            a.y = fma(t.y, v.x, a.y);
          }
      }
      __builtin_amdgcn_wave_barrier();  // Q is rewritten by the next planes
    }
  }
  double2 *out = M + (t0 + (long)b1 * nunocc + b2) * ldm + col0;
#pragma unroll
  for (int j = 0; j < PW_FOUT; j++) {
    const int ig = l + 64 * j;
    if (ig < ngptm) out[ig] = acc[j];
  }
}

// Forward pass z and gather of the mixed-basis IPWs (oidx: box index
// gz*256 + gy*16 + gx of imati.(Gm - g_k) per IPW).
__global__ __launch_bounds__(NT) void pw_pair_z_kernel(const double2 *__restrict__ slab,
                                                       int nunocc, int N3,
                                                       const int *__restrict__ oidx, int ngptm,
                                                       double2 *__restrict__ M, long ldm, long t0,
                                                       int col0) {
  __shared__ double2 tz[64][PW_BOX];  // N3 <= 64
  const int tid = threadIdx.x;
  for (int e = tid; e < N3 * PW_BOX; e += NT) {
    const int z = e / PW_BOX, g = e % PW_BOX;
    tz[z][g] = twid((long)(g - PW_OFF) * z, N3, -1.0);
  }
  __syncthreads();
  const long pair = blockIdx.x;  // b1*nunocc + b2
  const double2 *S = slab + pair * N3 * PW_BOX * PW_BOX;
  double2 *out = M + (t0 + pair) * ldm + col0;
  for (int ig = tid; ig < ngptm; ig += NT) {
    const int b = oidx[ig];
    const int gz = b >> 8, gyx = b & 255;
    double ar = 0.0, ai = 0.0;
    for (int z = 0; z < N3; z++) {
      const double2 t = tz[z][gz];
      const double2 v = S[(long)z * PW_BOX * PW_BOX + gyx];
      ar += t.x * v.x - t.y * v.y;
      ai += t.x * v.y + t.y * v.x;
    }
    out[ig] = make_double2(ar, ai);
  }
}

int box_index(const int *imati, const int *v, int &idx, int *gout = nullptr) {
  int g3[3];
  int *g = gout ? gout : g3;
  for (int i = 0; i < 3; i++) {
    g[i] = imati[3 * i] * v[0] + imati[3 * i + 1] * v[1] + imati[3 * i + 2] * v[2];
    if (g[i] < -PW_OFF || g[i] >= PW_BOX - PW_OFF) return 1;
  }
  idx = ((g[2] + PW_OFF) * PW_BOX + (g[1] + PW_OFF)) * PW_BOX + (g[0] + PW_OFF);
  return 0;
}

template <int NB>
int launch_pw(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cpw, int nband,
              int nk, const int *k1, const int *k2, const int *oidx_dev, int b1lo, int nocc,
              int b2lo, int nunocc, void *M, long ldm, int col0, void *ws) {
  const int N1 = p->nfft[0], N2 = p->nfft[1], N3 = p->nfft[2];
  const long NN = (long)N1 * N2 * N3;
  double2 *grids = (double2 *)ws;
  double2 *slab = grids + (long)(nocc + nunocc) * NN;
  const int *ngpt_dev = p->gidx + (long)p->nkpt * p->maxgpt;
  const bool joint = p->ngptm <= 64 * PW_FOUT && N1 == N2 &&
                     (N1 == 16 || N1 == 20 || N1 == 24 || N1 == 28 || N1 == 32 || N1 == 36 ||
                      N1 == 40 || N1 == 48);
  // The joint kernel's last round of workgroups leaves most CUs idle (1,600
  // workgroups at 16 x 400 states are 3.1 rounds of 512).  When it runs on
  // the context stream, outside graph capture, the k points alternate
  // between that stream and the auxiliary one, each with its own grids (the
  // second set in the slab region, which the joint kernel does not use), so
  // the transforms and pair kernel of k+1 fill the tail of k's.
  bool dual = joint && nk >= 2 && st == ctx->stream &&
              (long)nocc * nunocc * N3 * PW_BOX * PW_BOX >= (long)(nocc + nunocc) * NN;
  if (dual) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    SPX_HIP(ctx, hipStreamIsCapturing(st, &cs));
    dual = cs == hipStreamCaptureStatusNone;
  }
  if (dual) {
    int rc = spx_aux_stream(ctx);
    if (rc) return rc;
    SPX_HIP(ctx, hipEventRecord(ctx->aux_ev[7], st));
    SPX_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->aux_ev[7], 0));
  }
  for (int ik = 0; ik < nk; ik++) {
    const hipStream_t s = dual && (ik & 1) ? ctx->aux : st;
    double2 *gk = dual && (ik & 1) ? slab : grids;
    hipLaunchKernelGGL(pw_inverse_kernel<NB>, dim3(nocc + nunocc), dim3(NT), 0, s,
                       (const double2 *)cpw, p->maxgpt, nband, p->gidx, ngpt_dev, k1[ik], b1lo,
                       nocc, k2[ik], b2lo, N1, N2, N3, 1.0 / std::sqrt(p->vol), 1.0 / (double)NN,
                       p->step, gk);
    SPX_LAUNCHED(ctx);
    if (joint) {
      const int ngrp = (nocc + 3) / 4;
      const dim3 grid((unsigned)(((nunocc + 7) / 8) * 8 * ngrp));
      const size_t tzb = (size_t)(N3 + 4) * PW_BOX * sizeof(double2);  // G <= 4 padding planes
      const int *ox = oidx_dev + (long)ik * p->ngptm;
      const long tt = (long)ik * nocc * nunocc;
      switch (N1) {
        case 16: hipLaunchKernelGGL((pw_pair_joint_kernel<16, 16>), grid, dim3(NT), tzb, s, gk, nocc, nunocc, N3, ox, p->ngptm, (double2 *)M, ldm, tt, col0); break;
        case 20: hipLaunchKernelGGL((pw_pair_joint_kernel<20, 20>), grid, dim3(NT), tzb, s, gk, nocc, nunocc, N3, ox, p->ngptm, (double2 *)M, ldm, tt, col0); break;
        case 24: hipLaunchKernelGGL((pw_pair_joint_kernel<24, 24>), grid, dim3(NT), tzb, s, gk, nocc, nunocc, N3, ox, p->ngptm, (double2 *)M, ldm, tt, col0); break;
        case 28: hipLaunchKernelGGL((pw_pair_joint_kernel<28, 28>), grid, dim3(NT), tzb, s, gk, nocc, nunocc, N3, ox, p->ngptm, (double2 *)M, ldm, tt, col0); break;
        case 32: hipLaunchKernelGGL((pw_pair_joint_kernel<32, 32>), grid, dim3(NT), tzb, s, gk, nocc, nunocc, N3, ox, p->ngptm, (double2 *)M, ldm, tt, col0); break;
        case 36: hipLaunchKernelGGL((pw_pair_joint_kernel<36, 36>), grid, dim3(NT), tzb, s, gk, nocc, nunocc, N3, ox, p->ngptm, (double2 *)M, ldm, tt, col0); break;
        case 40: hipLaunchKernelGGL((pw_pair_joint_kernel<40, 40>), grid, dim3(NT), tzb, s, gk, nocc, nunocc, N3, ox, p->ngptm, (double2 *)M, ldm, tt, col0); break;
        default: hipLaunchKernelGGL((pw_pair_joint_kernel<48, 48>), grid, dim3(NT), tzb, s, gk, nocc, nunocc, N3, ox, p->ngptm, (double2 *)M, ldm, tt, col0); break;
      }
      SPX_LAUNCHED(ctx);
      continue;
    }
    if (p->ngptm <= 64 * PW_FOUT) {
      hipLaunchKernelGGL(pw_pair_fused_kernel<NB>, dim3(nunocc, (nocc + 3) / 4), dim3(NT), 0, st,
                         grids, nocc, nunocc, N1, N2, N3, oidx_dev + (long)ik * p->ngptm, p->ngptm,
                         (double2 *)M, ldm, (long)ik * nocc * nunocc, col0);
      SPX_LAUNCHED(ctx);
      continue;
    }
    hipLaunchKernelGGL(pw_pair_xy_kernel<NB>, dim3(nunocc, (nocc + 3) / 4), dim3(NT), 0, st, grids,
                       nocc, nunocc, N1, N2, N3, slab);
    SPX_LAUNCHED(ctx);
    hipLaunchKernelGGL(pw_pair_z_kernel, dim3(nocc * nunocc), dim3(NT), 0, st, slab, nunocc, N3,
                       oidx_dev + (long)ik * p->ngptm, p->ngptm, (double2 *)M, ldm,
                       (long)ik * nocc * nunocc, col0);
    SPX_LAUNCHED(ctx);
  }
  if (dual) {
    SPX_HIP(ctx, hipEventRecord(ctx->aux_ev[8], ctx->aux));
    SPX_HIP(ctx, hipStreamWaitEvent(st, ctx->aux_ev[8], 0));
  }
  return SPEXHIP_OK;
}

}  // namespace

int spx_pw_plan_build(spexhip_ctx *ctx, const spexhip_pw_desc *d, spexhip_plan *p) {
  p->kind = 2;
  for (int i = 0; i < 3; i++) {
    SPX_CHECK(ctx, d->nfft[i] >= PW_BOX && d->nfft[i] <= 64,
              "pw plan: nfft[%d]=%d outside [16, 64]", i, d->nfft[i]);
    p->nfft[i] = d->nfft[i];
  }
  for (int i = 0; i < 9; i++) p->imati[i] = d->imati[i];
  SPX_CHECK(ctx, d->vol > 0 && d->nkpt > 0 && d->maxgpt > 0 && d->ngptm >= 0 && d->gpt && d->ngpt,
            "pw plan: bad descriptor");
  p->vol = d->vol;
  p->nkpt = d->nkpt;
  p->maxgpt = d->maxgpt;
  p->ngptm = d->ngptm;
  p->ngpt.assign(d->ngpt, d->ngpt + d->nkpt);
  p->gptm.assign(d->gptm, d->gptm + 3 * d->ngptm);
  std::vector<int> tab((size_t)d->nkpt * d->maxgpt + d->nkpt, 0);
  int lo[3] = {PW_BOX, PW_BOX, PW_BOX}, hi[3] = {-PW_BOX, -PW_BOX, -PW_BOX};
  for (int k = 0; k < d->nkpt; k++) {
    SPX_CHECK(ctx, d->ngpt[k] >= 0 && d->ngpt[k] <= d->maxgpt, "pw plan: ngpt[%d] out of range", k);
    for (int ig = 0; ig < d->ngpt[k]; ig++) {
      int idx, g[3];
      if (box_index(p->imati, d->gpt + 3 * ((long)k * d->maxgpt + ig), idx, g))
        return spx_seterr(ctx, SPEXHIP_E_INVALID,
                          "pw plan: G vector %d of k %d outside the -8..7 Fourier box", ig, k);
      tab[(size_t)k * d->maxgpt + ig] = idx;
      for (int i = 0; i < 3; i++) {
        lo[i] = std::min(lo[i], g[i]);
        hi[i] = std::max(hi[i], g[i]);
      }
    }
    tab[(size_t)d->nkpt * d->maxgpt + k] = d->ngpt[k];
  }
  // Without theta the pair product is a pure convolution of two states with
  // Fourier support of width W = hi - lo per axis, i.e. product frequencies
  // in [-W, W]; the outputs lie in the -8..7 box.  On a grid with
  // N >= W + 9 no product frequency aliases onto an output, so the DFT gives
  // the exact convolution -- the same numbers on every such grid, the
  // caller's included.  The plan then uses the smallest such grid (at least
  // the 16-point box); a caller grid that aliases is kept as given.
  // Sizes are rounded up to 16, 20, 24 or 32 (the in-plane sizes with a
  // specialised pair kernel) when that is still below the caller's grid.
  const bool theta = d->step_r || d->step_c;
  if (!theta)
    for (int i = 0; i < 3; i++) {
      int nmin = std::max(PW_BOX, hi[i] >= lo[i] ? hi[i] - lo[i] + PW_OFF + 1 : PW_BOX);
      for (int nice : {16, 20, 24, 32})
        if (nmin <= nice && nice <= p->nfft[i]) {
          nmin = nice;
          break;
        }
      if (p->nfft[i] > nmin) p->nfft[i] = nmin;
    }
  int rc = spx_alloc(ctx, sizeof(int) * tab.size(), (void **)&p->gidx);
  if (rc) return rc;
  p->owned.push_back(p->gidx);
  SPX_HIP(ctx, hipMemcpy(p->gidx, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
  if (theta) {
    const size_t nn = (size_t)p->nfft[0] * p->nfft[1] * p->nfft[2];
    std::vector<double2> th(nn);
    for (size_t i = 0; i < nn; i++)
      th[i] = d->step_c ? make_double2(d->step_c[2 * i], d->step_c[2 * i + 1])
                        : make_double2(d->step_r[i], 0.0);
    rc = spx_alloc(ctx, 16 * nn, (void **)&p->step);
    if (rc) return rc;
    p->owned.push_back(p->step);
    SPX_HIP(ctx, hipMemcpy(p->step, th.data(), 16 * nn, hipMemcpyHostToDevice));
  }
  return SPEXHIP_OK;
}

size_t spx_wfprod_pw_ws_core(const spexhip_plan *p, int nocc, int nunocc);
size_t spx_wfprod_pw_ws(const spexhip_plan *p, int nk, int nocc, int nunocc) {
  return spx_wfprod_pw_ws_core(p, nocc, nunocc) + 4 * (size_t)nk * p->ngptm + 256;
}

// Box indices of imati.(Gm - g2(k)) for every packet k and IPW (host).
int spx_pw_oidx(spexhip_ctx *ctx, const spexhip_plan *p, int nk, const int *g2, int *oidx) {
  for (int ik = 0; ik < nk; ik++)
    for (int ig = 0; ig < p->ngptm; ig++) {
      int v[3];
      for (int i = 0; i < 3; i++) v[i] = p->gptm[3 * ig + i] - g2[3 * ik + i];
      int idx;
      if (box_index(p->imati, v, idx))
        return spx_seterr(ctx, SPEXHIP_E_INVALID,
                          "wfprod_pw: IPW %d - umklapp of k %d outside the -8..7 Fourier box", ig, ik);
      oidx[(size_t)ik * p->ngptm + ig] = idx;
    }
  return SPEXHIP_OK;
}

// Workspace without the oidx table.
size_t spx_wfprod_pw_ws_core(const spexhip_plan *p, int nocc, int nunocc) {
  const long NN = (long)p->nfft[0] * p->nfft[1] * p->nfft[2];
  return 16 * ((size_t)(nocc + nunocc) * NN + (size_t)nocc * nunocc * p->nfft[2] * PW_BOX * PW_BOX);
}

// Asynchronous launches (no host synchronisation): k1, k2 are host arrays
// used as launch arguments, oidx_dev a device table from spx_pw_oidx.
int spx_wfprod_pw_async(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cpw,
                        int nband, int nk, const int *k1, const int *k2, const int *oidx_dev,
                        int b1lo, int nocc, int b2lo, int nunocc, void *M, long ldm, int col0,
                        void *ws) {
  for (int ik = 0; ik < nk; ik++)
    SPX_CHECK(ctx, k1[ik] >= 0 && k1[ik] < p->nkpt && k2[ik] >= 0 && k2[ik] < p->nkpt,
              "wfprod_pw: k index out of range");
  const int nbmax = (std::max(p->nfft[0], p->nfft[1]) + 15) / 16;
  switch (nbmax) {
    case 1:
    case 2:
      return launch_pw<2>(ctx, st, p, cpw, nband, nk, k1, k2, oidx_dev, b1lo, nocc, b2lo, nunocc, M, ldm, col0, ws);
    case 3:
      return launch_pw<3>(ctx, st, p, cpw, nband, nk, k1, k2, oidx_dev, b1lo, nocc, b2lo, nunocc, M, ldm, col0, ws);
    default:
      return launch_pw<4>(ctx, st, p, cpw, nband, nk, k1, k2, oidx_dev, b1lo, nocc, b2lo, nunocc, M, ldm, col0, ws);
  }
}

// Synchronous form: builds and uploads the oidx table into the tail of ws.
int spx_wfprod_pw(spexhip_ctx *ctx, hipStream_t st, const spexhip_plan *p, const void *cpw,
                  int nband, int nk, const int *k1, const int *k2, const int *g2, int b1lo,
                  int nocc, int b2lo, int nunocc, void *M, long ldm, int col0, void *ws) {
  std::vector<int> oidx((size_t)nk * p->ngptm);
  int rc = spx_pw_oidx(ctx, p, nk, g2, oidx.data());
  if (rc) return rc;
  int *oidx_dev = (int *)((char *)ws + spx_wfprod_pw_ws_core(p, nocc, nunocc));
  SPX_HIP(ctx, hipMemcpyAsync(oidx_dev, oidx.data(), 4 * oidx.size(), hipMemcpyHostToDevice, st));
  rc = spx_wfprod_pw_async(ctx, st, p, cpw, nband, nk, k1, k2, oidx_dev, b1lo, nocc, b2lo, nunocc,
                           M, ldm, col0, ws);
  SPX_HIP(ctx, hipStreamSynchronize(st));  // oidx is a host temporary
  return rc;
}
