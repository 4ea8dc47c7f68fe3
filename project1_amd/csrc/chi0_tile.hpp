// One output tile of the chi0 rank-k update (FP64 or split-FP32 MFMA),
// shared by the single-problem kernels (chi0.hip) and the persistent
// grouped kernel over many q-points (chi0_grouped.hip).
//
// chi(I,J,w) = alpha * sum_t wght(t,w) M(t,I) conj(M(t,J)) + beta * chi
// (susceptibility.f:1183-1240, HILBERT NONE spectral function :1780-1788).
//
// Tile shape: TM x TM complex outputs per 256-thread workgroup, 4 waves in a
// 2x2 layout, each wave owning (TM/32)^2 blocks of 16x16; K-slices of KC
// transitions staged through double-buffered LDS with the weight applied
// while staging.
//
// Split-FP32 form.  With B(t,I) = sqrt|D(t,w)| M(t,I) and s(I,w) its column
// maximum (colscale.hpp), every real component is staged as
//     u = B * R / s(I,w),  |u| <= R = 512,   h = rint(u),  l = fp32(u - h),
// the row side carrying the sign of D, so that chi = s_I s_J / R^2 *
// sum_t sgn(D) u_I conj(u_J).  h is an integer of at most 10 bits on a grid
// common to all transitions of a column, so h_I*h_J and their sums over a
// K-slice (2 KC products of magnitude <= R^2 per real accumulator, <= 2^24)
// are exact in FP32.  The remaining part h_I*l_J + l_I*x_J, x = fp32(h + l)
// ~ fp32(u), is 2^-9 smaller and goes into a second FP32 accumulator; both
// are flushed into FP64 after every K-slice.  Small entries keep full FP32
// relative precision in l, and no term of the product is dropped beyond
// the FP32 rounding of the cross term: 12 real MFMAs per complex step, as
// for the plain hi/lo split, but without its O(2^-24) error on hi*hi.
#pragma once
#include "common.hpp"

namespace chi0t {

constexpr int NT = 256;

__device__ inline void tri_index(int x, int &bi, int &bj) {
  int j = (int)((sqrtf(8.f * (float)x + 1.f) - 1.f) * 0.5f);
  while ((j + 1) * (j + 2) / 2 <= x) j++;
  while (j * (j + 1) / 2 > x) j--;
  bj = j;
  bi = x - j * (j + 1) / 2;
}

constexpr double SPLIT_R = 512.0;  // grid levels of h per column maximum

// u = x * inv (|u| <= SPLIT_R) -> h = rint(u), l = fp32(u - h) (u - h is
// exact in FP64).
__device__ inline void split_fx(double x, double inv, float &h, float &l) {
  const double u = x * inv;
  const double r = rint(u);
  h = (float)r;
  l = (float)(u - r);
}

template <int TM, int KC, bool SPLIT>
struct Lds;
template <int TM, int KC>
struct Lds<TM, KC, false> {
  double Lr[2][KC][TM], Li[2][KC][TM], Rr[2][KC][TM], Ri[2][KC][TM];
};
template <int TM, int KC>
struct Lds<TM, KC, true> {
  float Lrh[2][KC][TM], Lrl[2][KC][TM], Lih[2][KC][TM], Lil[2][KC][TM];
  float Rrh[2][KC][TM], Rrl[2][KC][TM], Rih[2][KC][TM], Ril[2][KC][TM];
};

struct Problem {
  long n;                 // transitions
  int nb, nw;             // basis size, frequencies
  const double2 *M;       // products
  long ldm;
  const double *wght;     // wght[w*n + t]
  double2 *chi;           // chi[w*ldc*nb + I + ldc*J]
  long ldc;
  double alpha, beta;
  const double *scale;    // split-FP32: s(I,w) at scale[w*nb + I] (colscale.hpp)
};

// Computes tile (bi, bj) of frequency w.  Must be called by all 256 threads
// of the workgroup (it synchronises on LDS).
template <int TM, int KC, bool COLMAJOR, bool SPLIT>
__device__ void tile(const Problem &P, int bi, int bj, int w, Lds<TM, KC, SPLIT> &s) {
  constexpr int WB = TM / 32;             // 16x16 blocks per wave per dimension
  constexpr int PER = KC * TM / NT;       // elements per thread per stage
  static_assert(PER >= 1 && KC % 4 == 0, "tile shape");
  static_assert(2.0 * KC * SPLIT_R * SPLIT_R <= 16777216.0, "h*h sums of a K-slice exact in FP32");
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1, lr = lane & 15, lk = lane >> 4;
  const int I0 = bi * TM, J0 = bj * TM;
  const long n = P.n;
  const int nb = P.nb;
  const double *D = P.wght + (long)w * n;

  d4_t accr[WB][WB], acci[WB][WB];
#pragma unroll
  for (int a = 0; a < WB; a++)
#pragma unroll
    for (int b = 0; b < WB; b++) {
      accr[a][b] = d4_t{0, 0, 0, 0};
      acci[a][b] = d4_t{0, 0, 0, 0};
    }

  double2 va[PER], vb[PER];
  double dv[PER];
  const long nst = (n + KC - 1) / KC;
  auto coords = [&](int e, int &k, int &i) {
    if (COLMAJOR) {
      k = e % KC;
      i = e / KC;
    } else {
      k = e / TM;
      i = e % TM;
    }
  };
  // split form: R / s of the columns this thread stages (0 for padding and
  // all-zero columns)
  double inva[SPLIT ? PER : 1], invb[SPLIT ? PER : 1];
  if constexpr (SPLIT) {
    const double *S = P.scale + (long)w * nb;
#pragma unroll
    for (int r = 0; r < PER; r++) {
      int k, i;
      coords(tid + NT * r, k, i);
      const double sa = I0 + i < nb ? S[I0 + i] : 0.0, sb = J0 + i < nb ? S[J0 + i] : 0.0;
      inva[r] = sa > 0.0 ? SPLIT_R / sa : 0.0;
      invb[r] = sb > 0.0 ? SPLIT_R / sb : 0.0;
    }
  }
  auto fetch = [&](long st) {
    const long t0 = st * KC;
#pragma unroll
    for (int r = 0; r < PER; r++) {
      int k, i;
      coords(tid + NT * r, k, i);
      const long t = t0 + k;
      const int ca = I0 + i, cb = J0 + i;
      double2 za = make_double2(0.0, 0.0), zb = make_double2(0.0, 0.0);
      if (t < n) {
        if (ca < nb) za = COLMAJOR ? P.M[t + P.ldm * (long)ca] : P.M[t * P.ldm + ca];
        if (cb < nb) zb = COLMAJOR ? P.M[t + P.ldm * (long)cb] : P.M[t * P.ldm + cb];
      }
      va[r] = za;
      vb[r] = zb;
      dv[r] = t < n ? D[t] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < PER; r++) {
      int k, i;
      coords(tid + NT * r, k, i);
      if constexpr (SPLIT) {
        // the same sqrt|D| as the scale pass, so |sq*M| <= s bit for bit
        const double sq = sqrt(fabs(dv[r]));
        const double sa = dv[r] < 0.0 ? -sq : sq;
        float h, l;
        split_fx(sa * va[r].x, inva[r], h, l);
        s.Lrh[buf][k][i] = h;
        s.Lrl[buf][k][i] = l;
        split_fx(sa * va[r].y, inva[r], h, l);
        s.Lih[buf][k][i] = h;
        s.Lil[buf][k][i] = l;
        split_fx(sq * vb[r].x, invb[r], h, l);
        s.Rrh[buf][k][i] = h;
        s.Rrl[buf][k][i] = l;
        split_fx(sq * vb[r].y, invb[r], h, l);
        s.Rih[buf][k][i] = h;
        s.Ril[buf][k][i] = l;
      } else {
        s.Lr[buf][k][i] = dv[r] * va[r].x;
        s.Li[buf][k][i] = dv[r] * va[r].y;
        s.Rr[buf][k][i] = vb[r].x;
        s.Ri[buf][k][i] = vb[r].y;
      }
    }
  };

  __syncthreads();  // LDS may still be read by the previous tile
  fetch(0);
  store(0);
  __syncthreads();
  for (long st = 0; st < nst; st++) {
    const int buf = st & 1;
    if (st + 1 < nst) fetch(st + 1);
    if constexpr (SPLIT) {
      f4_t hr[WB][WB], hi_[WB][WB], xr[WB][WB], xi[WB][WB];
#pragma unroll
      for (int a = 0; a < WB; a++)
#pragma unroll
        for (int b = 0; b < WB; b++) {
          hr[a][b] = f4_t{0, 0, 0, 0};
          hi_[a][b] = f4_t{0, 0, 0, 0};
          xr[a][b] = f4_t{0, 0, 0, 0};
          xi[a][b] = f4_t{0, 0, 0, 0};
        }
#pragma unroll
      for (int kk = 0; kk < KC / 4; kk++) {
        const int k = kk * 4 + lk;
        float arh[WB], arl[WB], aih[WB], ail[WB], brh[WB], brl[WB], bih[WB], bil[WB];
        float brx[WB], bix[WB];
#pragma unroll
        for (int a = 0; a < WB; a++) {
          const int ia = wr * (TM / 2) + a * 16 + lr, jb = wc * (TM / 2) + a * 16 + lr;
          arh[a] = s.Lrh[buf][k][ia];
          arl[a] = s.Lrl[buf][k][ia];
          aih[a] = s.Lih[buf][k][ia];
          ail[a] = s.Lil[buf][k][ia];
          brh[a] = s.Rrh[buf][k][jb];
          brl[a] = s.Rrl[buf][k][jb];
          bih[a] = s.Rih[buf][k][jb];
          bil[a] = s.Ril[buf][k][jb];
          brx[a] = brh[a] + brl[a];
          bix[a] = bih[a] + bil[a];
        }
        // Re = ar br + ai bi, Im = ai br - ar bi with a = h + l on the row
        // side: h*h exactly into (hr, hi_), h*l + l*x into (xr, xi)
#pragma unroll
        for (int a = 0; a < WB; a++)
#pragma unroll
          for (int b = 0; b < WB; b++) {
            hr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(arh[a], brh[b], hr[a][b], 0, 0, 0);
            hr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[a], bih[b], hr[a][b], 0, 0, 0);
            hi_[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[a], brh[b], hi_[a][b], 0, 0, 0);
            hi_[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-arh[a], bih[b], hi_[a][b], 0, 0, 0);
            xr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(arh[a], brl[b], xr[a][b], 0, 0, 0);
            xr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(arl[a], brx[b], xr[a][b], 0, 0, 0);
            xr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[a], bil[b], xr[a][b], 0, 0, 0);
            xr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ail[a], bix[b], xr[a][b], 0, 0, 0);
            xi[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[a], brl[b], xi[a][b], 0, 0, 0);
            xi[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ail[a], brx[b], xi[a][b], 0, 0, 0);
            xi[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-arh[a], bil[b], xi[a][b], 0, 0, 0);
            xi[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-arl[a], bix[b], xi[a][b], 0, 0, 0);
          }
      }
      // FP64 correction pass of this K-slice
#pragma unroll
      for (int a = 0; a < WB; a++)
#pragma unroll
        for (int b = 0; b < WB; b++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            accr[a][b][r] += (double)hr[a][b][r] + (double)xr[a][b][r];
            acci[a][b][r] += (double)hi_[a][b][r] + (double)xi[a][b][r];
          }
    } else {
#pragma unroll
      for (int kk = 0; kk < KC / 4; kk++) {
        const int k = kk * 4 + lk;
        double ar[WB], ai[WB], br[WB], bim[WB];
#pragma unroll
        for (int a = 0; a < WB; a++) {
          ar[a] = s.Lr[buf][k][wr * (TM / 2) + a * 16 + lr];
          ai[a] = s.Li[buf][k][wr * (TM / 2) + a * 16 + lr];
          br[a] = s.Rr[buf][k][wc * (TM / 2) + a * 16 + lr];
          bim[a] = s.Ri[buf][k][wc * (TM / 2) + a * 16 + lr];
        }
#pragma unroll
        for (int a = 0; a < WB; a++)
#pragma unroll
          for (int b = 0; b < WB; b++) {
            // (ar + i ai)(br - i bim) = ar br + ai bim + i(ai br - ar bim)
            accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], accr[a][b], 0, 0, 0);
            accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bim[b], accr[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], acci[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[a], bim[b], acci[a][b], 0, 0, 0);
          }
      }
    }
    if (st + 1 < nst) store(buf ^ 1);
    __syncthreads();
  }

  // Result layouts: v_mfma_f64_16x16x4f64 holds column lane%16, rows
  // lane/16 + 4r; v_mfma_f32_16x16x4f32 holds column lane%16, rows
  // 4*(lane/16) + r (the FP64 accumulators of the split form follow it).
  double2 *C = P.chi + (long)w * P.ldc * nb;
#pragma unroll
  for (int a = 0; a < WB; a++)
#pragma unroll
    for (int b = 0; b < WB; b++) {
      const int J = J0 + wc * (TM / 2) + b * 16 + lr;
      if (J >= nb) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int I = I0 + wr * (TM / 2) + a * 16 + (SPLIT ? 4 * lk + r : lk + 4 * r);
        if (I >= nb) continue;
        double2 *p = C + I + P.ldc * (long)J;
        double f = P.alpha;
        if constexpr (SPLIT)
          f *= (P.scale[(long)w * nb + I] * (1.0 / SPLIT_R)) *
               (P.scale[(long)w * nb + J] * (1.0 / SPLIT_R));
        double2 o = make_double2(f * accr[a][b][r], f * acci[a][b][r]);
        if (P.beta != 0.0) {
          const double2 old = *p;
          o.x += P.beta * old.x;
          o.y += P.beta * old.y;
        }
        *p = o;
      }
    }
}

}  // namespace chi0t
