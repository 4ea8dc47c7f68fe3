// Gamma-point (q = 0) head and wings: of chi0 (susceptibility side) and of
// eps and eps^-1 (dielectric side).
//
// Reference, susceptibility side: at q = 0 the Coulomb matrix is singular
// in the direction of q, and the k.p expansion replaces the G = 0 product by
// the momentum matrix.  dcprod(t, 1:3) = <u|p|o> / sqrt(vol) / (e_u - e_o)
// (susceptibility.f:847-873, momentum_matrix in dwavefproducts.f:89), and
// the pair loop adds (susceptibility.f:1092-1157)
//     head(k1,k2,w) = sum_t D(t,w) dcprod(t,k1) conj(dcprod(t,k2)),
//     wing(k2,i,w)  = sum_t D(t,w) cprod(t,i)   conj(dcprod(t,k2)).
// With the augmented matrix [cprod | dcprod] these are the last three
// columns of the same pair sum, so the kernel treats the three dcprod
// columns as extra columns c = nb..nb+2 of M: one thread per column c, the
// transitions of a workgroup staged through LDS 16 at a time (both layouts
// of M are read coalesced into the same [t][c] tile), the three dcprod
// values and the weights of a transition broadcast from LDS, and per thread
// z_k = M(t,c) conj(P(t,k)) followed by the weighted sums over w.  M is read
// from HBM once per pass of up to 16 (6 complex) frequencies: the kernel is
// HBM-bound (16 B per M element per pass against 3 complex products and
// 3 x nw weighted sums).  Partial sums per K-split are reduced in a fixed
// order (deterministic).  The head (three columns, from P alone) has its own
// kernel: one thread per transition, the six products p_k1 conj(p_k2), a
// shuffle and LDS reduction per block and a fixed-order sum over blocks.
//
// Real frequencies (headc, wingc, :1112-1125, :1144-1151) carry complex
// weights Dr + i Di: wingc(k2,i,1) = sum Dc M(t,i) conj(P(t,k2)) and
// wingc(k2,i,2) = sum Dc conj(M(t,i)) P(t,k2) = sum Dc conj(z); the head is
// the first form on the dcprod columns.
//
// Dielectric side (imaginary axis, correlation.f:570-629 and :749-759):
// before eps is formed the first basis function's row and column of chi and
// wing(:,1,:) are set to zero, head = 1 - 4 pi head, wing(:,i,:) =
// -sqrt(4 pi) sqcoul(i) wing(:,i,:); after eps is inverted,
//     wing_new = -wing conj(eps^-1),  head += conj(wing) wing_new^T,
// the Schur complement of the bordered matrix [[head, conj(wing)],
// [wing^T, eps]] and its off-diagonal blocks.
#include "common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>

namespace {

constexpr int HW_NT = 256;  // augmented columns per workgroup, one per thread
constexpr int HW_TT = 16;   // transitions per LDS stage

__global__ void momentum_scale_kernel(int nk, int nocc, int nunocc, const double *__restrict__ eo,
                                      const double *__restrict__ eu, double sv,
                                      double2 *__restrict__ P, long ldp) {
  const long n = (long)nk * nocc * nunocc;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (!eo) {  // MULTDIFF (multdiff_ > 0): the numerator only
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double2 p = P[t + ldp * c];
      P[t + ldp * c] = make_double2(p.x / sv, p.y / sv);
    }
    return;
  }
  const int b2 = (int)(t % nunocc);
  const long r = t / nunocc;
  const int b1 = (int)(r % nocc);
  const int k = (int)(r / nocc);
  const double d = eu[(long)k * nunocc + b2] - eo[(long)k * nocc + b1];
  const bool ok = fabs(d) > 1e-8;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double2 p = P[t + ldp * c];
    // dcprod / sqrt(vol), then / rdum: two roundings, as the reference
    p = ok ? make_double2((p.x / sv) / d, (p.y / sv) / d) : make_double2(0.0, 0.0);
    P[t + ldp * c] = p;
  }
}

// Partial sums of one K-split: part[((s*NWB + w)*3 + k)*NV + v][c] (row
// stride pstride >= ncolM), NV = 1
// (real weights) or 2 (complex: the M conj(P) and conj(M) P forms), for
// the ncolM columns of M.  Each stage of HW_TT transitions is fetched into
// registers (branch-free, indices clamped into the arrays, so the loads of
// a stage are all in flight together) while the previous stage is summed
// from LDS, then committed to LDS between two barriers.  Rows past the end
// of the split get zero weights, so their clamped loads add nothing.
template <bool CPLX, int NWB, bool CM>
__global__ __launch_bounds__(HW_NT) void headwing_kernel(
    long n, int ncolM, const double2 *__restrict__ M, long ldm, const double2 *__restrict__ P,
    long ldp, int nw, int w0, const double *__restrict__ Dr, const double *__restrict__ Di,
    long ldd, long kchunk, long pstride, double2 *__restrict__ part) {
  constexpr int NV = CPLX ? 2 : 1;
  constexpr int ND = CPLX ? 2 : 1;
  constexpr int NDW = HW_TT * ND * NWB;  // weights per stage
  static_assert(NDW <= HW_NT && 3 * HW_TT <= HW_NT, "one weight and one P value per thread");
  __shared__ double2 Ms[HW_TT][HW_NT + 1];  // +1: the colmajor commit hits distinct banks
  __shared__ double2 Ps[HW_TT][3];
  __shared__ double Ds[HW_TT][ND * NWB];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * HW_NT;
  const int c = c0 + tid;
  const int cl = c < ncolM ? c : ncolM - 1;
  const long tb = (long)blockIdx.y * kchunk;
  const long te = tb + kchunk < n ? tb + kchunk : n;
  // this thread's share of the P and weight stage
  const int pj = tid / 3, pk = tid % 3;
  const int dj = tid / (ND * NWB), dd = (tid / NWB) % ND, dw = tid % NWB;
  const double *dsrc = dd ? Di : Dr;
  const long dwc = (long)(w0 + dw < nw ? w0 + dw : nw - 1) * ldd;
  // the in-flight stage, one vector value (an array here would live in
  // scratch memory)
  typedef double stage_t __attribute__((ext_vector_type(2 * HW_TT)));
  stage_t mr;
  double2 pr = make_double2(0.0, 0.0);
  double dr0 = 0.0;
#define HW_CLAMPT(t) ((t) < te ? (t) : te - 1)
// registers <- stage at t0 (P and weights first: their wait must not cover
// the M loads)
#define HW_FETCH(t0_)                                                             \
  do {                                                                            \
    const long f0 = (t0_);                                                        \
    if (tid < 3 * HW_TT) pr = P[HW_CLAMPT(f0 + pj) + ldp * pk];                   \
    if (tid < NDW) dr0 = dsrc[HW_CLAMPT(f0 + dj) + dwc];                          \
    __builtin_amdgcn_sched_barrier(0);                                            \
    _Pragma("unroll") for (int j = 0; j < HW_TT; j++) {                           \
      if constexpr (CM) {                                                         \
        const int e = tid + HW_NT * j;                                            \
        const int col = c0 + e / HW_TT < ncolM ? c0 + e / HW_TT : ncolM - 1;      \
        const double2 v = M[(long)col * ldm + HW_CLAMPT(f0 + e % HW_TT)];         \
        mr[2 * j] = v.x;                                                          \
        mr[2 * j + 1] = v.y;                                                      \
      } else {                                                                    \
        const double2 v = M[HW_CLAMPT(f0 + j) * ldm + cl];                        \
        mr[2 * j] = v.x;                                                          \
        mr[2 * j + 1] = v.y;                                                      \
      }                                                                           \
    }                                                                             \
  } while (0)
// LDS <- registers (between two barriers)
#define HW_COMMIT(t0_)                                                            \
  do {                                                                            \
    const long f0 = (t0_);                                                        \
    if (tid < 3 * HW_TT) Ps[pj][pk] = f0 + pj < te ? pr : make_double2(0.0, 0.0); \
    if (tid < NDW) Ds[dj][dd * NWB + dw] = (f0 + dj < te && w0 + dw < nw) ? dr0 : 0.0; \
    _Pragma("unroll") for (int j = 0; j < HW_TT; j++) {                           \
      if constexpr (CM) {                                                         \
        const int e = tid + HW_NT * j;                                            \
        Ms[e % HW_TT][e / HW_TT] = make_double2(mr[2 * j], mr[2 * j + 1]);        \
      } else {                                                                    \
        Ms[j][tid] = make_double2(mr[2 * j], mr[2 * j + 1]);                      \
      }                                                                           \
    }                                                                             \
  } while (0)
  double2 acc[NWB][3][NV];
#pragma unroll
  for (int w = 0; w < NWB; w++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int v = 0; v < NV; v++) acc[w][k][v] = make_double2(0.0, 0.0);
  const bool live = c0 + (tid & ~63) < ncolM;  // uniform per wave
  if (tb < te) {
    HW_FETCH(tb);
    HW_COMMIT(tb);
  }
  __syncthreads();
  for (long t0 = tb; t0 < te; t0 += HW_TT) {
    const long tn = t0 + HW_TT;
    HW_FETCH(tn < te ? tn : t0);  // the last stage fetches itself again (never committed)
    if (live) {
#pragma unroll 2
      for (int j = 0; j < HW_TT; j++) {
        const double2 m = Ms[j][tid];
        double2 z[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const double2 p = Ps[j][k];
          z[k] = make_double2(m.x * p.x + m.y * p.y, m.y * p.x - m.x * p.y);  // m conj(p)
        }
#pragma unroll
        for (int w = 0; w < NWB; w++) {
          const double dr = Ds[j][w];
          if constexpr (CPLX) {
            const double di = Ds[j][NWB + w];
#pragma unroll
            for (int k = 0; k < 3; k++) {
              acc[w][k][0].x += dr * z[k].x - di * z[k].y;  // (dr + i di) z
              acc[w][k][0].y += dr * z[k].y + di * z[k].x;
              acc[w][k][1].x += dr * z[k].x + di * z[k].y;  // (dr + i di) conj(z)
              acc[w][k][1].y += di * z[k].x - dr * z[k].y;
            }
          } else {
#pragma unroll
            for (int k = 0; k < 3; k++) {
              acc[w][k][0].x += dr * z[k].x;
              acc[w][k][0].y += dr * z[k].y;
            }
          }
        }
      }
    }
    __syncthreads();  // the stage has been read
    if (tn < te) HW_COMMIT(tn);
    __syncthreads();
  }
  if (c >= ncolM) return;
  double2 *ps = part + (long)blockIdx.y * NWB * 3 * NV * pstride;
#pragma unroll
  for (int w = 0; w < NWB; w++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int v = 0; v < NV; v++) ps[(((long)w * 3 + k) * NV + v) * pstride + c] = acc[w][k][v];
#undef HW_CLAMPT
#undef HW_FETCH
#undef HW_COMMIT
}

// out = beta*out + alpha*sum_s part (fixed order): real weights
// wing[k + 3*(c + ncolM*w)], complex wing[k + 3*(c + ncolM*(v + 2*w))].
__global__ void headwing_reduce_kernel(int ncolM, int nv, int nwb, int nwv, int w0, int nsplit,
                                       long ncolp, const double2 *__restrict__ part, double alpha,
                                       double beta, double2 *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;  // (w*3 + k)*nv + v
  if (c >= ncolM) return;
  const int v = r % nv, k = (r / nv) % 3, w = r / (3 * nv);
  if (w >= nwv) return;
  const long stride = (long)nwb * 3 * nv * ncolp;
  double sx = 0.0, sy = 0.0;
  for (int s = 0; s < nsplit; s++) {
    const double2 q = part[s * stride + (long)r * ncolp + c];
    sx += q.x;
    sy += q.y;
  }
  const long wg = w0 + w;
  double2 *o = out + k + 3 * ((long)c + (long)ncolM * (nv == 2 ? v + 2 * wg : wg));
  double2 val = make_double2(alpha * sx, alpha * sy);
  if (beta != 0.0) {
    const double2 old = *o;
    val.x += beta * old.x;
    val.y += beta * old.y;
  }
  *o = val;
}

// The head: sums over t of D(t,w) z_j(t) for the six pairs j = (k1 <= k2),
// z = P(t,k1) conj(P(t,k2)).  One thread per transition (grid-stride),
// NWH frequencies per launch; per block the sums are reduced across each
// wave with shuffles, then across the four waves in LDS, into
// part[block][(w*6 + j)*NV + v].
constexpr int HD_NT = 256, HD_NWH = 4;

template <bool CPLX>
__global__ __launch_bounds__(HD_NT) void head_partial_kernel(
    long n, const double2 *__restrict__ P, long ldp, int nw, int w0,
    const double *__restrict__ Dr, const double *__restrict__ Di, long ldd,
    double2 *__restrict__ part) {
  constexpr int NV = CPLX ? 2 : 1;
  constexpr int NR = HD_NWH * 6 * NV;
  __shared__ double2 red[HD_NT / 64][NR];
  double2 acc[HD_NWH][6][NV];
#pragma unroll
  for (int w = 0; w < HD_NWH; w++)
#pragma unroll
    for (int j = 0; j < 6; j++)
#pragma unroll
      for (int v = 0; v < NV; v++) acc[w][j][v] = make_double2(0.0, 0.0);
  for (long t = (long)blockIdx.x * HD_NT + threadIdx.x; t < n; t += (long)gridDim.x * HD_NT) {
    const double2 p0 = P[t], p1 = P[t + ldp], p2 = P[t + 2 * ldp];
    const double2 pk[3] = {p0, p1, p2};
    double2 z[6];
    int j = 0;
#pragma unroll
    for (int k2 = 0; k2 < 3; k2++)
#pragma unroll
      for (int k1 = 0; k1 <= k2; k1++, j++)
        z[j] = make_double2(pk[k1].x * pk[k2].x + pk[k1].y * pk[k2].y,
                            pk[k1].y * pk[k2].x - pk[k1].x * pk[k2].y);  // p_k1 conj(p_k2)
#pragma unroll
    for (int w = 0; w < HD_NWH; w++) {
      const bool in = w0 + w < nw;
      const long off = t + ldd * (long)(in ? w0 + w : w0);
      const double dr = in ? Dr[off] : 0.0;
      if constexpr (CPLX) {
        const double di = in ? Di[off] : 0.0;
#pragma unroll
        for (int q = 0; q < 6; q++) {
          acc[w][q][0].x += dr * z[q].x - di * z[q].y;  // (dr + i di) z
          acc[w][q][0].y += dr * z[q].y + di * z[q].x;
          acc[w][q][1].x += dr * z[q].x + di * z[q].y;  // (dr + i di) conj(z)
          acc[w][q][1].y += di * z[q].x - dr * z[q].y;
        }
      } else {
#pragma unroll
        for (int q = 0; q < 6; q++) {
          acc[w][q][0].x += dr * z[q].x;
          acc[w][q][0].y += dr * z[q].y;
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int w = 0; w < HD_NWH; w++)
#pragma unroll
    for (int q = 0; q < 6; q++)
#pragma unroll
      for (int v = 0; v < NV; v++) {
        double x = acc[w][q][v].x, y = acc[w][q][v].y;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          x += __shfl_down(x, o, 64);
          y += __shfl_down(y, o, 64);
        }
        if (lane == 0) red[wv][(w * 6 + q) * NV + v] = make_double2(x, y);
      }
  __syncthreads();
  if (threadIdx.x < NR) {
    double2 sum = red[0][threadIdx.x];
#pragma unroll
    for (int r = 1; r < HD_NT / 64; r++) {
      sum.x += red[r][threadIdx.x].x;
      sum.y += red[r][threadIdx.x].y;
    }
    part[(long)blockIdx.x * NR + threadIdx.x] = sum;
  }
}

// head(k1,k2,w0+w) = beta*head + alpha*sum over blocks (fixed order), and
// the partner (k2,k1): the conjugate (real weights) or the conj(z) form.
// One wave per value r = (w*6 + j)*nv + v: lane l sums blocks l, l+64, ...,
// then the lanes are combined by shuffles (a fixed order).
__global__ __launch_bounds__(64) void head_reduce_kernel(int nv, int nwv, int w0, int nblk,
                                                         const double2 *__restrict__ part,
                                                         double alpha, double beta,
                                                         double2 *__restrict__ head) {
  const int r = blockIdx.x;
  const int nr = HD_NWH * 6 * nv;
  const int v = r % nv, j = (r / nv) % 6, w = r / (6 * nv);
  double sx = 0.0, sy = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) {
    const double2 q = part[(long)b * nr + r];
    sx += q.x;
    sy += q.y;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sx += __shfl_down(sx, o, 64);
    sy += __shfl_down(sy, o, 64);
  }
  if (threadIdx.x != 0 || w >= nwv) return;
  static constexpr int K1[6] = {0, 0, 1, 0, 1, 2}, K2[6] = {0, 1, 1, 2, 2, 2};
  const int k1 = K1[j], k2 = K2[j];
  auto put = [&](int a, int b, double x, double y) {
    double2 *o = head + a + 3 * b + 9 * (long)(w0 + w);
    double2 val = make_double2(alpha * x, alpha * y);
    if (beta != 0.0) {
      const double2 old = *o;
      val.x += beta * old.x;
      val.y += beta * old.y;
    }
    *o = val;
  };
  if (nv == 1) {
    put(k1, k2, sx, sy);
    if (k1 != k2) put(k2, k1, sx, -sy);
  } else if (v == 0) {
    put(k1, k2, sx, sy);
  } else if (k1 != k2) {
    put(k2, k1, sx, sy);
  }
}

// Row and column 0 of each chi to zero, wing(:,0,:) to zero, then head =
// 1 - 4 pi head and wing(:,i,:) = -sqrt(4 pi) sqcoul(i) wing(:,i,:).
__global__ void gamma_prepare_kernel(int nb, int nw, int nv, double2 *__restrict__ X, long ldx,
                                     double2 *__restrict__ head, double2 *__restrict__ wing,
                                     const double *__restrict__ sq, double fourpi, double sfp) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y;
  if (g < nb) {
    double2 *Xw = X + ldx * nb * (long)w;
    Xw[g] = make_double2(0.0, 0.0);        // chi(g, 0)
    Xw[ldx * g] = make_double2(0.0, 0.0);  // chi(0, g)
  }
  if (g < 3 * (long)nb) {  // nv wing forms per frequency (2: wingc(:,:,1:2,w))
    const int i = (int)(g / 3);
    const double s = i == 0 ? 0.0 : -sfp * sq[i];
    for (int v = 0; v < nv; v++) {
      double2 *p = wing + g + 3 * (long)nb * (v + nv * (long)w);
      *p = make_double2(s * p->x, s * p->y);
    }
  }
  if (g < 9) {
    double2 *p = head + g + 9 * (long)w;
    const double id = (g % 4 == 0) ? 1.0 : 0.0;  // diagonal of the 3 x 3
    *p = make_double2(id - fourpi * p->x, -fourpi * p->y);
  }
}

// tmp(k, j, w) = -sum_i wing(k, i, w) E(j, i, w) = -(wing E^T)(k, j): for a
// Hermitian E that is -wing conj(E) (correlation.f:756), for the
// real-frequency dielecc the first wing's -wingc(:,:,1) dielecc^T (:761).  64 columns j per workgroup,
// the four waves split i and are summed in LDS.
__global__ __launch_bounds__(256) void gamma_wing_kernel(int nb, const double2 *__restrict__ E,
                                                         long lde, const double2 *__restrict__ wing,
                                                         long wstr, double2 *__restrict__ tmp) {
  __shared__ double2 red[4][3][64];
  const int w = blockIdx.y;
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  const double2 *Ew = E + lde * nb * (long)w;
  const double2 *ww = wing + wstr * w;
  const int per = (nb + 3) / 4;
  const int i0 = q * per, i1 = i0 + per < nb ? i0 + per : nb;
  double2 a[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  if (j < nb) {
    for (int i = i0; i < i1; i++) {
      const double2 e = Ew[j + lde * (long)i];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double2 x = ww[k + 3 * (long)i];
        a[k].x += x.x * e.x - x.y * e.y;
        a[k].y += x.x * e.y + x.y * e.x;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) red[q][k][lane] = a[k];
  __syncthreads();
  if (q == 0 && j < nb) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double2 s = red[0][k][lane];
      for (int r = 1; r < 4; r++) {
        s.x += red[r][k][lane].x;
        s.y += red[r][k][lane].y;
      }
      tmp[k + 3 * ((long)j + (long)nb * w)] = make_double2(-s.x, -s.y);
    }
  }
}

// head(k1,k2,w) += sum_i conj(wing(k1,i,w)) tmp(k2,i,w), then wing = tmp.
// One workgroup per frequency; every thread reads and writes only its own
// columns i.
__global__ __launch_bounds__(256) void gamma_head_kernel(int nb, double2 *__restrict__ head,
                                                         double2 *__restrict__ wing,
                                                         const double2 *__restrict__ tmp) {
  __shared__ double2 red[256][9];
  const int w = blockIdx.x;
  double2 a[9];
#pragma unroll
  for (int e = 0; e < 9; e++) a[e] = make_double2(0.0, 0.0);
  double2 *ww = wing + 3 * (long)nb * w;
  const double2 *tw = tmp + 3 * (long)nb * w;
  for (int i = threadIdx.x; i < nb; i += 256) {
    double2 o[3], nw3[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      o[k] = ww[k + 3 * (long)i];
      nw3[k] = tw[k + 3 * (long)i];
    }
#pragma unroll
    for (int k2 = 0; k2 < 3; k2++)
#pragma unroll
      for (int k1 = 0; k1 < 3; k1++) {
        // conj(o[k1]) * nw3[k2]
        a[k1 + 3 * k2].x += o[k1].x * nw3[k2].x + o[k1].y * nw3[k2].y;
        a[k1 + 3 * k2].y += o[k1].x * nw3[k2].y - o[k1].y * nw3[k2].x;
      }
#pragma unroll
    for (int k = 0; k < 3; k++) ww[k + 3 * (long)i] = nw3[k];
  }
#pragma unroll
  for (int e = 0; e < 9; e++) red[threadIdx.x][e] = a[e];
  __syncthreads();
  if (threadIdx.x < 9) {
    double sx = 0.0, sy = 0.0;
    for (int r = 0; r < 256; r++) {
      sx += red[r][threadIdx.x].x;
      sy += red[r][threadIdx.x].y;
    }
    double2 *h = head + threadIdx.x + 9 * (long)w;
    *h = make_double2(h->x + sx, h->y + sy);
  }
}

template <bool CPLX, int NWB, bool CM>
void launch_hw(dim3 grid, hipStream_t st, long n, int ncolM, const void *M, long ldm, const void *P,
               long ldp, int nw, int w0, const void *Dr, const void *Di, long ldd, long kchunk,
               void *part) {
  hipLaunchKernelGGL((headwing_kernel<CPLX, NWB, CM>), grid, dim3(HW_NT), 0, st, n, ncolM,
                     (const double2 *)M, ldm, (const double2 *)P, ldp, nw, w0, (const double *)Dr,
                     (const double *)Di, ldd, kchunk, (long)ncolM, (double2 *)part);
}

template <bool CPLX, int NWB>
void launch_hw_layout(bool cm, dim3 grid, hipStream_t st, long n, int ncolM, const void *M, long ldm,
                      const void *P, long ldp, int nw, int w0, const void *Dr, const void *Di,
                      long ldd, long kchunk, void *part) {
  if (cm)
    launch_hw<CPLX, NWB, true>(grid, st, n, ncolM, M, ldm, P, ldp, nw, w0, Dr, Di, ldd, kchunk, part);
  else
    launch_hw<CPLX, NWB, false>(grid, st, n, ncolM, M, ldm, P, ldp, nw, w0, Dr, Di, ldd, kchunk, part);
}

// frequencies per pass: accumulators (3 x NWB x NV complex) plus the
// in-flight stage (16 complex) within 256 VGPRs
constexpr int HW_MAXW_REAL = 12, HW_MAXW_CPLX = 6;

int nwb_for(bool cplx, int rem) {
  if (cplx) return rem <= 2 ? 2 : rem <= 4 ? 4 : 6;
  return rem <= 4 ? 4 : rem <= 8 ? 8 : 12;
}

// The wing sums over the columns of M against P, all frequency passes,
// into out.
int hw_sums(spexhip_ctx *ctx, long n, int ncolM, const void *M, long ldm, bool cm, const void *P,
            long ldp, int nw, const void *Dr, const void *Di, long ldd, double alpha, double beta,
            void *out) {
  const bool cplx = Di != nullptr;
  const int nv = cplx ? 2 : 1;
  const int ncb = (ncolM + HW_NT - 1) / HW_NT;
  const long ncolp = ncolM;  // row stride of the partial sums
  // two workgroups fit a CU (LDS): one round of 2 x 256 workgroups, split
  // evenly over the column blocks, each split at least 4 stages long
  long nsplit = 512 / ncb;
  const long maxsplit = (n + 4 * HW_TT - 1) / (4 * HW_TT);
  if (nsplit > maxsplit) nsplit = maxsplit;
  if (nsplit < 1) nsplit = 1;
  long kchunk = (n + nsplit - 1) / nsplit;
  kchunk = (kchunk + HW_TT - 1) / HW_TT * HW_TT;
  if (kchunk < HW_TT) kchunk = HW_TT;
  nsplit = n > 0 ? (n + kchunk - 1) / kchunk : 1;
  const int nwbmax = cplx ? HW_MAXW_CPLX : HW_MAXW_REAL;
  void *part;
  int rc = spx_scratch(ctx, 16 * (size_t)nsplit * nwbmax * 3 * nv * ncolp, &part);
  if (rc) return rc;
  for (int w0 = 0; w0 < nw; w0 += nwbmax) {
    const int rem = nw - w0 < nwbmax ? nw - w0 : nwbmax;
    const int nwb = nwb_for(cplx, rem);
    if (n > 0) {
      const dim3 grid((unsigned)ncb, (unsigned)nsplit);
#define HW_GO(C, B) \
  launch_hw_layout<C, B>(cm, grid, ctx->stream, n, ncolM, M, ldm, P, ldp, nw, w0, Dr, Di, ldd, kchunk, part)
      if (cplx) {
        switch (nwb) {
          case 2: HW_GO(true, 2); break;
          case 4: HW_GO(true, 4); break;
          default: HW_GO(true, 6); break;
        }
      } else {
        switch (nwb) {
          case 4: HW_GO(false, 4); break;
          case 8: HW_GO(false, 8); break;
          default: HW_GO(false, 12); break;
        }
      }
#undef HW_GO
      SPX_LAUNCHED(ctx);
    } else {
      rc = spx_fill_zero(ctx, ctx->stream, part, 16 * (size_t)nwbmax * 3 * nv * ncolp);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(headwing_reduce_kernel,
                       dim3((unsigned)((ncolM + 255) / 256), (unsigned)(nwb * 3 * nv)), dim3(256),
                       0, ctx->stream, ncolM, nv, nwb, rem, w0,
                       (int)(n > 0 ? nsplit : 1), ncolp, (const double2 *)part, alpha, beta,
                       (double2 *)out);
    SPX_LAUNCHED(ctx);
  }
  return SPEXHIP_OK;
}

// Shared by the real- and complex-weight entry points: the wings as the
// sums over the columns of M, the head from P alone.
int headwings(spexhip_ctx *ctx, long n, int nb, const void *M, long ldm, int colmajor,
              const void *P, long ldp, int nw, const void *Dr, const void *Di, const void *Dhr,
              const void *Dhi, long ldd, double alpha, double beta, void *head, void *wing) {
  const bool cplx = Di != nullptr;
  const int nv = cplx ? 2 : 1;
  SPX_CHECK(ctx, n >= 0 && nb > 0 && nw > 0 && ldp >= n && ldd >= n, "headwings: bad sizes");
  SPX_CHECK(ctx, colmajor ? ldm >= n : ldm >= nb, "headwings: ldm too small");
  if (!Dhr) Dhr = Dr;
  if (cplx && !Dhi) Dhi = Di;
  SPX_CHECK(ctx, spx_owns(ctx, head, 16 * 9 * (size_t)nw) &&
                     spx_owns(ctx, wing, 16 * 3 * (size_t)nb * nv * nw),
            "headwings: head (3x3xnw) or wing (3 x nb x %s nw) not context allocations", cplx ? "2 x" : "");
  if (n > 0) {
    const size_t mbytes = colmajor ? 16 * ((size_t)ldm * (nb - 1) + n)
                                   : 16 * ((size_t)ldm * (n - 1) + nb);
    const size_t dbytes = 8 * ((size_t)ldd * (nw - 1) + n);
    SPX_CHECK(ctx, spx_owns(ctx, M, mbytes), "headwings: M is not a context allocation of the stated shape");
    SPX_CHECK(ctx, spx_owns(ctx, P, 16 * ((size_t)ldp * 2 + n)),
              "headwings: P is not a context allocation of n x 3 complex (ld ldp)");
    SPX_CHECK(ctx, spx_owns(ctx, Dr, dbytes) && spx_owns(ctx, Dhr, dbytes) &&
                       (!cplx || (spx_owns(ctx, Di, dbytes) && spx_owns(ctx, Dhi, dbytes))),
              "headwings: weights are not context allocations of n x nw (ld ldd)");
  }
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  int rc = hw_sums(ctx, n, nb, M, ldm, colmajor != 0, P, ldp, nw, Dr, Di, ldd, alpha, beta,
                   wing);
  if (rc) return rc;
  // the head: its own two-stage reduction (a second scratch region, after
  // the wing sums on the same stream)
  const int nblk = (int)std::min<long>(1024, std::max<long>(1, (n + HD_NT - 1) / HD_NT));
  const int nr = HD_NWH * 6 * nv;
  void *hpart;
  rc = spx_scratch(ctx, 16 * (size_t)nblk * nr, &hpart);
  if (rc) return rc;
  for (int w0 = 0; w0 < nw; w0 += HD_NWH) {
    if (n > 0) {
      if (cplx)
        hipLaunchKernelGGL(head_partial_kernel<true>, dim3(nblk), dim3(HD_NT), 0, ctx->stream, n,
                           (const double2 *)P, ldp, nw, w0, (const double *)Dhr,
                           (const double *)Dhi, ldd, (double2 *)hpart);
      else
        hipLaunchKernelGGL(head_partial_kernel<false>, dim3(nblk), dim3(HD_NT), 0, ctx->stream, n,
                           (const double2 *)P, ldp, nw, w0, (const double *)Dhr, nullptr, ldd,
                           (double2 *)hpart);
      SPX_LAUNCHED(ctx);
    } else {
      rc = spx_fill_zero(ctx, ctx->stream, hpart, 16 * (size_t)nr);
      if (rc) return rc;
    }
    const int nwv = nw - w0 < HD_NWH ? nw - w0 : HD_NWH;
    hipLaunchKernelGGL(head_reduce_kernel, dim3(nr), dim3(64), 0, ctx->stream, nv, nwv, w0,
                       n > 0 ? nblk : 1, (const double2 *)hpart, alpha, beta, (double2 *)head);
    SPX_LAUNCHED(ctx);
  }
  return SPEXHIP_OK;
}

// tmp(k, j, w) = -sum_i wing(k, i, w) E(i, j, w): the second real-frequency
// wing, -wingc(:,:,2) dielecc (correlation.f:763).  One wave per column j,
// the lanes split i, shuffles combine them (fixed order).
__global__ __launch_bounds__(64) void gamma_wing_t_kernel(int nb, const double2 *__restrict__ E,
                                                          long lde, const double2 *__restrict__ wing,
                                                          long wstr, double2 *__restrict__ tmp) {
  const int j = blockIdx.x, w = blockIdx.y, lane = threadIdx.x;
  const double2 *Ej = E + lde * nb * (long)w + lde * (long)j;
  const double2 *ww = wing + wstr * w;
  double a[3][2] = {};
  for (int i = lane; i < nb; i += 64) {
    const double2 e = Ej[i];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double2 x = ww[k + 3 * (long)i];
      a[k][0] += x.x * e.x - x.y * e.y;
      a[k][1] += x.x * e.y + x.y * e.x;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) a[k][c] += __shfl_down(a[k][c], o, 64);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++)
      tmp[k + 3 * ((long)j + (long)nb * w)] = make_double2(-a[k][0], -a[k][1]);
  }
}

// Real frequencies (correlation.f:761-763): headc(k1,k2,w) += sum_i
// wingc(k1,i,2,w) tmp1(k2,i,w) (the old second wing, the new first one, no
// conjugation), then wingc(:,:,1) = tmp1 and wingc(:,:,2) = tmp2.  One
// workgroup per frequency; each thread reads and writes only its own i.
__global__ __launch_bounds__(256) void gamma_head_cplx_kernel(int nb, double2 *__restrict__ head,
                                                              double2 *__restrict__ wing,
                                                              const double2 *__restrict__ t1,
                                                              const double2 *__restrict__ t2) {
  __shared__ double2 red[256][9];
  const int w = blockIdx.x;
  double2 a[9];
#pragma unroll
  for (int e = 0; e < 9; e++) a[e] = make_double2(0.0, 0.0);
  double2 *w1 = wing + 3 * (long)nb * (2 * (long)w), *w2 = w1 + 3 * (long)nb;
  const double2 *n1 = t1 + 3 * (long)nb * w, *n2 = t2 + 3 * (long)nb * w;
  for (int i = threadIdx.x; i < nb; i += 256) {
    double2 o[3], u[3], v[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      o[k] = w2[k + 3 * (long)i];
      u[k] = n1[k + 3 * (long)i];
      v[k] = n2[k + 3 * (long)i];
    }
#pragma unroll
    for (int k2 = 0; k2 < 3; k2++)
#pragma unroll
      for (int k1 = 0; k1 < 3; k1++) {
        a[k1 + 3 * k2].x += o[k1].x * u[k2].x - o[k1].y * u[k2].y;
        a[k1 + 3 * k2].y += o[k1].x * u[k2].y + o[k1].y * u[k2].x;
      }
#pragma unroll
    for (int k = 0; k < 3; k++) {
      w1[k + 3 * (long)i] = u[k];
      w2[k + 3 * (long)i] = v[k];
    }
  }
#pragma unroll
  for (int e = 0; e < 9; e++) red[threadIdx.x][e] = a[e];
  __syncthreads();
  if (threadIdx.x < 9) {
    double sx = 0.0, sy = 0.0;
    for (int r = 0; r < 256; r++) {
      sx += red[r][threadIdx.x].x;
      sy += red[r][threadIdx.x].y;
    }
    double2 *h = head + threadIdx.x + 9 * (long)w;
    *h = make_double2(h->x + sx, h->y + sy);
  }
}

// E(0,0,w) = trace(head(:,:,w))/3, or vec^T head vec with a unit vector.
__global__ void gamma_head_scalar_kernel(int nw, double2 *__restrict__ E, long lde, long nb,
                                         const double2 *__restrict__ head, int use_vec, double v0,
                                         double v1, double v2) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const double2 *h = head + 9 * (long)w;
  double2 val;
  if (use_vec) {
    const double v[3] = {v0, v1, v2};
    double x = 0.0, y = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        x += v[a] * h[a + 3 * b].x * v[b];
        y += v[a] * h[a + 3 * b].y * v[b];
      }
    val = make_double2(x, y);
  } else {
    val = make_double2((h[0].x + h[4].x + h[8].x) / 3, (h[0].y + h[4].y + h[8].y) / 3);
  }
  E[lde * nb * (long)w] = val;
}

// MULTDIFF ON at any q (susceptibility.f:912-925, :1193): the first column
// of M scaled by the band-energy difference, next to the unscaled column,
// as the two momentum-like columns of the head/wing sums.
// P(t,0) = M(t,0) (e_unocc(k,b2) - e_occ(k,b1)), P(t,1) = M(t,0), P(t,2) = 0.
__global__ void multdiff_cols_kernel(int nocc, int nunocc, long n, const double2 *__restrict__ M,
                                     long ldm, int colmajor, const double *__restrict__ eo,
                                     const double *__restrict__ eu, double2 *__restrict__ P,
                                     long ldp) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long per = (long)nocc * nunocc;
  const long k = t / per;
  const int b1 = (int)((t - k * per) / nunocc), b2 = (int)(t % nunocc);
  const double2 m = colmajor ? M[t] : M[t * ldm];
  const double d = eu[k * nunocc + b2] - eo[k * nocc + b1];
  P[t] = make_double2(m.x * d, m.y * d);
  P[t + ldp] = m;
  P[t + 2 * ldp] = make_double2(0.0, 0.0);
}

// Row and column 0 of chi += alpha * (MULTDIFF sums - plain sums).  Wings
// of the two passes: wA(0,J) = sum D1 M(t,J) conj(P0), wB(1,J) = sum D0
// M(t,J) conj(M0); heads hA(0,0) = sum D2 |P0|^2, hB(1,1) = sum D0 |M0|^2.
// cplx: wing(k,i,s,w) with s = 0 for (i,0), s = 1 for (0,i); else the
// Hermitian (0,J) = conj of (J,0).
// inton: MULTDIFF INTON (multdiff_ = 2) instead, from one pass whose head
// weights are Dmd1 + Dmd2 and wing weights D0 (hB, wB unused): row and column
// 0 lose their plain sums, (0,0) becomes the sum with Dmd1 + Dmd2.
__global__ void multdiff_fix_kernel(int nb, int nw, int cplx, int inton,
                                    const double2 *__restrict__ hA,
                                    const double2 *__restrict__ wA,
                                    const double2 *__restrict__ hB,
                                    const double2 *__restrict__ wB, double alpha,
                                    double2 *__restrict__ chi, long ldchi) {
  const int J = blockIdx.x * blockDim.x + threadIdx.x, w = blockIdx.y;
  if (J >= nb || w >= nw) return;
  double2 *X = chi + ldchi * (long)nb * w;
  const int ns = cplx ? 2 : 1;
  const long wa = 3 * ((long)J + (long)nb * ns * w), wb = 1 + wa;  // k = 0 and k = 1
  if (J == 0) {
    // (0,0) and the plain (1,1) sum (inton: the k = 1 wing of column 0)
    const double2 a = hA[9 * w], b = inton ? wA[wb] : hB[4 + 9 * w];
    X[0].x += alpha * (a.x - b.x);
    X[0].y += alpha * (a.y - b.y);
    return;
  }
  const double2 lo = inton ? make_double2(-wA[wb].x, -wA[wb].y)
                           : make_double2(wA[wa].x - wB[wb].x, wA[wa].y - wB[wb].y);  // (J,0)
  double2 up;
  if (cplx) {
    const long s1 = 3L * nb;
    up = inton ? make_double2(-wA[wb + s1].x, -wA[wb + s1].y)
               : make_double2(wA[wa + s1].x - wB[wb + s1].x, wA[wa + s1].y - wB[wb + s1].y);
  } else {
    up = make_double2(lo.x, -lo.y);
  }
  double2 &xl = X[J], &xu = X[ldchi * (long)J];
  xl.x += alpha * lo.x;
  xl.y += alpha * lo.y;
  xu.x += alpha * up.x;
  xu.y += alpha * up.y;
}

__global__ void weight_sum_kernel(long n, int nw, long ldd, const double *__restrict__ A,
                                  const double *__restrict__ B, double *__restrict__ S) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long o = t + ldd * blockIdx.y;
  S[o] = A[o] + B[o];
}

// The MULTDIFF update of chi's row and column 0 after the plain pair sum.
// ON (inton = false): D1/D2 are the Hilbert folds with wghtS / frqs^l.
// INTON (inton = true): D1/D2 are the folds with the md = 1 and md = 2
// Hilbert weights (getwghtS(l), hilbert.f:106-300).
int multdiff(spexhip_ctx *ctx, bool inton, int nk, int nocc, int nunocc, const void *eo,
             const void *eu, int nb, const void *M, long ldm, int colmajor, int nw,
             const void *D0r, const void *D0i, const void *D1r, const void *D1i, const void *D2r,
             const void *D2i, long ldd, double alpha, void *chi, int ldchi) {
  const bool cplx = D0i != nullptr;
  SPX_CHECK(ctx, nk > 0 && nocc > 0 && nunocc > 0 && nb > 0 && nw > 0 && ldchi >= nb,
            "multdiff: bad sizes");
  SPX_CHECK(ctx, D0r && D1r && D2r && (cplx == (D1i != nullptr)) && (cplx == (D2i != nullptr)),
            "multdiff: missing weights");
  const long n = (long)nk * nocc * nunocc;
  SPX_CHECK(ctx, ldd >= n, "multdiff: ldd < n");
  SPX_CHECK(ctx, spx_owns(ctx, eo, 8 * (size_t)nk * nocc) &&
                     spx_owns(ctx, eu, 8 * (size_t)nk * nunocc) &&
                     spx_owns(ctx, chi, 16 * (size_t)ldchi * nb * nw),
            "multdiff: energies or chi are not context allocations of the stated shapes");
  const size_t dbytes = 8 * ((size_t)ldd * (nw - 1) + n);
  SPX_CHECK(ctx, spx_owns(ctx, D1r, dbytes) && spx_owns(ctx, D2r, dbytes) &&
                     (!cplx || (spx_owns(ctx, D1i, dbytes) && spx_owns(ctx, D2i, dbytes))),
            "multdiff: weights are not context allocations of n x nw (ld ldd)");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const int ns = cplx ? 2 : 1;
  void *P = nullptr, *hA = nullptr, *hB = nullptr, *wA = nullptr, *wB = nullptr, *Sr = nullptr,
       *Si = nullptr;
  int rc = spx_alloc(ctx, 16 * 3 * (size_t)n, &P);
  if (!rc) rc = spx_alloc(ctx, 16 * 9 * (size_t)nw, &hA);
  if (!rc) rc = spx_alloc(ctx, 16 * 3 * (size_t)nb * ns * nw, &wA);
  if (!rc && !inton) rc = spx_alloc(ctx, 16 * 9 * (size_t)nw, &hB);
  if (!rc && !inton) rc = spx_alloc(ctx, 16 * 3 * (size_t)nb * ns * nw, &wB);
  if (!rc && inton) rc = spx_alloc(ctx, 8 * (size_t)ldd * nw, &Sr);
  if (!rc && inton && cplx) rc = spx_alloc(ctx, 8 * (size_t)ldd * nw, &Si);
  if (!rc) {
    hipLaunchKernelGGL(multdiff_cols_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, nocc, nunocc, n, (const double2 *)M, ldm, colmajor,
                       (const double *)eo, (const double *)eu, (double2 *)P, n);
    if (inton) {
      const dim3 g((unsigned)((n + 255) / 256), (unsigned)nw);
      hipLaunchKernelGGL(weight_sum_kernel, g, dim3(256), 0, ctx->stream, n, nw, ldd,
                         (const double *)D1r, (const double *)D2r, (double *)Sr);
      if (cplx)
        hipLaunchKernelGGL(weight_sum_kernel, g, dim3(256), 0, ctx->stream, n, nw, ldd,
                           (const double *)D1i, (const double *)D2i, (double *)Si);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "multdiff: %s", hipGetErrorString(e));
  }
  if (inton) {
    // wings with D0 (the plain sums of row 0, to be removed), head with
    // Dmd1 + Dmd2 (element (0,0) of passes ll = 1 and ll = 2)
    if (!rc)
      rc = headwings(ctx, n, nb, M, ldm, colmajor, P, n, nw, D0r, D0i, Sr, Si, ldd, 1.0, 0.0, hA,
                     wA);
  } else {
    // wings with D1 (1/frqs), head with D2 (1/frqs^2); plain sums of the
    // same row with D0
    if (!rc)
      rc = headwings(ctx, n, nb, M, ldm, colmajor, P, n, nw, D1r, D1i, D2r, D2i, ldd, 1.0, 0.0, hA,
                     wA);
    if (!rc)
      rc = headwings(ctx, n, nb, M, ldm, colmajor, P, n, nw, D0r, D0i, D0r, D0i, ldd, 1.0, 0.0, hB,
                     wB);
  }
  if (!rc) {
    hipLaunchKernelGGL(multdiff_fix_kernel, dim3((unsigned)((nb + 63) / 64), (unsigned)nw),
                       dim3(64), 0, ctx->stream, nb, nw, (int)cplx, (int)inton,
                       (const double2 *)hA, (const double2 *)wA,
                       (const double2 *)(inton ? hA : hB), (const double2 *)(inton ? wA : wB),
                       alpha, (double2 *)chi, (long)ldchi);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "multdiff: %s", hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(ctx->stream);
  for (void *p : {P, hA, hB, wA, wB, Sr, Si})
    if (p) spx_free(ctx, p);
  return rc;
}

}  // namespace

extern "C" {

int spexhip_gamma_head_scalar_dev(spexhip_ctx *ctx, int nb, int nw, void *E, int lde,
                                  const void *head, const double *vec) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && lde >= nb, "gamma_head_scalar: bad sizes");
  SPX_CHECK(ctx, spx_owns(ctx, E, 16 * (size_t)lde * nb * nw) &&
                     spx_owns(ctx, head, 16 * 9 * (size_t)nw),
            "gamma_head_scalar: arrays are not context allocations of the stated shapes");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(gamma_head_scalar_kernel, dim3((nw + 63) / 64), dim3(64), 0, ctx->stream, nw,
                     (double2 *)E, (long)lde, (long)nb, (const double2 *)head, vec ? 1 : 0,
                     vec ? vec[0] : 0.0, vec ? vec[1] : 0.0, vec ? vec[2] : 0.0);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_momentum_scale_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc, const void *e_occ,
                               const void *e_unocc, double vol, void *P, long ldp) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nk > 0 && nocc > 0 && nunocc > 0 && vol > 0, "momentum_scale: bad sizes");
  const long n = (long)nk * nocc * nunocc;
  SPX_CHECK(ctx, ldp >= n, "momentum_scale: ldp < n");
  SPX_CHECK(ctx, (e_occ == nullptr) == (e_unocc == nullptr),
            "momentum_scale: give both energy arrays or neither");
  SPX_CHECK(ctx, (!e_occ || (spx_owns(ctx, e_occ, 8 * (size_t)nk * nocc) &&
                             spx_owns(ctx, e_unocc, 8 * (size_t)nk * nunocc))) &&
                     spx_owns(ctx, P, 16 * ((size_t)ldp * 2 + n)),
            "momentum_scale: arrays are not context allocations of the stated shapes");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(momentum_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     ctx->stream, nk, nocc, nunocc, (const double *)e_occ,
                     (const double *)e_unocc, std::sqrt(vol), (double2 *)P, ldp);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_headwings_dev(spexhip_ctx *ctx, long n, int nb, const void *M, long ldm, int colmajor,
                          const void *P, long ldp, int nw, const void *D, const void *Dh, long ldd,
                          double alpha, double beta, void *head, void *wing) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, D != nullptr, "headwings: null weights");
  return headwings(ctx, n, nb, M, ldm, colmajor, P, ldp, nw, D, nullptr, Dh, nullptr, ldd, alpha,
                   beta, head, wing);
}

int spexhip_headwings_cplx_dev(spexhip_ctx *ctx, long n, int nb, const void *M, long ldm,
                               int colmajor, const void *P, long ldp, int nw, const void *Dr,
                               const void *Di, const void *Dhr, const void *Dhi, long ldd,
                               double alpha, double beta, void *head, void *wing) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, Dr != nullptr && Di != nullptr, "headwings_cplx: null weights");
  SPX_CHECK(ctx, (Dhr == nullptr) == (Dhi == nullptr),
            "headwings_cplx: give both head weights or neither");
  return headwings(ctx, n, nb, M, ldm, colmajor, P, ldp, nw, Dr, Di, Dhr, Dhi, ldd, alpha, beta,
                   head, wing);
}

int spexhip_multdiff_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc, const void *e_occ_dev,
                         const void *e_unocc_dev, int nb, const void *M_dev, long ldm, int colmajor,
                         int nw, const void *D0_dev, const void *D1_dev, const void *D2_dev,
                         long ldd, double alpha, void *chi_dev, int ldchi) {
  if (!ctx) return SPEXHIP_E_INVALID;
  return multdiff(ctx, false, nk, nocc, nunocc, e_occ_dev, e_unocc_dev, nb, M_dev, ldm, colmajor,
                  nw, D0_dev, nullptr, D1_dev, nullptr, D2_dev, nullptr, ldd, alpha, chi_dev,
                  ldchi);
}

int spexhip_multdiff_inton_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc,
                               const void *e_occ_dev, const void *e_unocc_dev, int nb,
                               const void *M_dev, long ldm, int colmajor, int nw,
                               const void *D0_dev, const void *Dmd1_dev, const void *Dmd2_dev,
                               long ldd, double alpha, void *chi_dev, int ldchi) {
  if (!ctx) return SPEXHIP_E_INVALID;
  return multdiff(ctx, true, nk, nocc, nunocc, e_occ_dev, e_unocc_dev, nb, M_dev, ldm, colmajor,
                  nw, D0_dev, nullptr, Dmd1_dev, nullptr, Dmd2_dev, nullptr, ldd, alpha, chi_dev,
                  ldchi);
}

int spexhip_multdiff_cplx_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc,
                              const void *e_occ_dev, const void *e_unocc_dev, int nb,
                              const void *M_dev, long ldm, int colmajor, int nw,
                              const void *D0r_dev, const void *D0i_dev, const void *D1r_dev,
                              const void *D1i_dev, const void *D2r_dev, const void *D2i_dev,
                              long ldd, double alpha, void *chi_dev, int ldchi) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, D0i_dev && D1i_dev && D2i_dev, "multdiff_cplx: null imaginary weights");
  return multdiff(ctx, false, nk, nocc, nunocc, e_occ_dev, e_unocc_dev, nb, M_dev, ldm, colmajor,
                  nw, D0r_dev, D0i_dev, D1r_dev, D1i_dev, D2r_dev, D2i_dev, ldd, alpha, chi_dev,
                  ldchi);
}

int spexhip_multdiff_inton_cplx_dev(spexhip_ctx *ctx, int nk, int nocc, int nunocc,
                                    const void *e_occ_dev, const void *e_unocc_dev, int nb,
                                    const void *M_dev, long ldm, int colmajor, int nw,
                                    const void *D0r_dev, const void *D0i_dev,
                                    const void *Dmd1r_dev, const void *Dmd1i_dev,
                                    const void *Dmd2r_dev, const void *Dmd2i_dev, long ldd,
                                    double alpha, void *chi_dev, int ldchi) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, D0i_dev && Dmd1i_dev && Dmd2i_dev, "multdiff_inton_cplx: null imaginary weights");
  return multdiff(ctx, true, nk, nocc, nunocc, e_occ_dev, e_unocc_dev, nb, M_dev, ldm, colmajor,
                  nw, D0r_dev, D0i_dev, Dmd1r_dev, Dmd1i_dev, Dmd2r_dev, Dmd2i_dev, ldd, alpha,
                  chi_dev, ldchi);
}

int spexhip_gamma_prepare_dev(spexhip_ctx *ctx, int nb, int nw, void *X, int ldx, void *head,
                              void *wing, const void *sqcoul) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && ldx >= nb, "gamma_prepare: bad sizes");
  SPX_CHECK(ctx, spx_owns(ctx, X, 16 * (size_t)ldx * nb * nw) &&
                     spx_owns(ctx, head, 16 * 9 * (size_t)nw) &&
                     spx_owns(ctx, wing, 16 * 3 * (size_t)nb * nw) &&
                     spx_owns(ctx, sqcoul, 8 * (size_t)nb),
            "gamma_prepare: arrays are not context allocations of the stated shapes");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const double fourpi = 4 * M_PI;
  hipLaunchKernelGGL(gamma_prepare_kernel, dim3((unsigned)((3 * (long)nb + 255) / 256), nw),
                     dim3(256), 0, ctx->stream, nb, nw, 1, (double2 *)X, (long)ldx, (double2 *)head,
                     (double2 *)wing, (const double *)sqcoul, fourpi, std::sqrt(fourpi));
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_gamma_inverse_dev(spexhip_ctx *ctx, int nb, int nw, const void *Einv, int lde,
                              void *head, void *wing) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && lde >= nb, "gamma_inverse: bad sizes");
  SPX_CHECK(ctx, spx_owns(ctx, Einv, 16 * (size_t)lde * nb * nw) &&
                     spx_owns(ctx, head, 16 * 9 * (size_t)nw) &&
                     spx_owns(ctx, wing, 16 * 3 * (size_t)nb * nw),
            "gamma_inverse: arrays are not context allocations of the stated shapes");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *tmp;
  int rc = spx_scratch(ctx, 16 * 3 * (size_t)nb * nw, &tmp);
  if (rc) return rc;
  hipLaunchKernelGGL(gamma_wing_kernel, dim3((unsigned)((nb + 63) / 64), nw), dim3(256), 0,
                     ctx->stream, nb, (const double2 *)Einv, (long)lde, (const double2 *)wing,
                     3 * (long)nb, (double2 *)tmp);
  SPX_LAUNCHED(ctx);
  hipLaunchKernelGGL(gamma_head_kernel, dim3(nw), dim3(256), 0, ctx->stream, nb, (double2 *)head,
                     (double2 *)wing, (const double2 *)tmp);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_gamma_prepare_cplx_dev(spexhip_ctx *ctx, int nb, int nw, void *X, int ldx, void *head,
                                   void *wing, const void *sqcoul) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && ldx >= nb, "gamma_prepare_cplx: bad sizes");
  SPX_CHECK(ctx, spx_owns(ctx, X, 16 * (size_t)ldx * nb * nw) &&
                     spx_owns(ctx, head, 16 * 9 * (size_t)nw) &&
                     spx_owns(ctx, wing, 16 * 6 * (size_t)nb * nw) &&
                     spx_owns(ctx, sqcoul, 8 * (size_t)nb),
            "gamma_prepare_cplx: arrays are not context allocations of the stated shapes");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const double fourpi = 4 * M_PI;
  hipLaunchKernelGGL(gamma_prepare_kernel, dim3((unsigned)((3 * (long)nb + 255) / 256), nw),
                     dim3(256), 0, ctx->stream, nb, nw, 2, (double2 *)X, (long)ldx, (double2 *)head,
                     (double2 *)wing, (const double *)sqcoul, fourpi, std::sqrt(fourpi));
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

int spexhip_gamma_inverse_cplx_dev(spexhip_ctx *ctx, int nb, int nw, const void *Einv, int lde,
                                   void *head, void *wing) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && lde >= nb, "gamma_inverse_cplx: bad sizes");
  SPX_CHECK(ctx, spx_owns(ctx, Einv, 16 * (size_t)lde * nb * nw) &&
                     spx_owns(ctx, head, 16 * 9 * (size_t)nw) &&
                     spx_owns(ctx, wing, 16 * 6 * (size_t)nb * nw),
            "gamma_inverse_cplx: arrays are not context allocations of the stated shapes");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  void *tmp;
  int rc = spx_scratch(ctx, 16 * 6 * (size_t)nb * nw, &tmp);
  if (rc) return rc;
  double2 *t1 = (double2 *)tmp, *t2 = t1 + 3 * (size_t)nb * nw;
  // t1 = -wingc(:,:,1) dielecc^T ; t2 = -wingc(:,:,2) dielecc
  hipLaunchKernelGGL(gamma_wing_kernel, dim3((unsigned)((nb + 63) / 64), nw), dim3(256), 0,
                     ctx->stream, nb, (const double2 *)Einv, (long)lde, (const double2 *)wing,
                     6 * (long)nb, t1);
  SPX_LAUNCHED(ctx);
  hipLaunchKernelGGL(gamma_wing_t_kernel, dim3((unsigned)nb, nw), dim3(64), 0, ctx->stream, nb,
                     (const double2 *)Einv, (long)lde, (const double2 *)wing + 3 * (long)nb,
                     6 * (long)nb, t2);
  SPX_LAUNCHED(ctx);
  hipLaunchKernelGGL(gamma_head_cplx_kernel, dim3(nw), dim3(256), 0, ctx->stream, nb,
                     (double2 *)head, (double2 *)wing, (const double2 *)t1, (const double2 *)t2);
  SPX_LAUNCHED(ctx);
  return SPEXHIP_OK;
}

}  // extern "C"

extern "C" {

// Host drop-in for the head and wing sums of one memory packet
// (susceptibility.f:1092-1157): host arrays in the reference layouts, the
// outputs ACCUMULATED across packets like the reference's headh/wingh and
// headc/wingc.  No spin factor (applied once by the caller, :1336).
int spexhip_susceptibility_headwings(spexhip_ctx *ctx, long n, int nb, const double *cprod,
                                     long ldc, const double *dcprod, long lddc, int nfrqh,
                                     const double *wh, const double *whh, int nfrqc,
                                     const double *wcr, const double *wci, const double *wchr,
                                     const double *wchi, double *headh, double *wingh,
                                     double *headc, double *wingc) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, n >= 0 && nb > 0 && nfrqh >= 0 && nfrqc >= 0 && ldc >= n && lddc >= n,
            "susceptibility_headwings: bad arguments");
  SPX_CHECK(ctx, (nfrqh == 0 || (wh && headh && wingh)) &&
                     (nfrqc == 0 || (wcr && wci && headc && wingc)),
            "susceptibility_headwings: missing weights or outputs");
  SPX_CHECK(ctx, (wchr == nullptr) == (wchi == nullptr),
            "susceptibility_headwings: give both real-frequency head weights or neither");
  if (n == 0 || nfrqh + nfrqc == 0) return SPEXHIP_OK;
  SPX_CHECK(ctx, cprod && dcprod, "susceptibility_headwings: null input");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const int nf = nfrqh > nfrqc ? nfrqh : nfrqc;
  const size_t mb = 16 * (size_t)n * nb, pb = 16 * 3 * (size_t)n, db = 8 * (size_t)n * nf,
               hb = 16 * 9 * (size_t)nf, wb = 16 * 3 * (size_t)nb * 2 * nf;
  void *dM = nullptr, *dP = nullptr, *dW[4] = {}, *dH = nullptr, *dWg = nullptr;
  int rc = spx_alloc(ctx, mb, &dM);
  if (!rc) rc = spx_alloc(ctx, pb, &dP);
  for (int i = 0; i < 4 && !rc; i++) rc = spx_alloc(ctx, db, &dW[i]);
  if (!rc) rc = spx_alloc(ctx, hb, &dH);
  if (!rc) rc = spx_alloc(ctx, wb, &dWg);
  auto up = [&](void *dst, const double *src, long rows, long ld, long cols, size_t esz) {
    if (rc) return;
    hipError_t e = hipMemcpy2DAsync(dst, esz * rows, src, esz * ld, esz * rows, cols,
                                    hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "susceptibility_headwings: upload failed: %s",
                      hipGetErrorString(e));
  };
  auto down = [&](void *dst, const void *src, size_t bytes) {
    if (rc) return;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
      rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "susceptibility_headwings: download failed: %s",
                      hipGetErrorString(e));
  };
  up(dM, cprod, n, ldc, nb, 16);
  up(dP, dcprod, n, lddc, 3, 16);
  std::vector<double> h, w;
  if (nfrqh) {
    up(dW[0], wh, n, n, nfrqh, 8);
    if (whh) up(dW[1], whh, n, n, nfrqh, 8);
    if (!rc)
      rc = spexhip_headwings_dev(ctx, n, nb, dM, n, 1, dP, n, nfrqh, dW[0], whh ? dW[1] : nullptr,
                                 n, 1.0, 0.0, dH, dWg);
    h.resize(18 * (size_t)nfrqh);
    w.resize(6 * (size_t)nb * nfrqh);
    down(h.data(), dH, 8 * h.size());
    down(w.data(), dWg, 8 * w.size());
    for (int f = 0; !rc && f < nfrqh; f++) {  // packed upper, ij = k1 + k2(k2+1)/2
      int ij = 0;
      for (int k2 = 0; k2 < 3; k2++)
        for (int k1 = 0; k1 <= k2; k1++, ij++) {
          headh[2 * (6 * (size_t)f + ij)] += h[2 * (k1 + 3 * k2 + 9 * (size_t)f)];
          headh[2 * (6 * (size_t)f + ij) + 1] += h[2 * (k1 + 3 * k2 + 9 * (size_t)f) + 1];
        }
    }
    for (size_t i = 0; !rc && i < w.size(); i++) wingh[i] += w[i];
  }
  if (nfrqc && !rc) {
    up(dW[0], wcr, n, n, nfrqc, 8);
    up(dW[1], wci, n, n, nfrqc, 8);
    if (wchr) {
      up(dW[2], wchr, n, n, nfrqc, 8);
      up(dW[3], wchi, n, n, nfrqc, 8);
    }
    if (!rc)
      rc = spexhip_headwings_cplx_dev(ctx, n, nb, dM, n, 1, dP, n, nfrqc, dW[0], dW[1],
                                      wchr ? dW[2] : nullptr, wchr ? dW[3] : nullptr, n, 1.0, 0.0,
                                      dH, dWg);
    h.resize(18 * (size_t)nfrqc);
    w.resize(12 * (size_t)nb * nfrqc);
    down(h.data(), dH, 8 * h.size());
    down(w.data(), dWg, 8 * w.size());
    for (size_t i = 0; !rc && i < h.size(); i++) headc[i] += h[i];
    for (size_t i = 0; !rc && i < w.size(); i++) wingc[i] += w[i];
  }
  (void)hipStreamSynchronize(ctx->stream);
  for (void *p : {dM, dP, dW[0], dW[1], dW[2], dW[3], dH, dWg})
    if (p) spx_free(ctx, p);
  return rc;
}

// Host drop-in for correlation.f:570-759 at the Gamma point on the
// imaginary axis: chi (packed upper, nb(nb+1)/2 x nw, after the spin
// factor) -> eps^-1 (packed), with the head (packed 6 x nw) and wings
// (3 x nb x nw) replaced by those of the inverse.  v^1/2 is sqrt(coul)
// (the Coulomb eigenbasis).  SPEXHIP_E_INVALID if the head update is not
// Hermitian to 1e-6 (packmat's test, w_packmat.inc:52-73).
int spexhip_dielectric_inverse_gamma(spexhip_ctx *ctx, int nb, int nw, const double *chi_packed,
                                     const double *sqcoul, double *head, double *wing,
                                     double *einv_packed) {
  if (!ctx) return SPEXHIP_E_INVALID;
  SPX_CHECK(ctx, nb > 0 && nw > 0 && chi_packed && sqcoul && head && wing && einv_packed,
            "dielectric_inverse_gamma: bad arguments");
  SPX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t xb = 16 * (size_t)nb * nb * nw, hb = 16 * 9 * (size_t)nw,
               wb = 16 * 3 * (size_t)nb * nw;
  void *X = nullptr, *V = nullptr, *dH = nullptr, *dW = nullptr;
  int rc = spx_alloc(ctx, xb, &X);
  if (!rc) rc = spx_alloc(ctx, 8 * (size_t)nb, &V);
  if (!rc) rc = spx_alloc(ctx, hb, &dH);
  if (!rc) rc = spx_alloc(ctx, wb, &dW);
  std::vector<double> hf(18 * (size_t)nw);
  for (int f = 0; f < nw; f++) {  // packed upper -> full Hermitian 3 x 3
    int ij = 0;
    for (int k2 = 0; k2 < 3; k2++)
      for (int k1 = 0; k1 <= k2; k1++, ij++) {
        const double re = head[2 * (6 * (size_t)f + ij)], im = head[2 * (6 * (size_t)f + ij) + 1];
        hf[2 * (k1 + 3 * k2 + 9 * (size_t)f)] = re;
        hf[2 * (k1 + 3 * k2 + 9 * (size_t)f) + 1] = im;
        hf[2 * (k2 + 3 * k1 + 9 * (size_t)f)] = re;
        hf[2 * (k2 + 3 * k1 + 9 * (size_t)f) + 1] = k1 == k2 ? im : -im;
      }
  }
  if (!rc) rc = spexhip_unpack_h2d(ctx, nb, nw, chi_packed, X, nb);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(V, sqcoul, 8 * (size_t)nb, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dH, hf.data(), hb, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dW, wing, wb, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "upload: %s", hipGetErrorString(e));
  }
  if (!rc) rc = spexhip_gamma_prepare_dev(ctx, nb, nw, X, nb, dH, dW, V);
  if (!rc) rc = spexhip_dielectric_dev(ctx, nb, nw, X, nb, 0, V, X, nb);
  if (!rc) rc = spexhip_inverse_hpd_dev(ctx, nb, nw, X, nb);
  if (!rc) rc = spexhip_gamma_inverse_dev(ctx, nb, nw, X, nb, dH, dW);
  if (!rc) rc = spexhip_pack_upper_d2h(ctx, nb, nw, X, nb, einv_packed, 0);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(hf.data(), dH, hb, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(wing, dW, wb, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = spx_seterr(ctx, SPEXHIP_E_DEVICE, "download: %s", hipGetErrorString(e));
  }
  for (int f = 0; !rc && f < nw; f++) {
    int ij = 0;
    for (int k2 = 0; k2 < 3 && !rc; k2++)
      for (int k1 = 0; k1 <= k2; k1++, ij++) {
        const double r1 = hf[2 * (k1 + 3 * k2 + 9 * (size_t)f)], i1 = hf[2 * (k1 + 3 * k2 + 9 * (size_t)f) + 1];
        const double r2 = hf[2 * (k2 + 3 * k1 + 9 * (size_t)f)], i2 = -hf[2 * (k2 + 3 * k1 + 9 * (size_t)f) + 1];
        double dev = std::fabs(r1 - r2) + std::fabs(i1 - i2);
        const double mag = std::fabs(r1) + std::fabs(r2) + std::fabs(i1) + std::fabs(i2);
        if (mag > 1) dev /= mag;
        if (dev > 1e-6) {
          rc = spx_seterr(ctx, SPEXHIP_E_INVALID,
                          "dielectric_inverse_gamma: head not Hermitian (deviation %.1e, frequency %d)",
                          dev, f);
          break;
        }
        head[2 * (6 * (size_t)f + ij)] = r1;
        head[2 * (6 * (size_t)f + ij) + 1] = i1;
      }
  }
  (void)hipStreamSynchronize(ctx->stream);
  for (void *p : {X, V, dH, dW})
    if (p) spx_free(ctx, p);
  return rc;
}

}  // extern "C"
