// chi0 pair sum in the Ozaki int8 form (exact integer accumulation).
//
// Same quantity as chi0.hip (susceptibility.f:1159-1251 with the HILBERT
// NONE spectral function :1780-1788):
//     chi(I,J,w) = sum_t D(t,w) M(t,I) conj(M(t,J)).
// When D(.,w) has one sign (always on the imaginary axis for HILBERT NONE,
// D = -2d/(w^2+d^2)/nkpt < 0), write D = sgn |D| and B(t,I) = sqrt|D(t,w)|
// M(t,I), so chi = sgn * B^T conj(B) -- a Hermitian rank-k update of one
// matrix per frequency, with the weight dynamic range halved on each side.
// Each real component of B is split per column I with scale
// s(I,w) = max_t max(|Re B|, |Im B|):
//     B/s = sum_{p=1..4} a_p w_p + r,  w_p = 1/(127*254^(p-1)),
//     a_1 in [-127,127], a_2..a_4 in [-127,127], |r| <= 0.5*w_4,
// (rounding to nearest at every level), so the representation error is
// below 2.4e-10 * s.  Products of slices with p+q <= 5 are accumulated
// exactly in int32 on v_mfma_i32_16x16x64_i8 -- one int32 accumulator per
// level l = p+q, since w_p w_q depends on p+q only -- and flushed into FP64
// accumulators (weight w_l) every 4096 transitions, far below the int32
// overflow bound (4 pairs * 2 products * 127^2 * 4096 < 2^31).
// Re chi = sum Br Br + Bi Bi,  Im chi = sum Bi Br - Br Bi (negated Br on
// the row side), then chi = sgn * s_I s_J * (...).
//
// Layout of the slice array (one frequency):
//   S[chunk c][part][slice p][I (nbp)][64 bytes of t = 64c .. 64c+63],
//   part 0 = Re, 1 = Im;  so a 64-transition chunk of a 64-column panel is
//   8 contiguous 4 KB blocks.
#include "common.hpp"

typedef int i4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int NS = 4;      // slices
constexpr int TK = 64;     // transitions per chunk (MFMA K)
constexpr int TT = 64;     // output tile edge
constexpr int NT = 256;
constexpr int ROWB = 80;   // LDS row stride in bytes (64 + 16 pad: no bank conflicts)
constexpr int FLUSH = 64;  // chunks between int32 -> FP64 flushes

__device__ inline unsigned neg_bytes(unsigned w) {
  // per-byte two's complement negation without cross-byte carries
  const unsigned nw = ~w;
  return ((nw & 0x7F7F7F7Fu) + 0x01010101u) ^ (nw & 0x80808080u);
}

// sqd[w][t] = sqrt|D(t,w)|, and sign flags per frequency (bit0: D>0 seen,
// bit1: D<0 seen).
__global__ void ozk_sqd_kernel(long n, int nw, const double *__restrict__ D,
                               double *__restrict__ sqd, int *__restrict__ flags) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * nw) return;
  const double d = D[g];
  sqd[g] = sqrt(fabs(d));
  const int w = (int)(g / n);
  if (d > 0) atomicOr(flags + w, 1);
  if (d < 0) atomicOr(flags + w, 2);
}

// scale[w][I] = max_t sqd[w][t] * max(|Re M|, |Im M|), via atomicMax on the
// bit pattern (non-negative doubles order like unsigned integers).
template <bool COLMAJOR>
__global__ void ozk_scale_kernel(long n, int nb, const double2 *__restrict__ M, long ldm, int nw,
                                 const double *__restrict__ sqd, long tchunk,
                                 unsigned long long *__restrict__ scale) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  const long t0 = (long)blockIdx.y * tchunk;
  const long t1 = t0 + tchunk < n ? t0 + tchunk : n;
  double m[16];
#pragma unroll
  for (int w = 0; w < 16; w++) m[w] = 0.0;
  for (long t = t0; t < t1; t++) {
    const double2 v = COLMAJOR ? M[t + ldm * (long)I] : M[t * ldm + I];
    const double a = fmax(fabs(v.x), fabs(v.y));
#pragma unroll
    for (int w = 0; w < 16; w++)
      if (w < nw) m[w] = fmax(m[w], a * sqd[(long)w * n + t]);
  }
#pragma unroll
  for (int w = 0; w < 16; w++)
    if (w < nw) atomicMax(scale + (long)w * nb + I, (unsigned long long)__double_as_longlong(m[w]));
}

__device__ inline void slice4(double u, signed char a[NS]) {
  // u in [-127, 127]
  double r = u;
#pragma unroll
  for (int p = 0; p < NS; p++) {
    const double q = rint(r);
    a[p] = (signed char)(int)q;
    r = (r - q) * 254.0;
  }
}

// Slices for frequencies w0 .. w0+nwg-1 into S (nwg consecutive arrays of
// sbytes each).  One thread: one column I, 16 consecutive transitions.
template <bool COLMAJOR>
__global__ void ozk_slice_kernel(long n, int nb, int nbp, const double2 *__restrict__ M, long ldm,
                                 int w0, int nwg, const double *__restrict__ sqd,
                                 const double *__restrict__ scale, unsigned char *__restrict__ S,
                                 long sbytes) {
  const int I = blockIdx.y * blockDim.x + threadIdx.x;
  if (I >= nbp) return;
  const long c = blockIdx.x >> 2;
  const int g16 = (blockIdx.x & 3) * 16;
  const long tb = c * TK + g16;
  double2 v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const long t = tb + j;
    v[j] = (I < nb && t < n) ? (COLMAJOR ? M[t + ldm * (long)I] : M[t * ldm + I])
                             : make_double2(0.0, 0.0);
  }
  for (int wg = 0; wg < nwg; wg++) {
    const int w = w0 + wg;
    const double s = I < nb ? scale[(long)w * nb + I] : 0.0;
    const double inv = s > 0.0 ? 127.0 / s : 0.0;
    unsigned char out[2][NS][16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const long t = tb + j;
      const double f = t < n ? sqd[(long)w * n + t] * inv : 0.0;
      signed char a[NS];
      slice4(v[j].x * f, a);
#pragma unroll
      for (int p = 0; p < NS; p++) out[0][p][j] = (unsigned char)a[p];
      slice4(v[j].y * f, a);
#pragma unroll
      for (int p = 0; p < NS; p++) out[1][p][j] = (unsigned char)a[p];
    }
    unsigned char *base = S + (long)wg * sbytes + c * (2L * NS * nbp * TK);
#pragma unroll
    for (int part = 0; part < 2; part++)
#pragma unroll
      for (int p = 0; p < NS; p++) {
        uint4 q;
        q.x = out[part][p][0] | (out[part][p][1] << 8) | (out[part][p][2] << 16) | ((unsigned)out[part][p][3] << 24);
        q.y = out[part][p][4] | (out[part][p][5] << 8) | (out[part][p][6] << 16) | ((unsigned)out[part][p][7] << 24);
        q.z = out[part][p][8] | (out[part][p][9] << 8) | (out[part][p][10] << 16) | ((unsigned)out[part][p][11] << 24);
        q.w = out[part][p][12] | (out[part][p][13] << 8) | (out[part][p][14] << 16) | ((unsigned)out[part][p][15] << 24);
        *(uint4 *)(base + ((long)(part * NS + p) * nbp + I) * TK + g16) = q;
      }
  }
}

__device__ inline void tri_index(int x, int &bi, int &bj) {
  int j = (int)((sqrtf(8.f * (float)x + 1.f) - 1.f) * 0.5f);
  while ((j + 1) * (j + 2) / 2 <= x) j++;
  while (j * (j + 1) / 2 > x) j--;
  bj = j;
  bi = x - j * (j + 1) / 2;
}

// Rank-k update on int8 slices.  LDS per chunk: row panel (Re, Im, -Re) and
// column panel (Re, Im), NS slices each, 64 rows of 64 bytes (stride ROWB).
__global__ __launch_bounds__(NT) void ozk_herk_kernel(
    long nchunk, int nb, int nbp, const unsigned char *__restrict__ S, long sbytes, int w0,
    const double *__restrict__ scale, const double *__restrict__ sgn, double2 *__restrict__ chi,
    long ldc, double alpha, double beta) {
  __shared__ __attribute__((aligned(16))) unsigned char Ls[5 * NS * TT * ROWB];  // 100 KB
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1, lr = lane & 15, lk = lane >> 4;
  int bi, bj;
  tri_index(blockIdx.x, bi, bj);
  const int wg = blockIdx.y, w = w0 + wg;
  const int I0 = bi * TT, J0 = bj * TT;
  const unsigned char *Sw = S + (long)wg * sbytes;
  const long cstride = 2L * NS * nbp * TK;

  // chunk loader: 8 planes x 2 panels x 64 rows x 64 B = 64 KB = 256 threads x 16 uint4
  uint4 pre[16];
  auto fetch = [&](long c) {
    const unsigned char *base = Sw + c * cstride;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int e = tid + NT * r;      // 0 .. 4095 uint4 slots
      const int panel = e >> 11;       // 0: rows (I), 1: cols (J)
      const int q = e & 2047;
      const int plane = q >> 8;        // part*NS + p
      const int row = (q >> 2) & 63;
      const int seg = q & 3;
      const int col = (panel ? J0 : I0) + row;
      pre[r] = *(const uint4 *)(base + ((long)plane * nbp + col) * TK + seg * 16);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int e = tid + NT * r;
      const int panel = e >> 11;
      const int q = e & 2047;
      const int plane = q >> 8;
      const int row = (q >> 2) & 63;
      const int seg = q & 3;
      // planes 0..NS-1 Re, NS..2NS-1 Im; row panel also keeps -Re at 2NS..3NS-1
      const int lp = panel ? 3 * NS + plane : plane;
      *(uint4 *)(Ls + (lp * TT + row) * ROWB + seg * 16) = pre[r];
      if (!panel && plane < NS) {
        uint4 ng;
        ng.x = neg_bytes(pre[r].x);
        ng.y = neg_bytes(pre[r].y);
        ng.z = neg_bytes(pre[r].z);
        ng.w = neg_bytes(pre[r].w);
        *(uint4 *)(Ls + ((2 * NS + plane) * TT + row) * ROWB + seg * 16) = ng;
      }
    }
  };

  i4_t ar[NS][2][2], ai[NS][2][2];  // level p+q = 2..5 stored at index p+q-2
  d4_t fr[2][2], fi[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      fr[a][b] = d4_t{0, 0, 0, 0};
      fi[a][b] = d4_t{0, 0, 0, 0};
#pragma unroll
      for (int l = 0; l < NS; l++) {
        ar[l][a][b] = i4_t{0, 0, 0, 0};
        ai[l][a][b] = i4_t{0, 0, 0, 0};
      }
    }
  auto flush = [&]() {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          // level weights w_p w_q = 1/(127^2 254^(p+q-2)); small terms first
          double sr = 0.0, si = 0.0;
#pragma unroll
          for (int l = NS - 1; l >= 0; l--) {
            double lw = 1.0 / (127.0 * 127.0);
#pragma unroll
            for (int e = 0; e < l; e++) lw *= 1.0 / 254.0;
            sr += lw * (double)ar[l][a][b][r];
            si += lw * (double)ai[l][a][b][r];
          }
          fr[a][b][r] += sr;
          fi[a][b][r] += si;
        }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int l = 0; l < NS; l++) {
          ar[l][a][b] = i4_t{0, 0, 0, 0};
          ai[l][a][b] = i4_t{0, 0, 0, 0};
        }
  };

  fetch(0);
  stash();
  __syncthreads();
  for (long c = 0; c < nchunk; c++) {
    if (c + 1 < nchunk) fetch(c + 1);
    // column operands for all slices: Re (planes 3NS+q) and Im (planes 4NS+q)
    i4_t Br[NS][2], Bi[NS][2];
#pragma unroll
    for (int q = 0; q < NS; q++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int row = wc * 32 + b * 16 + lr;
        Br[q][b] = *(const i4_t *)(Ls + ((3 * NS + q) * TT + row) * ROWB + lk * 16);
        Bi[q][b] = *(const i4_t *)(Ls + ((4 * NS + q) * TT + row) * ROWB + lk * 16);
      }
#pragma unroll
    for (int p = 0; p < NS; p++) {
      i4_t Ar[2], Ai[2], An[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int row = wr * 32 + a * 16 + lr;
        Ar[a] = *(const i4_t *)(Ls + ((0 * NS + p) * TT + row) * ROWB + lk * 16);
        Ai[a] = *(const i4_t *)(Ls + ((1 * NS + p) * TT + row) * ROWB + lk * 16);
        An[a] = *(const i4_t *)(Ls + ((2 * NS + p) * TT + row) * ROWB + lk * 16);
      }
#pragma unroll
      for (int q = 0; q < NS - p; q++) {
        const int l = p + q;  // slices p+1, q+1: level p+q+2 -> index p+q
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            ar[l][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ar[a], Br[q][b], ar[l][a][b], 0, 0, 0);
            ar[l][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ai[a], Bi[q][b], ar[l][a][b], 0, 0, 0);
            ai[l][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ai[a], Br[q][b], ai[l][a][b], 0, 0, 0);
            ai[l][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(An[a], Bi[q][b], ai[l][a][b], 0, 0, 0);
          }
      }
    }
    if ((c + 1) % FLUSH == 0) flush();
    __syncthreads();
    if (c + 1 < nchunk) {
      stash();
      __syncthreads();
    }
  }
  flush();

  // v_mfma_i32_16x16x64_i8 result layout: column j = lane%16, rows 4*(lane/16)+r.
  double2 *C = chi + (long)w * ldc * nb;
  const double sg = sgn[w] * alpha;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int J = J0 + wc * 32 + b * 16 + lr;
      if (J >= nb) continue;
      const double sJ = scale[(long)w * nb + J];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int I = I0 + wr * 32 + a * 16 + 4 * lk + r;
        if (I >= nb) continue;
        const double f = sg * scale[(long)w * nb + I] * sJ;
        double2 *p = C + I + ldc * (long)J;
        double2 o = make_double2(f * fr[a][b][r], f * fi[a][b][r]);
        if (beta != 0.0) {
          const double2 old = *p;
          o.x += beta * old.x;
          o.y += beta * old.y;
        }
        *p = o;
      }
    }
}

}  // namespace

// Workspace of the Ozaki form for a frequency group of nwg.
size_t spx_ozaki_slice_bytes(long n, int nb) {
  const long nchunk = (n + TK - 1) / TK;
  const int nbp = ((nb + TT - 1) / TT) * TT;
  return (size_t)nchunk * 2 * NS * nbp * TK;
}

// Returns SPEXHIP_E_INVALID (with a message) when some frequency has
// weights of both signs; the caller then uses another precision.
int spx_chi0_ozaki(spexhip_ctx *ctx, hipStream_t st, long n, int nb, const void *M, long ldm,
                   int colmajor, int nw, const void *wght, double alpha, double beta, void *chi,
                   int ldchi, size_t ws_limit) {
  SPX_CHECK(ctx, nw <= 16, "chi0 ozaki: at most 16 frequencies per call (got %d)", nw);
  if (n == 0) return spx_chi0_launch(ctx, st, 0, nb, M, ldm, colmajor, nw, wght,
                                     SPEXHIP_PREC_FP64, alpha, beta, chi, ldchi);
  const long nchunk = (n + TK - 1) / TK;
  const int nbp = ((nb + TT - 1) / TT) * TT;
  const size_t sb = spx_ozaki_slice_bytes(n, nb);
  // frequency group size bounded by the workspace limit
  int nwg = nw;
  while (nwg > 1 && sb * nwg > ws_limit) nwg--;
  size_t fixed = 8 * (size_t)n * nw + 8 * (size_t)nb * nw + 8 * (size_t)nw + 4 * 16;
  fixed = (fixed + 255) & ~(size_t)255;
  void *ws;
  int rc = spx_scratch(ctx, fixed + sb * nwg, &ws);
  if (rc) return rc;
  double *sqd = (double *)ws;
  double *scale = sqd + (size_t)n * nw;
  double *sgn = scale + (size_t)nb * nw;
  int *flags = (int *)(sgn + nw);
  unsigned char *S = (unsigned char *)ws + fixed;
  SPX_HIP(ctx, hipMemsetAsync(flags, 0, 4 * 16, st));
  SPX_HIP(ctx, hipMemsetAsync(scale, 0, 8 * (size_t)nb * nw, st));
  {
    const long tot = n * nw;
    hipLaunchKernelGGL(ozk_sqd_kernel, dim3((tot + 255) / 256), dim3(256), 0, st, n, nw,
                       (const double *)wght, sqd, flags);
    SPX_LAUNCHED(ctx);
  }
  int hflags[16];
  SPX_HIP(ctx, hipMemcpyAsync(hflags, flags, 4 * nw, hipMemcpyDeviceToHost, st));
  SPX_HIP(ctx, hipStreamSynchronize(st));
  double hsgn[16];
  for (int w = 0; w < nw; w++) {
    if (hflags[w] == 3)
      return spx_seterr(ctx, SPEXHIP_E_INVALID,
                        "chi0 ozaki: weights of frequency %d have both signs", w);
    hsgn[w] = hflags[w] == 1 ? 1.0 : -1.0;
  }
  SPX_HIP(ctx, hipMemcpyAsync(sgn, hsgn, 8 * nw, hipMemcpyHostToDevice, st));
  {
    const long tchunk = 4096;
    dim3 grid((nb + 255) / 256, (n + tchunk - 1) / tchunk);
    if (colmajor)
      hipLaunchKernelGGL(ozk_scale_kernel<true>, grid, dim3(256), 0, st, n, nb, (const double2 *)M,
                         ldm, nw, sqd, tchunk, (unsigned long long *)scale);
    else
      hipLaunchKernelGGL(ozk_scale_kernel<false>, grid, dim3(256), 0, st, n, nb, (const double2 *)M,
                         ldm, nw, sqd, tchunk, (unsigned long long *)scale);
    SPX_LAUNCHED(ctx);
  }
  const int nbt = nbp / TT;
  if (ctx->phase_ev[0]) SPX_HIP(ctx, hipEventRecord(ctx->phase_ev[0], st));
  for (int w0 = 0; w0 < nw; w0 += nwg) {
    const int g = nw - w0 < nwg ? nw - w0 : nwg;
    dim3 sgrid(nchunk * 4, (nbp + 255) / 256);
    if (colmajor)
      hipLaunchKernelGGL(ozk_slice_kernel<true>, sgrid, dim3(256), 0, st, n, nb, nbp,
                         (const double2 *)M, ldm, w0, g, sqd, scale, S, (long)sb);
    else
      hipLaunchKernelGGL(ozk_slice_kernel<false>, sgrid, dim3(256), 0, st, n, nb, nbp,
                         (const double2 *)M, ldm, w0, g, sqd, scale, S, (long)sb);
    SPX_LAUNCHED(ctx);
    if (w0 == 0 && ctx->phase_ev[1]) SPX_HIP(ctx, hipEventRecord(ctx->phase_ev[1], st));
    dim3 hgrid(nbt * (nbt + 1) / 2, g);
    hipLaunchKernelGGL(ozk_herk_kernel, hgrid, dim3(NT), 0, st, nchunk, nb, nbp, S, (long)sb, w0,
                       scale, sgn, (double2 *)chi, (long)ldchi, alpha, beta);
    SPX_LAUNCHED(ctx);
  }
  if (ctx->phase_ev[2]) SPX_HIP(ctx, hipEventRecord(ctx->phase_ev[2], st));
  ctx->phase_groups = (nw + nwg - 1) / nwg;
  // the host copy of the signs must outlive the async copy
  SPX_HIP(ctx, hipStreamSynchronize(st));
  return SPEXHIP_OK;
}
