// chi0 pair sum in the Ozaki int8 form (exact integer accumulation).
//
// Same quantity as chi0.hip (susceptibility.f:1159-1251 with the HILBERT
// NONE spectral function :1780-1788):
//     chi(I,J,w) = sum_t D(t,w) M(t,I) conj(M(t,J)).
// When D(.,w) has one sign (always on the imaginary axis for HILBERT NONE,
// D = -2d/(w^2+d^2)/nkpt < 0), write D = sgn |D| and B(t,I) = sqrt|D(t,w)|
// M(t,I), so chi = sgn * B^T conj(B) -- a Hermitian rank-k update of one
// matrix per frequency, with the weight dynamic range split between the
// two sides.  Each real component of B is split per column I with scale
// s(I,w) = max_t max(|Re B|, |Im B|):
//     B/s = sum_{p=1..4} a_p w_p + r,  w_p = 1/(127*254^(p-1)),
//     a_p in [-127,127], |r| <= 0.5*w_4,
// (rounding to nearest at every level), so the representation error is
// below 2.4e-10 * s.  Products of slices with p+q <= 5, plus the real
// part of the pair (3,3), are accumulated exactly in int32 on
// v_mfma_i32_16x16x64_i8 -- one int32 accumulator per level l = p+q, since
// w_p w_q depends on p+q only -- and flushed into FP64 (weight w_l) every
// 16384 transitions, within the int32 bound: a level gathers at most
// 4 pairs x 2 products of magnitude <= 127^2 per transition, and
// 8 * 127^2 * 16384 = 2.114e9 < 2^31 - 1.  Of the dropped pairs
// (p+q >= 6), those with p != q are products of independent digits and
// average out over the transitions; (3,3) is kept because on the diagonal
// (and for correlated columns) it is a sum of squares, a bias of about
// 254^-4/3 (s/rms)^2 relative to chi.
//
//     Re chi = sum Br_I Br_J + Bi_I Bi_J,  Im chi = sum Bi_I Br_J + Br_I (-Bi_J),
// then chi = sgn * s_I s_J * (...).
//
// Slice layout (one frequency), "MFMA-ready": for 64-transition chunk c,
// plane pl = part*4 + slice (part 0 = Re, 1 = Im) and 16-column block ib,
// a 1 KB block holding lane l's operand bytes at offset 16*l, lane l =
// row (l%16) of the block and transitions 64c + 16*(l/16) .. +15 -- the
// operand layout of v_mfma_i32_16x16x64_i8.  The rank-k kernel copies
// these blocks straight into LDS with global_load_lds_dwordx4 (one block
// per wave-instruction) and reads each operand with one conflict-free
// 16-byte LDS load per lane.
#include <algorithm>
#include <type_traits>

#include "colscale.hpp"
#include "common.hpp"

typedef int i4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int NS = 4;      // slices
constexpr int TK = 64;     // transitions per chunk (MFMA K)
constexpr int TT = 64;     // output tile edge
constexpr int NT = 256;
constexpr int FLUSH = 256;  // chunks (of 64 transitions) between int32 -> FP64 flushes
static_assert(8LL * 127 * 127 * 64 * FLUSH <= 2147483647LL, "int32 level accumulators overflow");
constexpr int PFD = 1;     // L2 prefetch distance in chunks (1: none beyond the double buffer)
constexpr int OZK_WAVES = 8;  // waves per rank-k workgroup (4 or 8)
constexpr int OZK_KSPLIT = 2;  // transition parts: slicing part i+1 overlaps the rank-k of part i
static_assert(OZK_KSPLIT >= 1 && OZK_KSPLIT <= 8, "one context event per part");
constexpr int NPL = 2 * NS;  // planes per chunk

__device__ inline long blk_off(long c, int pl, int ib, int nib) {
  return (((long)c * NPL + pl) * nib + ib) * 1024L;
}

// sqd[w][t] = sqrt|D(t,w)| for w < nw, 0 on the padding rows nw <= w <
// nwp (blockIdx.y), and sign flags per frequency (bit0: D>0 seen, bit1:
// D<0 seen).  Grid-stride over t: one LDS atomic per wave and one global
// atomic per workgroup, so few workgroups contend for each flag.
__global__ __launch_bounds__(256) void ozk_sqd_kernel(long n, int nw, const double *__restrict__ D,
                                                      double *__restrict__ sqd,
                                                      int *__restrict__ flags) {
  __shared__ int fl;
  const int w = blockIdx.y;
  if (threadIdx.x == 0) fl = 0;
  __syncthreads();
  bool pos = false, neg = false;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const double d = w < nw ? D[(long)w * n + t] : 0.0;
    sqd[(long)w * n + t] = sqrt(fabs(d));
    pos |= d > 0.0;
    neg |= d < 0.0;
  }
  const unsigned long long bp = __ballot(pos), bn = __ballot(neg);
  if ((threadIdx.x & 63) == 0 && (bp || bn)) atomicOr(&fl, (bp ? 1 : 0) | (bn ? 2 : 0));
  __syncthreads();
  if (threadIdx.x == 0 && fl) atomicOr(flags + w, fl);
}

// Digits of u in [-127, 127]: a[p] = rint(r_p), r_{p+1} = (r_p - a[p]) 254.
// Adding 1.5 * 2^52 to |r| < 2^51 rounds it to an integer (to nearest, ties
// to even, as rint) that the low mantissa bits of the sum then hold in two's
// complement, so one FP64 add replaces the rounding and the conversion; the
// sum minus 1.5 * 2^52 is rint(r) exactly.  Contraction stays off: every
// step rounds as written (four FP64 adds and multiplies per digit).
__device__ inline void slice4(double u, int a[NS]) {
#pragma clang fp contract(off)
  constexpr double kRound = 6755399441055744.0;  // 1.5 * 2^52
  double r = u;
#pragma unroll
  for (int p = 0; p < NS; p++) {
    const double s = r + kRound;
    a[p] = __double2loint(s);
    if (p + 1 < NS) r = (r - (s - kRound)) * 254.0;
  }
}

// Slices of frequencies w0 .. w0+nwg-1 (nwg arrays of sbytes each) for the
// chunks c0, c0+1, ...  One wave per (chunk, 16-column block); lane l = row
// l%16, transitions 64c + 16(l/16) + 0..15.
template <bool COLMAJOR>
__global__ __launch_bounds__(NT) void ozk_slice_kernel(long n, int nb, int nbp,
                                                       const double2 *__restrict__ M, long ldm,
                                                       int w0, int nwg,
                                                       const double *__restrict__ sqd,
                                                       const double *__restrict__ scale,
                                                       unsigned char *__restrict__ S, long sbytes,
                                                       long c0, long nwaves) {
  const long wid = (long)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  if (wid >= nwaves) return;
  const int nib = nbp / 16;
  const long c = c0 + wid / nib;
  const int ib = (int)(wid % nib);
  const int l = threadIdx.x & 63, row = l & 15, g = l >> 4;
  const int I = ib * 16 + row;
  const long tb = c * TK + 16 * g;
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
    unsigned int q[NPL][4];
#pragma unroll
    for (int pl = 0; pl < NPL; pl++)
#pragma unroll
      for (int d = 0; d < 4; d++) q[pl][d] = 0u;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const long t = tb + j;
      const double f = t < n ? sqd[(long)w * n + t] * inv : 0.0;
      int a[NS];
      slice4(v[j].x * f, a);
#pragma unroll
      for (int p = 0; p < NS; p++) q[p][j >> 2] |= ((unsigned)a[p] & 0xFFu) << (8 * (j & 3));
      slice4(v[j].y * f, a);
#pragma unroll
      for (int p = 0; p < NS; p++) q[NS + p][j >> 2] |= ((unsigned)a[p] & 0xFFu) << (8 * (j & 3));
    }
    unsigned char *base = S + (long)wg * sbytes;
#pragma unroll
    for (int pl = 0; pl < NPL; pl++) {
      uint4 o;
      o.x = q[pl][0];
      o.y = q[pl][1];
      o.z = q[pl][2];
      o.w = q[pl][3];
      *(uint4 *)(base + blk_off(c, pl, ib, nib) + 16 * l) = o;
    }
  }
}

// Bytewise negation of 16 int8 in [-127, 127] (two's complement per byte,
// no carries between bytes): -x = (~x & 0x7f..) + 1 with the sign bit of ~x
// xor-ed back in.
__device__ inline i4_t neg_i8x16(i4_t x) {
  i4_t r;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const unsigned n = ~(unsigned)x[k];
    r[k] = (int)(((n & 0x7f7f7f7fu) + 0x01010101u) ^ (n & 0x80808080u));
  }
  return r;
}

__device__ inline void tri_index(int x, int &bi, int &bj) {
  int j = (int)((sqrtf(8.f * (float)x + 1.f) - 1.f) * 0.5f);
  while ((j + 1) * (j + 2) / 2 <= x) j++;
  while (j * (j + 1) / 2 > x) j--;
  bj = j;
  bi = x - j * (j + 1) / 2;
}

// Logical tile index -> (frequency in the group, 64x64 tile bi <= bj):
// frequency-major, and within a frequency 4x4 super-blocks of tiles,
// column of super-blocks by column, so that consecutive tiles share row
// blocks of slices.
__device__ inline void tile_of(long t0, int T, int nbt, int &wg, int &bi, int &bj) {
  wg = (int)(t0 / T);
  int t = (int)(t0 % T);
  bi = bj = 0;
  const int ns = (nbt + 3) / 4;
  for (int SJ = 0; SJ < ns; SJ++) {
    const int cols = min(4, nbt - 4 * SJ);
    for (int SI = 0; SI <= SJ; SI++) {
      const int rows = min(4, nbt - 4 * SI);
      const int cnt = SI < SJ ? rows * cols : rows * (rows + 1) / 2;
      if (t < cnt) {
        if (SI < SJ) {
          bi = 4 * SI + t / cols;
          bj = 4 * SJ + t % cols;
        } else {
          int i, jj;
          tri_index(t, i, jj);
          bi = 4 * SI + i;
          bj = 4 * SI + jj;
        }
        return;
      }
      t -= cnt;
    }
  }
}

typedef __attribute__((address_space(3))) void lds_void;

// Rank-k update on int8 slices: 64x64 tile per workgroup of NWV waves, a
// 2 x NWV/2 grid of 32-row x 64/(NWV/2)-column wave tiles (16x16 blocks).
// Per chunk the workgroup stages 8 planes x 4 row blocks + 8 planes x 4
// column blocks (64 KB) into one of two LDS buffers; each wave issues
// 64/NWV global_load_lds_dwordx4 (1 KB each).  With NWV = 8 two waves share
// each SIMD, so one wave's barrier and LDS waits hide behind the other's
// MFMAs.
//
// Persistent: one workgroup per CU (G of them, a multiple of 8) walks the
// tiles in rounds of G.  Every tile streams the same K, so the tiles of a
// round run side by side, and within a round XCD x (workgroups x, x+8, ...)
// holds G/8 consecutive tiles of the logical order: tiles that share row
// blocks read them through one L2 at nearly the same time.
template <int NWV>
__global__ __launch_bounds__(NWV * 64) void ozk_herk_kernel(
    long nchunk, int nb, int nbp, int ntile, int ngroup, const unsigned char *__restrict__ S,
    long sbytes, int w0, const double *__restrict__ scale, const double *__restrict__ sgn,
    double2 *__restrict__ chi, long ldc, double alpha, double beta) {
  constexpr int WCN = NWV / 2;     // waves across the columns
  constexpr int BB = 4 / WCN;      // 16-column blocks per wave
  constexpr int EPW = 64 / NWV;    // 1 KB copies per wave per chunk
  constexpr int PFI = EPW / 8;     // prefetch instructions per wave per chunk
  static_assert(NWV == 4 || NWV == 8, "4 or 8 waves");
  // [buf][panel][plane][block of 4][1 KB], then 256 B of prefetch scratch
  constexpr int LPF = 2 * 2 * NPL * 4 * 1024;
  __shared__ __attribute__((aligned(1024))) unsigned char Ls[LPF + 256];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv / WCN, wc = wv % WCN;
  const int nib = nbp / 16, nbt = nbp / TT;
  const long ntot = (long)ntile * ngroup, G = gridDim.x;
  const long slot = (long)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
  // current tile (set at the top of each round)
  int bi = 0, bj = 0, wg = 0, w = 0;
  const unsigned char *Sw = S;

  // wave wv loads the blocks e = wv*EPW .. of the 64 per chunk:
  // e -> panel = e/32, plane = (e/4)%8, block = e%4
  auto issue = [&](long c, int buf) {
#pragma unroll
    for (int r = 0; r < EPW; r++) {
      const int e = wv * EPW + r;
      const int panel = e >> 5, pl = (e >> 2) & 7, blk = e & 3;
      const int ibk = (panel ? bj : bi) * 4 + blk;
      const unsigned char *src = Sw + blk_off(c, pl, ibk, nib) + 16 * lane;
      unsigned char *dst = Ls + ((((buf * 2 + panel) * NPL + pl) * 4 + blk) * 1024);
      __builtin_amdgcn_global_load_lds((const void *)src, (lds_void *)dst, 16, 0, 0);
    }
  };
  // L2 prefetch of chunk c: one dword of each 128-byte line of this wave's
  // blocks (8 lines per block, 64 per instruction), dropped into a scratch
  // LDS word per lane.  The double buffer alone gives a copy one chunk of
  // MFMA work to land; with the prefetch PFD-1 chunks ahead the copies hit
  // L2.
  auto prefetch = [&](long c) {
#pragma unroll
    for (int h = 0; h < PFI; h++) {
      const int e = wv * EPW + h * 8 + (lane >> 3);
      const int panel = e >> 5, pl = (e >> 2) & 7, blk = e & 3;
      const int ibk = (panel ? bj : bi) * 4 + blk;
      const unsigned char *src = Sw + blk_off(c, pl, ibk, nib) + 128 * (lane & 7);
      __builtin_amdgcn_global_load_lds((const void *)src, (lds_void *)(Ls + LPF), 4, 0, 0);
    }
  };
  // s_waitcnt with vmcnt <= n (loads retire in issue order, so the copies
  // issued before the last n prefetch instructions have landed), lgkmcnt 0.
  auto wait_copies = [&](int n) {
    if (n >= 4) __builtin_amdgcn_s_waitcnt(0x74);
    else if (n == 3) __builtin_amdgcn_s_waitcnt(0x73);
    else if (n == 2) __builtin_amdgcn_s_waitcnt(0x72);
    else if (n == 1) __builtin_amdgcn_s_waitcnt(0x71);
    else __builtin_amdgcn_s_waitcnt(0x70);
  };
  auto opnd = [&](int buf, int panel, int pl, int blk) -> i4_t {
    return *(const i4_t *)(Ls + ((((buf * 2 + panel) * NPL + pl) * 4 + blk) * 1024) + 16 * lane);
  };

  // level l = p+q (0-based slices) -> one int32 accumulator for Re and one
  // for Im per 16x16 block; Im = Ai.Br + Ar.(-Bi).  r4: Re of the pair of
  // third slices (0-based p = q = 2), the one level-4 term that does not
  // average out (on the diagonal it is a sum of squares).
  i4_t re[NS][2][BB], im[NS][2][BB], r4[2][BB];
  auto zero_acc = [&]() {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < BB; b++) {
        r4[a][b] = i4_t{0, 0, 0, 0};
#pragma unroll
        for (int l = 0; l < NS; l++) {
          re[l][a][b] = i4_t{0, 0, 0, 0};
          im[l][a][b] = i4_t{0, 0, 0, 0};
        }
      }
  };
  // The FP64 accumulation lives in the output itself: each flush adds the
  // scaled level sums to chi (the first one also applies beta), so no FP64
  // accumulators occupy registers during the int8 loop.  The workgroup owns
  // its tile, so the read-modify-write needs no synchronisation.
  // v_mfma_i32_16x16x64_i8 result layout: column j = lane%16, rows 4*(lane/16)+r.
  const int lr = lane & 15, lk = lane >> 4;
  int I0 = 0, J0 = 0;
  double2 *C = chi;
  double sg = 0.0;
  auto flush = [&](bool first) {
#pragma unroll
    for (int b = 0; b < BB; b++) {
      const int J = J0 + b * 16;
      if (J >= nb) continue;
      const double sJ = sg * scale[(long)w * nb + J];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int I = I0 + a * 16 + r;
          if (I >= nb) continue;
          // level weights w_p w_q = 1/(127^2 254^(p+q)); small terms first
          double sr = (double)r4[a][b][r] * (1.0 / (127.0 * 127.0 * 254.0 * 254.0 * 254.0 * 254.0));
          double si = 0.0;
#pragma unroll
          for (int l = NS - 1; l >= 0; l--) {
            double lw = 1.0 / (127.0 * 127.0);
#pragma unroll
            for (int e = 0; e < l; e++) lw *= 1.0 / 254.0;
            sr += lw * (double)re[l][a][b][r];
            si += lw * (double)im[l][a][b][r];
          }
          const double f = sJ * scale[(long)w * nb + I];
          double2 *p = C + I + ldc * (long)J;
          double2 o = make_double2(f * sr, f * si);
          if (!first || beta != 0.0) {
            const double2 old = *p;
            const double bb = first ? beta : 1.0;
            o.x += bb * old.x;
            o.y += bb * old.y;
          }
          *p = o;
        }
    }
  };

  // One chunk of 64 transitions.  FIRST: the first chunk after a flush --
  // every accumulator's first MFMA takes C = 0 (an inline constant), so the
  // accumulators are never zeroed through registers and stay in AGPRs.
  auto chunk = [&](long c, auto first_tag) {
    constexpr bool FIRST = decltype(first_tag)::value;
    const i4_t z4 = {0, 0, 0, 0};
    const int buf = c & 1;
    // this wave's copies of chunk c have landed (prefetches issued after
    // them may still be in flight) ...
    wait_copies(PFD <= 1 ? 0
                : c == 0 ? PFI * (int)(nchunk - 1 < PFD - 1 ? nchunk - 1 : PFD - 1)
                         : (c - 1 + PFD < nchunk ? PFI : 0));
    // ... and every other wave's (a plain barrier: the wait above already
    // made this wave's copies visible, and a fence would drain the
    // prefetches too)
    __builtin_amdgcn_s_barrier();
    if (PFD > 1 && c + PFD < nchunk) prefetch(c + PFD);
    // every column-side operand of the chunk stays in registers; the row
    // side is read once per slice
    // Im = Ai.Br + Ar.(-Bi): the bytewise negation is done once per chunk on
    // the column side (NS blocks) instead of once per slice on the row side
    i4_t Br[NS][BB], Bi[NS][BB], Bn[NS][BB];
#pragma unroll
    for (int q = 0; q < NS; q++)
#pragma unroll
      for (int b = 0; b < BB; b++) {
        Br[q][b] = opnd(buf, 1, q, wc * BB + b);
        Bi[q][b] = opnd(buf, 1, NS + q, wc * BB + b);
        Bn[q][b] = neg_i8x16(Bi[q][b]);
      }
#pragma unroll
    for (int p = 0; p < NS; p++) {
      i4_t Ar[2], Ai[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        Ar[a] = opnd(buf, 0, p, wr * 2 + a);
        Ai[a] = opnd(buf, 0, NS + p, wr * 2 + a);
      }
#pragma unroll
      for (int q = 0; q < NS - p; q++) {
        const int l = p + q;
        const bool z = FIRST && p == 0;  // first pair of level l in this order
        // independent accumulators first, the dependent second halves after
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < BB; b++) {
            re[l][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ar[a], Br[q][b], z ? z4 : re[l][a][b], 0, 0, 0);
            im[l][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ai[a], Br[q][b], z ? z4 : im[l][a][b], 0, 0, 0);
          }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < BB; b++) {
            re[l][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ai[a], Bi[q][b], re[l][a][b], 0, 0, 0);
            im[l][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ar[a], Bn[q][b], im[l][a][b], 0, 0, 0);
          }
      }
      // copies of chunk c+1 into buffer buf^1 (last read before the barrier),
      // issued once the first slice's MFMAs are queued: right after the
      // barrier they would hold back the MFMA pipe while both waves of a
      // SIMD issue their copies; later than this they have too little time
      // to land (measured)
      if (p == 0 && c + 1 < nchunk) issue(c + 1, buf ^ 1);
      if (p == 2) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < BB; b++) {
            r4[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ar[a], Br[2][b], FIRST ? z4 : r4[a][b], 0, 0, 0);
            r4[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Ai[a], Bi[2][b], r4[a][b], 0, 0, 0);
          }
      }
    }
  };

  for (long tix = slot; tix < ntot; tix += G) {
    tile_of(tix, ntile, nbt, wg, bi, bj);
    w = w0 + wg;
    Sw = S + (long)wg * sbytes;
    I0 = bi * TT + wr * 32 + 4 * lk;
    J0 = bj * TT + wc * (16 * BB) + lr;
    C = chi + (long)w * ldc * nb;
    sg = sgn[w] * alpha;
    __syncthreads();  // the previous tile's LDS reads are done before new copies land
    issue(0, 0);
    for (long c = 1; c < PFD && c < nchunk; c++) prefetch(c);
    for (long c0 = 0; c0 < nchunk; c0 += FLUSH) {
      const long ce = c0 + FLUSH < nchunk ? c0 + FLUSH : nchunk;
      chunk(c0, std::true_type{});
      for (long c = c0 + 1; c < ce; c++) chunk(c, std::false_type{});
      flush(c0 == 0);
    }
    if (nchunk == 0) {  // no transitions: chi = beta * chi
      zero_acc();
      flush(true);
    }
  }
}

// sgn[w] = +1 / -1 from the sign flags; a frequency with weights of both
// signs gets sgn = 0 and raises *err (checked by the caller after the run).
__global__ void ozk_sign_kernel(int nw, const int *__restrict__ flags, double *__restrict__ sgn,
                                int *__restrict__ err) {
  const int w = threadIdx.x;
  if (w >= nw) return;
  const int f = flags[w];
  sgn[w] = f == 3 ? 0.0 : (f == 1 ? 1.0 : -1.0);
  if (f == 3) err[0] = 1;
}

__global__ void ozk_zero_kernel(unsigned long long *__restrict__ p, long count) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < count) p[g] = 0ull;
}

}  // namespace

// Slice bytes of one frequency.
size_t spx_ozaki_slice_bytes(long n, int nb) {
  const long nchunk = (n + TK - 1) / TK;
  const int nbp = ((nb + TT - 1) / TT) * TT;
  return (size_t)nchunk * NPL * nbp * TK;
}

// sqd rows are padded to a multiple of 4 frequencies (see col_scale_kernel)
static int ozk_nwp(int nw) { return (nw + 3) / 4 * 4; }

static size_t ozaki_fixed_bytes(long n, int nb, int nw) {
  size_t fixed = 8 * (size_t)n * ozk_nwp(nw) + 8 * (size_t)nb * nw + 8 * (size_t)nw + 4 * 16;
  return (fixed + 1023) & ~(size_t)1023;
}

// Workspace of the Ozaki form when nwg frequencies are sliced at a time.
size_t spx_ozaki_ws_bytes(long n, int nb, int nw, int nwg) {
  return ozaki_fixed_bytes(n, nb, nw) + spx_ozaki_slice_bytes(n, nb) * (size_t)nwg;
}

// Asynchronous Ozaki chi0 on stream st with caller workspace ws (ws_bytes >=
// spx_ozaki_ws_bytes(n, nb, nw, 1), ws 1 KB aligned).  No host
// synchronisation and no memset nodes, so it may be captured in a graph.
// *err_dev (device int, caller zeroed) becomes non-zero if some frequency
// has weights of both signs.
static int ozk_ncu(spexhip_ctx *ctx) {
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
      n <= 0)
    n = 256;
  return n;
}

int spx_chi0_ozaki_async(spexhip_ctx *ctx, hipStream_t st, long n, int nb, const void *M,
                         long ldm, int colmajor, int nw, const void *wght, double alpha,
                         double beta, void *chi, int ldchi, void *ws, size_t ws_bytes,
                         int *err_dev) {
  SPX_CHECK(ctx, nw >= 1 && nw <= 16, "chi0 ozaki: 1..16 frequencies per call (got %d)", nw);
  if (n == 0) return spx_chi0_launch(ctx, st, 0, nb, M, ldm, colmajor, nw, wght,
                                     SPEXHIP_PREC_FP64, alpha, beta, chi, ldchi);
  SPX_CHECK(ctx, ((uintptr_t)ws & 1023) == 0, "chi0 ozaki: workspace not 1 KB aligned");
  const long nchunk = (n + TK - 1) / TK;
  const int nbp = ((nb + TT - 1) / TT) * TT;
  const size_t sb = spx_ozaki_slice_bytes(n, nb);
  const size_t fixed = ozaki_fixed_bytes(n, nb, nw);
  SPX_CHECK(ctx, ws_bytes >= fixed + sb, "chi0 ozaki: workspace too small");
  const int nwg = (int)std::min<size_t>((size_t)nw, (ws_bytes - fixed) / sb);
  double *sqd = (double *)ws;
  const int nwp = ozk_nwp(nw);
  double *scale = sqd + (size_t)n * nwp;
  double *sgn = scale + (size_t)nb * nw;
  int *flags = (int *)(sgn + nw);
  unsigned char *S = (unsigned char *)ws + fixed;
  const bool timed = st == ctx->stream;
  // zero the sign flags (a fill kernel: capture-safe)
  hipLaunchKernelGGL(ozk_zero_kernel, dim3(1), dim3(64), 0, st, (unsigned long long *)flags, 8L);
  SPX_LAUNCHED(ctx);
  hipLaunchKernelGGL(ozk_sqd_kernel, dim3((unsigned)std::min<long>((n + 255) / 256, 128), nwp),
                     dim3(256), 0, st, n, nw, (const double *)wght, sqd, flags);
  SPX_LAUNCHED(ctx);
  hipLaunchKernelGGL(ozk_sign_kernel, dim3(1), dim3(64), 0, st, nw, flags, sgn, err_dev);
  SPX_LAUNCHED(ctx);
  {
    const int rc = spx_col_scales(ctx, st, n, nb, M, ldm, colmajor, nw, sqd, 0, scale);
    if (rc) return rc;
  }
  const int nbt = nbp / TT;
  const int nib = nbp / 16;
  const int ntile = nbt * (nbt + 1) / 2;
  const size_t cbytes = (size_t)NPL * nib * 1024;  // slice bytes of one chunk, one frequency
  auto slice = [&](hipStream_t s2, int w0, int g, long c0, long c1) {
    const long nwaves = (c1 - c0) * nib;
    dim3 sgrid((unsigned)((nwaves + NT / 64 - 1) / (NT / 64)));
    if (colmajor)
      hipLaunchKernelGGL(ozk_slice_kernel<true>, sgrid, dim3(NT), 0, s2, n, nb, nbp,
                         (const double2 *)M, ldm, w0, g, sqd, scale, S, (long)sb, c0, nwaves);
    else
      hipLaunchKernelGGL(ozk_slice_kernel<false>, sgrid, dim3(NT), 0, s2, n, nb, nbp,
                         (const double2 *)M, ldm, w0, g, sqd, scale, S, (long)sb, c0, nwaves);
  };
  auto herk = [&](int w0, int g, long c0, long c1, double b) {
    // persistent grid: one workgroup per CU (128 KB of LDS each), a multiple of 8
    const long want = ((long)ntile * g + 7) / 8 * 8;
    const int grid = (int)std::min<long>(want, std::max(8, ozk_ncu(ctx) / 8 * 8));
    hipLaunchKernelGGL(ozk_herk_kernel<OZK_WAVES>, dim3(grid), dim3(OZK_WAVES * 64), 0, st,
                       c1 - c0, nb, nbp, ntile, g, S + c0 * cbytes, (long)sb, w0, scale, sgn,
                       (double2 *)chi, (long)ldchi, alpha, b);
  };
  if (timed) SPX_HIP(ctx, hipEventRecord(ctx->phase_ev[0], st));
  // One frequency group on the context stream: the transitions run in
  // OZK_KSPLIT parts, and the slices of part i+1 (HBM-bound) are formed on a
  // second stream while the rank-k of part i (MFMA-bound) runs; every part
  // after the first accumulates (beta = 1).  Measured at 409,600 transitions:
  // the exposed slicing halves (13.8 -> 6.9 ms) but the overlapped half slows
  // the rank-k by 5.6 ms, a net 2.3 ms; three or four parts gain no more.
  const int parts = (nwg >= nw && st == ctx->stream && nchunk >= 2 * OZK_KSPLIT) ? OZK_KSPLIT : 1;
  if (parts > 1) {
    int rc = spx_aux_stream(ctx);
    if (rc) return rc;
    hipStream_t aux = ctx->aux;
    SPX_HIP(ctx, hipEventRecord(ctx->aux_ev[0], st));  // scales done
    SPX_HIP(ctx, hipStreamWaitEvent(aux, ctx->aux_ev[0], 0));
    for (int i = 0; i < parts; i++) {
      const long c0 = nchunk * i / parts, c1 = nchunk * (i + 1) / parts;
      slice(aux, 0, nw, c0, c1);
      SPX_LAUNCHED(ctx);
      SPX_HIP(ctx, hipEventRecord(ctx->aux_ev[1 + i], aux));
    }
    for (int i = 0; i < parts; i++) {
      const long c0 = nchunk * i / parts, c1 = nchunk * (i + 1) / parts;
      SPX_HIP(ctx, hipStreamWaitEvent(st, ctx->aux_ev[1 + i], 0));
      if (i == 0 && timed) SPX_HIP(ctx, hipEventRecord(ctx->phase_ev[1], st));
      herk(0, nw, c0, c1, i == 0 ? beta : 1.0);
      SPX_LAUNCHED(ctx);
    }
  } else {
    for (int w0 = 0; w0 < nw; w0 += nwg) {
      const int g = nw - w0 < nwg ? nw - w0 : nwg;
      slice(st, w0, g, 0, nchunk);
      SPX_LAUNCHED(ctx);
      if (w0 == 0 && timed) SPX_HIP(ctx, hipEventRecord(ctx->phase_ev[1], st));
      herk(w0, g, 0, nchunk, beta);
      SPX_LAUNCHED(ctx);
    }
  }
  if (timed) {
    SPX_HIP(ctx, hipEventRecord(ctx->phase_ev[2], st));
    ctx->phase_groups = (nw + nwg - 1) / nwg;
  }
  return SPEXHIP_OK;
}

// Synchronous form used by spexhip_chi0_dev: context scratch, then a check
// of the sign flag.
int spx_chi0_ozaki(spexhip_ctx *ctx, hipStream_t st, long n, int nb, const void *M, long ldm,
                   int colmajor, int nw, const void *wght, double alpha, double beta, void *chi,
                   int ldchi, size_t ws_limit) {
  SPX_CHECK(ctx, nw >= 1 && nw <= 16, "chi0 ozaki: 1..16 frequencies per call (got %d)", nw);
  int nwg = nw;
  while (nwg > 1 && spx_ozaki_ws_bytes(n, nb, nw, nwg) + 1024 > ws_limit) nwg--;
  const size_t wsb = spx_ozaki_ws_bytes(n, nb, nw, nwg) + 1024;
  void *ws;
  int rc = spx_scratch(ctx, wsb, &ws);
  if (rc) return rc;
  int *err = (int *)((char *)ws + wsb - 1024);
  SPX_HIP(ctx, hipMemsetAsync(err, 0, sizeof(int), st));
  rc = spx_chi0_ozaki_async(ctx, st, n, nb, M, ldm, colmajor, nw, wght, alpha, beta, chi, ldchi,
                            ws, wsb - 1024, err);
  if (rc) return rc;
  int h = 0;
  SPX_HIP(ctx, hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, st));
  SPX_HIP(ctx, hipStreamSynchronize(st));
  if (h)
    return spx_seterr(ctx, SPEXHIP_E_INVALID,
                      "chi0 ozaki: the weights of some frequency have both signs");
  return SPEXHIP_OK;
}
