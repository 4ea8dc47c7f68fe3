// Per-column, per-frequency scales of the balanced chi0 operands
//     B(t,I) = sqrt|D(t,w)| * M(t,I),
//     scale[w][I] = max_t sqrt|D(t,w)| * max(|Re M(t,I)|, |Im M(t,I)|),
// used by the compensated forms of the pair sum (split-FP32 and Ozaki int8),
// which both write B on a grid relative to its column maximum.  The weights
// come either as D itself (the square root is taken here, with the same
// expression the split-FP32 staging uses, so every staged |B| is <= its
// scale bit for bit) or already as sqrt|D| (the Ozaki workspace).  The
// kernels live in chi0.hip.
#pragma once
#include "common.hpp"

// Zeroes scale[nw][nb] and fills it (fill kernels only: capture-safe).
// from_d: W is D (rows of n transitions per frequency); else W is sqrt|D|
// in the same layout.
int spx_col_scales(spexhip_ctx *ctx, hipStream_t st, long n, int nb, const void *M, long ldm,
                   int colmajor, int nw, const double *W, int from_d, double *scale);
