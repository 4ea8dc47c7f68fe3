// FFT grid setup of the interstitial products (host side): the basis change
// imat of determine_imat (fft.f:163-238), the grid of fft_init /
// gcount_sphere (fft.f:36-51, 101-158) and the step function theta(r) on
// that grid (wavefproducts.f:576-590).  Integer and small FP64 work done
// once per run; the device plans take their results (spexhip_pw_desc).
//
// Conventions: rlat[i + 3*j] is component i of reciprocal basis vector j
// (the reference's rlat(i,j), column-major); imat and imati are row-major,
// imat[3*i + j] = imat(i,j), like spexhip_pw_desc.imati.
#include <cmath>
#include <vector>

#include "../../include/spexhip.h"

namespace {

// m(i,j) of a row-major integer matrix
inline int at(const int *m, int i, int j) { return m[3 * i + j]; }

// rlat1 = rlat * imat (column-major result, k summed in increasing order)
void lat_times(const double *rlat, const int *imat, double *rlat1) {
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += rlat[i + 3 * k] * (double)at(imat, k, j);
      rlat1[i + 3 * j] = s;
    }
}

double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Non-orthogonality of a basis, sum over the three pairs of cos^2 of their
// angle (fft.f:177-179, 197-199).
double nonorth(const double *r) {
  const double l0 = dot3(r, r), l1 = dot3(r + 3, r + 3), l2 = dot3(r + 6, r + 6);
  const double d01 = dot3(r, r + 3), d12 = dot3(r + 3, r + 6), d20 = dot3(r + 6, r);
  return d01 * d01 / (l0 * l1) + d12 * d12 / (l1 * l2) + d20 * d20 / (l2 * l0);
}

// Smallest 2^a 3^b 5^c 7^d >= n within the reference's exponent ranges
// (fft.f:136-146); 0 if there is none.
long smooth_size(long n) {
  long best = 0;
  long p2 = 1;
  for (int i2 = 0; i2 <= 16; i2++, p2 *= 2) {
    long p3 = p2;
    for (int i3 = 0; i3 <= 11; i3++, p3 *= 3) {
      long p5 = p3;
      for (int i5 = 0; i5 <= 7; i5++, p5 *= 5) {
        long p7 = p5;
        for (int i7 = 0; i7 <= 6; i7++, p7 *= 7)
          if (p7 >= n && (best == 0 || p7 < best)) best = p7;
      }
    }
  }
  return best;
}

// One pass of the unnormalised backward DFT along one axis of a complex
// grid (x fastest): a(.., m, ..) <- sum_g a(.., g, ..) e^{+2 pi i g m / N}.
void dft_axis(std::vector<double> &a, const int *N, int axis) {
  const long n = N[axis];
  const long stride = axis == 0 ? 1 : axis == 1 ? N[0] : (long)N[0] * N[1];
  const long total = (long)N[0] * N[1] * N[2];
  std::vector<double> cs(2 * n), line(2 * n);
  for (long r = 0; r < n; r++) {
    const double ang = 2.0 * M_PI * (double)r / (double)n;
    cs[2 * r] = std::cos(ang);
    cs[2 * r + 1] = std::sin(ang);
  }
  for (long base = 0; base < total; base++) {
    if ((base / stride) % n != 0) continue;  // first element of each line
    for (long g = 0; g < n; g++) {
      line[2 * g] = a[2 * (base + g * stride)];
      line[2 * g + 1] = a[2 * (base + g * stride) + 1];
    }
    for (long m = 0; m < n; m++) {
      double sr = 0.0, si = 0.0;
      for (long g = 0; g < n; g++) {
        const long r = (g * m) % n;
        const double c = cs[2 * r], s = cs[2 * r + 1];
        sr += line[2 * g] * c - line[2 * g + 1] * s;
        si += line[2 * g] * s + line[2 * g + 1] * c;
      }
      a[2 * (base + m * stride)] = sr;
      a[2 * (base + m * stride) + 1] = si;
    }
  }
}

}  // namespace

extern "C" {

int spexhip_determine_imat(const double *rlat, int *imat, int *imati) {
  if (!rlat || !imat || !imati) return SPEXHIP_E_INVALID;
  const int b = 2;
  int best[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double rdum0 = nonorth(rlat);
  if (!std::isfinite(rdum0)) return SPEXHIP_E_INVALID;
  int c[9];  // row-major candidate: c[3*r + s] = i_{r+1, s+1}
  double r1[9];
  // the reference's loop nest i11, i12, i13, i21, ..., i33 (outermost first)
  for (c[0] = -b; c[0] <= b; c[0]++)
    for (c[1] = -b; c[1] <= b; c[1]++)
      for (c[2] = -b; c[2] <= b; c[2]++)
        for (c[3] = -b; c[3] <= b; c[3]++)
          for (c[4] = -b; c[4] <= b; c[4]++)
            for (c[5] = -b; c[5] <= b; c[5]++)
              for (c[6] = -b; c[6] <= b; c[6]++)
                for (c[7] = -b; c[7] <= b; c[7]++)
                  for (c[8] = -b; c[8] <= b; c[8]++) {
                    const int det = c[0] * c[4] * c[8] + c[1] * c[5] * c[6] + c[2] * c[3] * c[7] -
                                    c[6] * c[4] * c[2] - c[7] * c[5] * c[0] - c[8] * c[3] * c[1];
                    if (det != 1) continue;
                    lat_times(rlat, c, r1);
                    const double rdum = nonorth(r1);
                    if (rdum < rdum0 - 1e-12) {
                      rdum0 = rdum;
                      for (int q = 0; q < 9; q++) best[q] = c[q];
                    }
                  }
  for (int q = 0; q < 9; q++) imat[q] = best[q];
  // imati = adjugate (det = 1): imati(i,j) = (-1)^(i+j) * minor(j,i)
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const int r0 = j == 0 ? 1 : 0, r1_ = j == 2 ? 1 : 2;  // rows left after deleting j
      const int c0 = i == 0 ? 1 : 0, c1 = i == 2 ? 1 : 2;   // columns left after deleting i
      const int minor = at(best, r0, c0) * at(best, r1_, c1) - at(best, r0, c1) * at(best, r1_, c0);
      imati[3 * i + j] = ((i + j) % 2 ? -1 : 1) * minor;
    }
  for (int i = 0; i < 3; i++)  // imat * imati = 1, as the reference checks (fft.f:217)
    for (int j = 0; j < 3; j++) {
      int s = 0;
      for (int k = 0; k < 3; k++) s += at(imat, i, k) * at(imati, k, j);
      if (s != (i == j)) return SPEXHIP_E_INVALID;
    }
  return SPEXHIP_OK;
}

int spexhip_fft_grid(const double *rlat, const int *imat, double gcutf, int *nfft, long *npoints) {
  if (!rlat || !imat || !nfft || !(gcutf >= 0.0) || !std::isfinite(gcutf))
    return SPEXHIP_E_INVALID;
  double r1[9];
  lat_times(rlat, imat, r1);
  const double g2 = gcutf * gcutf;
  long count = 0;
  long mx[3] = {-1000, -1000, -1000};
  // shells of constant |ix| + |iy| + |iz| = n until one holds no point
  for (long n = 0;; n++) {
    bool found = false;
    for (long ix = -n; ix <= n; ix++) {
      const long ry = n - std::labs(ix);
      for (long iy = -ry; iy <= ry; iy++) {
        const long iz0 = n - std::labs(ix) - std::labs(iy);
        for (int sgn = 0; sgn < (iz0 > 0 ? 2 : 1); sgn++) {
          const long iz = sgn ? -iz0 : iz0;
          double v[3], r = 0.0;
          for (int i = 0; i < 3; i++) {
            v[i] = (double)ix * r1[i] + (double)iy * r1[i + 3] + (double)iz * r1[i + 6];
            r += v[i] * v[i];
          }
          if (r <= g2) {
            count++;
            found = true;
            if (ix > mx[0]) mx[0] = ix;
            if (iy > mx[1]) mx[1] = iy;
            if (iz > mx[2]) mx[2] = iz;
          }
        }
      }
    }
    if (!found) break;
    if (n > (1 << 16)) return SPEXHIP_E_INVALID;
  }
  for (int i = 0; i < 3; i++) {
    const long m = 2 * mx[i] + 1;
    if (m > (1L << 16)) return SPEXHIP_E_INVALID;
    const long s = smooth_size(m);
    if (s < m || s > (1L << 30)) return SPEXHIP_E_INVALID;
    nfft[i] = (int)s;
  }
  if (npoints) *npoints = count;
  return SPEXHIP_OK;
}

int spexhip_step_grid(const int *nfft, const int *imat, const double *rlat, double gmax,
                      const double *cstep, const int *cext, double *theta) {
  if (!nfft || !imat || !rlat || !cstep || !cext || !theta || !(gmax >= 0.0))
    return SPEXHIP_E_INVALID;
  for (int i = 0; i < 3; i++)
    if (nfft[i] < 1 || nfft[i] > 4096 || cext[i] < 0) return SPEXHIP_E_INVALID;
  const int N[3] = {nfft[0], nfft[1], nfft[2]};
  const long NN = (long)N[0] * N[1] * N[2];
  const long w0 = 2L * cext[0] + 1, w1 = 2L * cext[1] + 1;
  std::vector<double> a(2 * NN, 0.0);
  const double gm2 = gmax * gmax;
  for (int k = 0; k < N[2]; k++)
    for (int j = 0; j < N[1]; j++)
      for (int i = 0; i < N[0]; i++) {
        // grid frequency, folded into (-N/2, N/2] (wavefproducts.f:579-581)
        int g[3] = {i, j, k};
        for (int d = 0; d < 3; d++)
          if (g[d] > N[d] / 2) g[d] -= N[d];
        int g1[3];
        for (int r = 0; r < 3; r++) g1[r] = at(imat, r, 0) * g[0] + at(imat, r, 1) * g[1] + at(imat, r, 2) * g[2];
        double v2 = 0.0;
        for (int r = 0; r < 3; r++) {
          const double v = rlat[r] * g1[0] + rlat[r + 3] * g1[1] + rlat[r + 6] * g1[2];
          v2 += v * v;
        }
        if (v2 > gm2) continue;
        for (int d = 0; d < 3; d++)
          if (g1[d] < -cext[d] || g1[d] > cext[d]) return SPEXHIP_E_INVALID;
        const long src = ((long)(g1[2] + cext[2]) * w1 + (g1[1] + cext[1])) * w0 + (g1[0] + cext[0]);
        const long dst = i + (long)N[0] * (j + (long)N[1] * k);
        a[2 * dst] = cstep[2 * src];
        a[2 * dst + 1] = cstep[2 * src + 1];
      }
  for (int axis = 0; axis < 3; axis++) dft_axis(a, N, axis);  // FFTW BACKWARD, unnormalised
  for (long q = 0; q < 2 * NN; q++) theta[q] = a[q];
  return SPEXHIP_OK;
}

}  // extern "C"
