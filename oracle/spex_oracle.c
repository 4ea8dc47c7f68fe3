/*
 * CPU oracle for the SPEX polarisation-function path (TEST INFRASTRUCTURE).
 * FP64 restatement of the reference arithmetic; each function cites the
 * reference lines it follows.  Used only by tests, smoke() and the CPU
 * baseline leg of bench.py -- never by the product library.
 */
#include "spex_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;
#define ZAT(p, i) ((zc *)(p))[i]

/* ------------------------------------------------------------------ */
/* susceptibility.f:954-973: wghtSr(i,ifrq) = -2*rdum/((frqh/img)**2 +   */
/* rdum**2)/nkpt/nsym1*nkpts(k), rdum = ene(i2,k+q) - ene(i1,k).        */
void or_weights_none(int nk, int nocc, int nunocc, const double *e_occ,
                     const double *e_unocc, int nkpt, int nsym1,
                     const int *nkpts, int nw, const double *omega,
                     double *wght) {
  long n = (long)nk * nocc * nunocc;
  long i = 0;
  for (int k = 0; k < nk; k++)
    for (int i1 = 0; i1 < nocc; i1++)
      for (int i2 = 0; i2 < nunocc; i2++, i++) {
        double rdum = e_unocc[(long)k * nunocc + i2] - e_occ[(long)k * nocc + i1];
        for (int f = 0; f < nw; f++) {
          double w = omega[f];
          wght[i + n * f] =
              -2 * rdum / (w * w + rdum * rdum) / nkpt / nsym1 * nkpts[k];
        }
      }
}

/* ------------------------------------------------------------------ */
/* susceptibility.f:1183-1240 (+ spectral_function :1780-1788):          */
/*   do j; do i<=j: re(t)+i*im(t) = cprod(t,i)*conjg(cprod(t,j))          */
/*   sus1(w)=ddot(re,wght(:,w)); sus2(w)=ddot(im,wght(:,w))                */
/*   susceph(ij,w) += sus1 + img*sus2,   ij = j(j-1)/2+i                   */
void or_chi0_pairsum(long n, int nb, const double *cprod, long ldc, int nw,
                     const double *wght, double *chi, long pair_lo,
                     long pair_hi, int nthreads) {
  long npack = (long)nb * (nb + 1) / 2;
  if (pair_lo < 0) pair_lo = 0;
  if (pair_hi < 0 || pair_hi > npack) pair_hi = npack;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    double *re = malloc(sizeof(double) * n);
    double *im = malloc(sizeof(double) * n);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (long ij = pair_lo; ij < pair_hi; ij++) {
      /* invert ij = j(j+1)/2 + i (0-based) */
      long j = (long)((sqrt(8.0 * (double)ij + 1.0) - 1.0) / 2.0);
      while ((j + 1) * (j + 2) / 2 <= ij) j++;
      while (j * (j + 1) / 2 > ij) j--;
      long i = ij - j * (j + 1) / 2;
      const zc *ci = (const zc *)cprod + i * ldc;
      const zc *cj = (const zc *)cprod + j * ldc;
      for (long t = 0; t < n; t++) {
        zc c = ci[t] * conj(cj[t]);
        re[t] = creal(c);
        im[t] = cimag(c);
      }
      for (int f = 0; f < nw; f++) {
        const double *w = wght + n * f;
        double s1 = 0, s2 = 0;
        for (long t = 0; t < n; t++) {
          s1 += re[t] * w[t];
          s2 += im[t] * w[t];
        }
        chi[2 * (ij + npack * f)] += s1;
        chi[2 * (ij + npack * f) + 1] += s2;
      }
    }
    free(re);
    free(im);
  }
}

/* Blocked Hermitian rank-k form of the same sum.  For each frequency the  */
/* weighted operand a(t,i) = w(t)*cprod(t,i) is formed per t-block and the */
/* upper triangle is accumulated column by column.                         */
void or_chi0_herk(long n, int nb, const double *cprod, long ldc, int nw,
                  const double *wght, double *chi, int nthreads) {
  long npack = (long)nb * (nb + 1) / 2;
  const long TB = 512;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
  for (int f = 0; f < nw; f++) {
    const double *w = wght + n * f;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (long j = 0; j < nb; j++) {
      zc *acc = calloc(j + 1, sizeof(zc));
      zc *cjw = malloc(sizeof(zc) * TB);
      const zc *cj = (const zc *)cprod + j * ldc;
      for (long t0 = 0; t0 < n; t0 += TB) {
        long tl = n - t0 < TB ? n - t0 : TB;
        for (long t = 0; t < tl; t++) cjw[t] = w[t0 + t] * conj(cj[t0 + t]);
        for (long i = 0; i <= j; i++) {
          const zc *ci = (const zc *)cprod + i * ldc + t0;
          zc s = 0;
          for (long t = 0; t < tl; t++) s += ci[t] * cjw[t];
          acc[i] += s;
        }
      }
      for (long i = 0; i <= j; i++) {
        long ij = j * (j + 1) / 2 + i;
        ZAT(chi, ij + npack * f) += acc[i];
      }
      free(acc);
      free(cjw);
    }
  }
}

void or_scale(long count, double *z, double s) {
  for (long i = 0; i < 2 * count; i++) z[i] *= s;
}

/* correlation.f:614-620:  suscep = -suscep;  p_diagmatdiag(sqcoul,suscep) */
/* (wrapper.f:1264-1278: mat(k)=diag(i)*diag(j)*mat(k)); p_plusdiag(1,.)   */
/* (wrapper.f:1317-1328: mat(j)=1+mat(j) at j=i(i+1)/2).                   */
void or_dielectric_diag(int nb, const double *sqcoul, double *mat) {
  long k = 0;
  for (int j = 0; j < nb; j++)
    for (int i = 0; i <= j; i++, k++) {
      zc v = -ZAT(mat, k);
      ZAT(mat, k) = sqcoul[i] * sqcoul[j] * v;
    }
  for (int i = 0; i < nb; i++) {
    long ii = (long)i * (i + 1) / 2 + i;
    ZAT(mat, ii) = 1.0 + ZAT(mat, ii);
  }
}

static void unpack_herm(int nb, const double *p, zc *full) {
  long k = 0;
  for (int j = 0; j < nb; j++)
    for (int i = 0; i <= j; i++, k++) {
      full[i + (long)nb * j] = ZAT(p, k);
      full[j + (long)nb * i] = conj(ZAT(p, k));
    }
}

void or_dielectric_dense(int nb, const double *vhalf, const double *chi_packed,
                         double *eps_packed) {
  zc *c = malloc(sizeof(zc) * nb * nb);
  zc *tmp = calloc((size_t)nb * nb, sizeof(zc));
  const zc *v = (const zc *)vhalf;
  unpack_herm(nb, chi_packed, c);
  /* tmp = chi * Vh */
  for (int j = 0; j < nb; j++)
    for (int l = 0; l < nb; l++) {
      zc b = v[l + (long)nb * j];
      const zc *cl = c + (long)nb * l;
      zc *tj = tmp + (long)nb * j;
      for (int i = 0; i < nb; i++) tj[i] += cl[i] * b;
    }
  /* eps(i,j) = delta_ij - sum_l Vh(i,l) tmp(l,j), upper triangle only */
  long k = 0;
  for (int j = 0; j < nb; j++)
    for (int i = 0; i <= j; i++, k++) {
      zc s = 0;
      for (int l = 0; l < nb; l++) s += v[i + (long)nb * l] * tmp[l + (long)nb * j];
      ZAT(eps_packed, k) = (i == j ? 1.0 : 0.0) - s;
    }
  free(c);
  free(tmp);
}

/* Inverse of an HPD packed-upper matrix via A = U^H U (upper Cholesky),   */
/* U^{-1} by back substitution, A^{-1} = U^{-1} U^{-H}.                     */
int or_inverse_hpd_packed(int nb, double *mat) {
  long N = nb;
  zc *a = malloc(sizeof(zc) * N * N);
  unpack_herm(nb, mat, a);
#define A(i, j) a[(i) + N * (j)]
  /* upper Cholesky, column by column (zpotrf 'U' unblocked order) */
  for (long j = 0; j < N; j++) {
    double d = creal(A(j, j));
    for (long k = 0; k < j; k++) d -= creal(conj(A(k, j)) * A(k, j));
    if (!(d > 0)) {
      free(a);
      return (int)(j + 1);
    }
    d = sqrt(d);
    A(j, j) = d;
    for (long i = j + 1; i < N; i++) {
      zc s = A(j, i);
      for (long k = 0; k < j; k++) s -= conj(A(k, j)) * A(k, i);
      A(j, i) = s / d;
    }
  }
  /* W = U^{-1} (upper), in place (ztrtri 'U','N' column order) */
  for (long j = 0; j < N; j++) {
    A(j, j) = 1.0 / A(j, j);
    zc ajj = -A(j, j);
    /* W(i,j) = -W(j,j) * sum_{k=i..j-1} W(i,k) U(k,j), i ascending so    */
    /* that U(k,j), k>i, is still in place when W(i,j) overwrites U(i,j). */
    for (long i = 0; i < j; i++) {
      zc s = 0;
      for (long k = i; k < j; k++) s += A(i, k) * A(k, j);
      A(i, j) = s;
    }
    for (long i = 0; i < j; i++) A(i, j) *= ajj;
  }
  /* inv = W W^H, upper triangle (zlauum 'U') */
  long p = 0;
  for (long j = 0; j < N; j++)
    for (long i = 0; i <= j; i++, p++) {
      zc s = 0;
      for (long k = j; k < N; k++) s += A(i, k) * conj(A(j, k));
      ZAT(mat, p) = s;
    }
#undef A
  free(a);
  return 0;
}

/* ------------------------------------------------------------------ */
/* getinput.f:1720-1736: fac(i)=i!, sfac(i)=sqrt(i!)                     */
static double fac_(int i) {
  double f = 1;
  for (int k = 2; k <= i; k++) f *= k;
  return f;
}
static double sfac_(int i) {
  double f = 1;
  for (int k = 1; k <= i; k++) f *= sqrt((double)k);
  return f;
}

/* numerics.f:666-700 */
double or_wigner3j(int l1, int l2, int l3, int m1, int m2, int m3) {
  int f1 = l3 - l2 + m1, f2 = l3 - l1 - m2, f3 = l1 + l2 - l3, f4 = l1 - m1,
      f5 = l2 + m2;
  int tmin = 0, tmax;
  if (-f1 > tmin) tmin = -f1;
  if (-f2 > tmin) tmin = -f2;
  tmax = f3;
  if (f4 < tmax) tmax = f4;
  if (f5 < tmax) tmax = f5;
  double w = 0;
  if (tmin <= tmax) {
    for (int t = tmin; t <= tmax; t++)
      w += ((t % 2) ? -1.0 : 1.0) /
           (fac_(t) * fac_(f1 + t) * fac_(f2 + t) * fac_(f3 - t) * fac_(f4 - t) *
            fac_(f5 - t));
    int e = l1 - l2 - m3;
    w *= ((e % 2 + 2) % 2 ? -1.0 : 1.0) * sfac_(l1 + l2 - l3) * sfac_(l1 - l2 + l3) *
         sfac_(-l1 + l2 + l3) / sfac_(l1 + l2 + l3 + 1) * sfac_(l1 + m1) *
         sfac_(l1 - m1) * sfac_(l2 + m2) * sfac_(l2 - m2) * sfac_(l3 + m3) *
         sfac_(l3 - m3);
  }
  return w;
}

/* numerics.f:709-727: integral of conj(Y_l1m1) Y_l2m2 conj(Y_l3m3) */
double or_gaunt(int l1, int l2, int l3, int m1, int m2, int m3) {
  if (m3 != m2 - m1) return 0;
  if (abs(m1) > l1 || abs(m2) > l2 || abs(m3) > l3) return 0;
  if ((l1 + l2 + l3) % 2) return 0;
  if (l3 < abs(l1 - l2) || l3 > l1 + l2) return 0;
  int e = m1 + m3;
  double sgn = ((e % 2 + 2) % 2) ? -1.0 : 1.0;
  return sgn * sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / M_PI / 4) *
         or_wigner3j(l1, l2, l3, -m1, m2, -m3) * or_wigner3j(l1, l2, l3, 0, 0, 0);
}

/* ------------------------------------------------------------------ */
/* wavefproducts.f:95-431 (SUBTYPE 2, noinv, no SOC).                    */
void or_wfprod_mt(int ntype, const or_mt_type *types, const double *cent,
                  const double *q, int nk, const int *k1, const int *k2,
                  const double *cmt, int maxlmindx, int ncent, int nband,
                  int b1lo, int nocc, int b2lo, int nunocc, double *cprod,
                  long ldc) {
  long n = (long)nk * nocc * nunocc;
  int nbasp = 0;
  for (int it = 0; it < ntype; it++) {
    int maxllm = 0;
    for (int ll = 0; ll <= types[it].lcutm; ll++) maxllm += (2 * ll + 1) * types[it].nindxm[ll];
    nbasp += types[it].neq * maxllm;
  }
  for (int col = 0; col < nbasp; col++)
    for (long t = 0; t < n; t++) ZAT(cprod, t + ldc * col) = 0;
#define CMT(lm, ic, b, k) \
  ZAT(cmt, (lm) + (long)maxlmindx * ((ic) + (long)ncent * ((b) + (long)nband * (k))))
  int llm0 = 0, ic = -1;
  for (int it = 0; it < ntype; it++) {
    const or_mt_type *T = &types[it];
    int lcut = T->lcut, lcutm = T->lcutm;
    int maxlm = 0, maxllm = 0;
    for (int l = 0; l <= lcut; l++) maxlm += (2 * l + 1) * T->nindx[l];
    for (int ll = 0; ll <= lcutm; ll++) maxllm += (2 * ll + 1) * T->nindxm[ll];
    /* lmstart(m,l) (:225-232) */
    int lmstart[64][129];
    {
      int lm = 0;
      for (int l = 0; l <= lcut; l++)
        for (int m = -l; m <= l; m++) {
          lmstart[l][m + 64] = lm;
          lm += T->nindx[l];
        }
    }
#define INTEG(a, b, c) T->integral[(a) + (long)T->d1 * ((b) + (long)T->d2 * (c))]
    zc *mat = malloc(sizeof(zc) * maxlm * maxllm);
    zc *cmt1 = malloc(sizeof(zc) * maxlm);
    for (int ieq = 0; ieq < T->neq; ieq++) {
      ic++;
      /* :252 cexp = exp(-img*2*pi*dot_product(kpt(:,ik0),cent(:,ic))) */
      double ph = q[0] * cent[3 * ic] + q[1] * cent[3 * ic + 1] + q[2] * cent[3 * ic + 2];
      zc cph = cexp(-I * 2 * M_PI * ph);
      for (int ik = 0; ik < nk; ik++) {
        for (int ib = 0; ib < nocc; ib++) {
          for (int lm = 0; lm < maxlm; lm++) cmt1[lm] = CMT(lm, ic, b1lo + ib, k1[ik]);
          for (long z = 0; z < (long)maxlm * maxllm; z++) mat[z] = 0;
          /* :294-339 Gaunt-weighted contraction */
          int llm = 0, lln = 0;
          for (int ll = 0; ll <= lcutm; ll++) {
            for (int mm = -ll; mm <= ll; mm++) {
              int lm = 0, ln = 0;
              for (int l = 0; l <= lcut; l++) {
                int nx = T->nindx[l];
                for (int m = -l; m <= l; m++) {
                  int m1 = m + mm;
                  int ln1 = 0;
                  int l1lo = abs(l - ll), l1hi = l + ll < lcut ? l + ll : lcut;
                  for (int l1 = l1lo; l1 <= l1hi; l1 += 2) {
                    int nx1 = T->nindx[l1];
                    if (l1 >= abs(m1)) {
                      /* gaunt1(l1,l,ll,m,mm) = gaunt(ll,l1,l,mm,m1,m) (:163) */
                      double gnt = or_gaunt(ll, l1, l, mm, m1, m);
                      if (gnt != 0) {
                        int lm1 = lmstart[l1][m1 + 64];
                        for (int nn = 0; nn < T->nindxm[ll]; nn++)
                          for (int n1 = 0; n1 < nx1; n1++) {
                            zc s = 0;
                            for (int nq = 0; nq < nx; nq++)
                              s += cmt1[lm + nq] * INTEG(ln + nq, ln1 + n1, lln + nn);
                            mat[(lm1 + n1) + (long)maxlm * (llm + nn)] += s * gnt;
                          }
                      }
                    }
                    ln1 += nx1;
                  }
                  lm += nx;
                }
                ln += nx;
              }
              llm += T->nindxm[ll];
            }
            lln += T->nindxm[ll];
          }
          /* :359 cprod(:,ib,ik,llm0+1:llm) += cexp*conjg(macmat(cmt2,mat)) */
          for (int b2 = 0; b2 < nunocc; b2++) {
            long t = b2 + (long)nunocc * (ib + (long)nocc * ik);
            for (int I2 = 0; I2 < maxllm; I2++) {
              zc s = 0;
              for (int lm = 0; lm < maxlm; lm++)
                s += conj(CMT(lm, ic, b2lo + b2, k2[ik])) * mat[lm + (long)maxlm * I2];
              ZAT(cprod, t + ldc * (llm0 + I2)) += cph * conj(s);
            }
          }
        }
      }
      llm0 += maxllm;
    }
    free(mat);
    free(cmt1);
#undef INTEG
  }
#undef CMT
}

/* ------------------------------------------------------------------ */
/* wavefproducts.f:766-997 (explicit convolution, SUBTYPE 2, noinv).     */
/* point(g) over the box of g = G2 - Gm + g2 (:795-841):                  */
/*  non-fullpw: point(g) = index of g in the G list of k1 (pgptinv), and  */
/*    cprod(b2,b1,k,ig) += dotprod(cpw3,cpw2)*sqrtvoli, cpw3 = c1(point)  */
/*    (:953-972);                                                          */
/*  fullpw: cpw0(G0) = sum_G1 cstep(G0-G1) c1(G1) (:870-888) and          */
/*    cprod += sqrtvoli*conjg(macmat(cpw2,cpw3)), cpw3 = cpw0(point)       */
/*    (:917-945).                                                          */
void or_wfprod_pw(int nk, const int *k1, const int *k2, const int *g2,
                  const int *gpt, const int *ngpt, int maxgpt,
                  const double *cpw, int nband, int ngptm, const int *gptm,
                  double vol, int fullpw, const double *cstep, int sb,
                  int b1lo, int nocc, int b2lo, int nunocc, double *cprod,
                  long ldc) {
  double sqrtvoli = 1 / sqrt(vol);
  long n = (long)nk * nocc * nunocc;
  for (int ig = 0; ig < ngptm; ig++)
    for (long t = 0; t < n; t++) ZAT(cprod, t + ldc * ig) = 0;
#define CPW(g, b, k) ZAT(cpw, (g) + (long)maxgpt * ((b) + (long)nband * (k)))
  int sw = 2 * sb + 1;
  for (int ik = 0; ik < nk; ik++) {
    int kk1 = k1[ik], kk2 = k2[ik];
    const int *G1 = gpt + 3L * maxgpt * kk1, *G2 = gpt + 3L * maxgpt * kk2;
    int n1 = ngpt[kk1], n2 = ngpt[kk2];
    /* box extents of g = G2 - Gm + g2 (def == 0 pass, :803-808) */
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) { lo[d] = 1 << 30; hi[d] = -(1 << 30); }
    for (int ig2 = 0; ig2 < n2; ig2++)
      for (int ig = 0; ig < ngptm; ig++)
        for (int d = 0; d < 3; d++) {
          int g = G2[3 * ig2 + d] - gptm[3 * ig + d] + g2[3 * ik + d];
          if (g < lo[d]) lo[d] = g;
          if (g > hi[d]) hi[d] = g;
        }
    long ex[3];
    for (int d = 0; d < 3; d++) ex[d] = hi[d] - lo[d] + 1;
    if (n2 == 0 || ngptm == 0) continue;
    int *point = calloc(ex[0] * ex[1] * ex[2], sizeof(int));
#define PIDX(g0, g1, gg2) (((long)((gg2) - lo[2]) * ex[1] + ((g1) - lo[1])) * ex[0] + ((g0) - lo[0]))
    if (!fullpw) {
      for (int ig1 = 0; ig1 < n1; ig1++) {
        int a = G1[3 * ig1], b = G1[3 * ig1 + 1], c = G1[3 * ig1 + 2];
        if (a < lo[0] || a > hi[0] || b < lo[1] || b > hi[1] || c < lo[2] || c > hi[2]) continue;
        point[PIDX(a, b, c)] = ig1 + 1;
      }
    }
    /* fullpw: every g in the box is a G0 of cpw0 */
    long nbox = ex[0] * ex[1] * ex[2];
    zc *cpw0 = fullpw ? malloc(sizeof(zc) * nbox) : NULL;
    for (int ib = 0; ib < nocc; ib++) {
      if (fullpw) {
        for (long p = 0; p < nbox; p++) {
          int g0 = (int)(p % ex[0]) + lo[0], gq1 = (int)((p / ex[0]) % ex[1]) + lo[1],
              gq2 = (int)(p / (ex[0] * ex[1])) + lo[2];
          zc s = 0;
          for (int ig1 = 0; ig1 < n1; ig1++) {
            int d0 = g0 - G1[3 * ig1], d1 = gq1 - G1[3 * ig1 + 1], d2 = gq2 - G1[3 * ig1 + 2];
            if (abs(d0) > sb || abs(d1) > sb || abs(d2) > sb) continue;
            s += ZAT(cstep, ((long)(d2 + sb) * sw + (d1 + sb)) * sw + (d0 + sb)) *
                 CPW(ig1, b1lo + ib, kk1);
          }
          cpw0[p] = s;
        }
      }
      for (int ig = 0; ig < ngptm; ig++) {
        for (int b2 = 0; b2 < nunocc; b2++) {
          long t = b2 + (long)nunocc * (ib + (long)nocc * ik);
          zc acc = 0;
          for (int ig2 = 0; ig2 < n2; ig2++) {
            int g[3];
            for (int d = 0; d < 3; d++)
              g[d] = G2[3 * ig2 + d] - gptm[3 * ig + d] + g2[3 * ik + d];
            long p = PIDX(g[0], g[1], g[2]);
            zc c1;
            if (!fullpw) {
              if (!point[p]) continue;
              c1 = CPW(point[p] - 1, b1lo + ib, kk1);
            } else {
              c1 = cpw0[p];
            }
            acc += conj(c1) * CPW(ig2, b2lo + b2, kk2);
          }
          ZAT(cprod, t + ldc * ig) += acc * sqrtvoli;
        }
      }
    }
    free(point);
    free(cpw0);
#undef PIDX
  }
#undef CPW
}
