/*
 * CPU oracle for the SPEX polarisation-function path (TEST INFRASTRUCTURE).
 *
 * Plain-C FP64 restatement of the reference arithmetic.  Only the test
 * suite, the smoke check and the CPU-baseline leg of bench.py may load it,
 * and only as the checker; the product library never links or calls it.
 *
 * Conventions shared with the product ABI:
 *   complex numbers are interleaved double[2] (Fortran complex(c_double));
 *   matrices are column-major like the Fortran reference;
 *   transition index t = b2 + nunocc*(b1 + nocc*k)  (b2 fastest,
 *   wavefproducts.f:120,490 and susceptibility.f:956-973);
 *   packed Hermitian storage is the upper triangle, ij = j*(j+1)/2 + i
 *   (0-based form of susceptibility.f:1210, ij=j(j-1)/2+i).
 */
#ifndef SPEX_ORACLE_H
#define SPEX_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* susceptibility.f:954-973 (HILBERT NONE). wght is (n x nw), t fastest. */
void or_weights_none(int nk, int nocc, int nunocc,
                     const double *e_occ,   /* [nk][nocc]   energies at k    */
                     const double *e_unocc, /* [nk][nunocc] energies at k+q  */
                     int nkpt, int nsym1, const int *nkpts, /* nkpts[k] */
                     int nw, const double *omega, double *wght);

/* susceptibility.f:1183-1240 with spectral_function HILBERT NONE :1780-1788.
 * Literal pair loop over packed pairs [pair_lo, pair_hi) in the reference
 * order (j outer, i<=j inner; the running pair counter equals ij).
 * chi is ACCUMULATED (+=) in packed-upper layout (npack x nw).            */
void or_chi0_pairsum(long n, int nb, const double *cprod, long ldc,
                     int nw, const double *wght, double *chi,
                     long pair_lo, long pair_hi, int nthreads);

/* Same quantity in blocked ZHERK form (second CPU baseline).
 * chi[ij + npack*w] += sum_t wght[t,w] cprod[t,i] conj(cprod[t,j]).       */
void or_chi0_herk(long n, int nb, const double *cprod, long ldc,
                  int nw, const double *wght, double *chi, int nthreads);

/* susceptibility.f:1331-1343: chi *= 2 for nspin=1. */
void or_scale(long count_complex, double *z, double s);

/* correlation.f:614-620 with p_diagmatdiag_dzpd1 (wrapper.f:1264-1278) and
 * p_plusdiag1_zp (wrapper.f:1317-1328): mat <- 1 - diag(sq) (mat) diag(sq). */
void or_dielectric_diag(int nb, const double *sqcoul, double *mat_packed);

/* Dense form: eps = 1 - Vh chi Vh with Vh a full Hermitian nb x nb matrix. */
void or_dielectric_dense(int nb, const double *vhalf, const double *chi_packed,
                         double *eps_packed);

/* correlation.f:709 Minverse -> inverse_zp1 (wrapper.f:2362-2374): in-place
 * inverse of an HPD packed-upper matrix.  Restated with a Cholesky
 * factorisation (what the ScaLAPACK path PZPOTRF/PZPOTRI does,
 * w_scalapack.inc:752-761).  Returns 0, or k>0 if the leading k x k minor
 * is not positive definite.                                                 */
int or_inverse_hpd_packed(int nb, double *mat_packed);

/* numerics.f:666-700 (Racah) and :709-727. */
double or_wigner3j(int l1, int l2, int l3, int m1, int m2, int m3);
double or_gaunt(int l1, int l2, int l3, int m1, int m2, int m3);

/* One atom type of the muffin-tin basis, wavefproducts.f:177-222. */
typedef struct {
  int lcut;           /* lcutp(itype) */
  const int *nindx;   /* nindx(0:lcut) */
  int lcutm;          /* lcutm(itype) */
  const int *nindxm;  /* nindxm(0:lcutm) */
  int neq;            /* atoms of this type */
  const double *integral; /* integral(d1,d2,d3) column-major, real */
  int d1, d2, d3;     /* its dimensions */
} or_mt_type;

/* wavefproducts.f:95-431 (wavefproducts2_mt, SUBTYPE 2, no SOC, noinv).
 * cmt(maxlmindx, ncent, nband, nkpt) column-major complex.  cprod is
 * (nunocc*nocc*nk) x nbasp complex column-major (t fastest), overwritten. */
void or_wfprod_mt(int ntype, const or_mt_type *types, const double *cent /*[ncent][3]*/,
                  const double *q /*[3] kpt(:,ik0)*/, int nk, const int *k1, const int *k2,
                  const double *cmt, int maxlmindx, int ncent, int nband,
                  int b1lo, int nocc, int b2lo, int nunocc,
                  double *cprod, long ldc);

/* wavefproducts.f:766-997 explicit convolution (SUBTYPE 2, noinv).
 *   G lists:  gpt[k][maxgpt][3], ngpt[k];  coefficients cpw(maxgpt,nband,nkpt)
 *   mixed basis IPWs gptm[ngptm][3];  umklapp g2[k][3] (gpt2).
 *   fullpw!=0 uses the step-function coefficients cstep on the box
 *   [-sb..sb]^3 (index ((z+sb)*(2sb+1)+(y+sb))*(2sb+1)+(x+sb)); otherwise
 *   G1 must lie in the G list of k1 (the gcut sphere).
 *   Output cprod(t, ig) for ig < ngptm, column-major, overwritten.         */
void or_wfprod_pw(int nk, const int *k1, const int *k2, const int *g2,
                  const int *gpt, const int *ngpt, int maxgpt,
                  const double *cpw, int nband,
                  int ngptm, const int *gptm, double vol,
                  int fullpw, const double *cstep, int sb,
                  int b1lo, int nocc, int b2lo, int nunocc,
                  double *cprod, long ldc);

#ifdef __cplusplus
}
#endif
#endif
