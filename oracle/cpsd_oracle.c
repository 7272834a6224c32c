/*
 * cpsd_oracle.c -- test infrastructure only (see oracle.h).
 *
 * Serial FP64 restatement of the reference SD step ccpsdv3/cpsd.c:33-698
 * (ispin = 1, fixed or moving ions) and the helpers it calls.  Sums are sequential loops; the reference's BLAS
 * calls (ddot/daxpy/dcopy/dgemm/idamax) are written out as plain loops.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ---- BLAS-equivalent loops ------------------------------------------- */

static double dot_s(int n, const double *x, int incx, const double *y, int incy)
{
   double s = 0.0;
   int i;
   for (i = 0; i < n; ++i) s += x[(size_t)i*incx] * y[(size_t)i*incy];
   return s;
}

static void axpy(size_t n, double a, const double *x, double *y)
{
   size_t i;
   for (i = 0; i < n; ++i) y[i] += a * x[i];
}

/* C(m x n) = beta*C + alpha*A(m x k)*B(k x n), column-major, lda = m etc. */
static void gemm_nn(int m, int n, int k, double alpha, const double *A,
                    size_t lda, const double *B, int ldb, double beta,
                    double *C, size_t ldc)
{
   int i, j, l;
   for (j = 0; j < n; ++j) {
      double *c = C + (size_t)j * ldc;
      if (beta == 0.0) for (i = 0; i < m; ++i) c[i] = 0.0;
      else if (beta != 1.0) for (i = 0; i < m; ++i) c[i] *= beta;
      for (l = 0; l < k; ++l) {
         double b = alpha * B[l + (size_t)j * ldb];
         const double *a = A + (size_t)l * lda;
         for (i = 0; i < m; ++i) c[i] += b * a[i];
      }
   }
}

/* ---- G-space reductions ---------------------------------------------- */

/* ccpsdv3/gcdotc.c:16-29: 2*sum over the whole half-complex array minus the
 * kx = 0 plane counted once (real and imaginary parts). */
double oracle_gcdotc(int nfft, const double *a, const double *b)
{
   int n2 = nfft * nfft, n3 = (nfft + 2) * nfft * nfft;
   return 2.0 * dot_s(n3, a, 1, b, 1)
        - dot_s(n2, a, nfft + 2, b, nfft + 2)
        - dot_s(n2, a + 1, nfft + 2, b + 1, nfft + 2);
}

/* ccpsdv3/gsdot.c:16-28 (real arrays of nfft3d entries) */
double oracle_gsdot(int nfft, const double *a, const double *b)
{
   int nh = nfft / 2 + 1;
   return 2.0 * dot_s(nh * nfft * nfft, a, 1, b, 1)
        - dot_s(nfft * nfft, a, nh, b, nh);
}

/* ccpsdv3/gssum.c:30-41 */
double oracle_gssum(int nfft, const double *a)
{
   int nh = nfft / 2 + 1, i, n3 = nh * nfft * nfft;
   double s1 = 0.0, s2 = 0.0;
   for (i = 0; i < n3; ++i) s1 += a[i];
   for (i = 0; i < nfft * nfft; ++i) s2 += a[(size_t)i * nh];
   return 2.0 * s1 - s2;
}

/* ---- structure factors (ccpsdv3/strfac.c) ---------------------------- */

/* Phase factors by repeated complex multiplication, ccpsdv3/strfac.c:10-77:
 * ex1[k] = e^{-i(b1.R+pi)k} (k = 0..N/2), ex2/ex3 over k = 0..N-1 with
 * ex[N-k] = conj(ex[k]); the Nyquist entry (k = N/2) is zeroed. */
void oracle_phafac(int nfft, const double *unitg, int nion, const double *r1,
                   double *ex1, double *ex2, double *ex3)
{
   const double pi = 4.0 * atan(1.0);
   int ii, k, nh = nfft / 2, d;
   for (ii = 0; ii < nion; ++ii) {
      double *e[3];
      e[0] = ex1 + (size_t)ii * (nfft + 2);
      e[1] = ex2 + (size_t)ii * 2 * nfft;
      e[2] = ex3 + (size_t)ii * 2 * nfft;
      for (d = 0; d < 3; ++d) {
         double sw = unitg[3*d] * r1[3*ii] + unitg[3*d+1] * r1[3*ii+1]
                   + unitg[3*d+2] * r1[3*ii+2] + pi;
         double cr = cos(sw), ci = -sin(sw);
         double *x = e[d];
         x[0] = 1.0; x[1] = 0.0;
         for (k = 1; k <= nh; ++k) {
            double pr = x[2*(k-1)], pim = x[2*(k-1)+1];
            x[2*k]   = pr * cr - pim * ci;
            x[2*k+1] = pr * ci + pim * cr;
            if (d > 0) {
               x[2*(nfft-k)]   =  x[2*k];
               x[2*(nfft-k)+1] = -x[2*k+1];
            }
         }
         x[nfft] = 0.0; x[nfft+1] = 0.0;
      }
   }
}

/* exi(kx,ky,kz) = ex1[kx] * (ex2[ky] * ex3[kz]), ccpsdv3/strfac.c:93-123 */
void oracle_strfac(int nfft, const double *ex1, const double *ex2,
                   const double *ex3, double *exi)
{
   int i, j, k, nh = nfft / 2 + 1;
   size_t kk = 0;
   for (k = 0; k < nfft; ++k)
      for (j = 0; j < nfft; ++j) {
         double tr = ex2[2*j] * ex3[2*k] - ex2[2*j+1] * ex3[2*k+1];
         double ti = ex2[2*j] * ex3[2*k+1] + ex2[2*j+1] * ex3[2*k];
         for (i = 0; i < nh; ++i) {
            exi[kk]   = ex1[2*i] * tr - ex1[2*i+1] * ti;
            exi[kk+1] = ex1[2*i+1] * tr + ex1[2*i] * ti;
            kk += 2;
         }
      }
}

/* ---- LDA, spin-restricted -------------------------------------------- */

/* Slater exchange, ccpsdv3/dirac_exchange.c:64-92 (alpha = 2/3 default at
 * :14, constants :18-24). */
void oracle_dirac_exchange(int n, const double *rho, double *vx, double *ex)
{
   const double alpha = 0.6666666666666666667;
   const double onethird = 0.33333333333333333333;
   const double pi = 3.14159265358979311600;
   int i;
   for (i = 0; i < n; ++i) {
      double nn = rho[i] + rho[i];
      double n13 = pow(3.0 * nn / pi, onethird);
      ex[i] = -(9.0 / 8.0) * alpha * n13;
      vx[i] = -(3.0 / 2.0) * alpha * n13;
   }
}

/* VWN correlation added onto the exchange arrays, ccpsdv3/vosko.c:99-167
 * with ispin = 1 (nup = ndown, xi = 0, so the spin-interpolation terms are
 * exactly zero).  Constants are the reference's literals (:24-71, :90-97). */
void oracle_vosko(int n, const double *dn, double *xcp, double *xce)
{
   const double bp = 3.727440e+00, cp = 1.293520e+01;
   const double cp1 = 3.109070e-02, cp2 = 9.690228e-04, cp3 = 1.049800e-01;
   const double cp4 = 3.878329e-02, cp5 = 3.075995e+00, cp6 = 1.863720e+00;
   const double dp1 = 6.218140e-02, dp2 = 1.938045e-03;
   const double dp4 = -3.205972e-02, dp5 = -1.192972e-01, dp6 = 1.863720e+00;
   const double dp7 = 9.461748e+00;
   const double onesixth = 0.166666666666667, onethird = 0.333333333333333;
   const double rs_scale = 0.620350490899400;
   const double small_number = 1.0e-30;
   int i;
   for (i = 0; i < n; ++i) {
      double nup = dn[i] + small_number;
      double nt = nup + nup;
      double rs = rs_scale / pow(nt, onethird);
      double x = sqrt(rs);
      double xxp = rs + bp * x + cp;
      double dxxp = 2.0 * x + bp;
      double ecp = cp1 * log(rs / xxp) + cp2 * log((x + cp3) * (x + cp3) / xxp)
                 + cp4 * atan(cp5 / (x + cp6));
      double ucp = ecp - onesixth * x * (dp1 / x + dp2 / (x + cp3)
                 + dp4 * dxxp / xxp + dp5 / ((x + dp6) * (x + dp6) + dp7));
      xce[i] += ecp;
      xcp[i] += ucp;
   }
}

/* erfc approximation of ccpsdv3/cpsd.c:11-17 used at :541-543:
 * (1 + b1 w + ... + b6 w^6)^(-16) */
double oracle_erfc_as(double w)
{
   const double B1 = 0.0705230784, B2 = 0.0422820123, B3 = 0.0092705272;
   const double B4 = 0.0001520143, B5 = 0.0002765672, B6 = 0.0000430638;
   double aa = 1.0 + w * (B1 + w * (B2 + w * (B3 + w * (B4 + w * (B5 + w * B6)))));
   double e = 1.0 / (aa * aa * aa * aa);
   return e * e * e * e;
}

/* ---- the SD step ------------------------------------------------------ */

static int g_lag_iters = 0;
int oracle_last_lagrange_iterations(void) { return g_lag_iters; }

/* ---- preconditioned SD (test infrastructure for the library's
 * cpsd_dev_set_preconditioner; no counterpart in the reference) ---------- */
static const double *g_precond_K;   /* [nfft3d], one factor per complex G */
static const double *g_precond_h;   /* [ne*ne] column-major */
static double *g_precond_D;         /* [ne*ne] column-major, the removed c1 component */

void oracle_set_preconditioner(const double *K, const double *h)
{
   g_precond_K = K;
   g_precond_h = h;
}

static void ensure_precond_D(int ne)
{
   static int cap = 0;
   if (ne * ne > cap) {
      free(g_precond_D);
      g_precond_D = (double *) malloc((size_t)ne * ne * sizeof(double));
      cap = ne * ne;
   }
}

/* The overlaps of the preconditioned step.  K o R is not orthogonal to c1
 * where K varies, which moves <c1|c2> away from the identity and slows the
 * Lagrange fixed point; the step's component along c1 is taken out first:
 * c2' = c2 - c1 D, D = <c1|c2> - I, and the s matrices are formed from
 *   A12' = A12 - A11 D,  A22' = A22 - D^T A12 - A12^T D + D^T A11 D
 * (A_ab(i,j) = <a_i|b_j>, all ne x ne, column-major).  The constraint then
 * adds dte c1 (X - D/dte) to c2. */
static void precond_overlaps(int nfft, size_t n2, int ne, double dte, const double *c1, const double *c2,
                             double *s11, double *s12, double *s21, double *s22, double *D)
{
   const size_t kk = (size_t)ne * ne;
   double *A11 = (double *) malloc(kk * sizeof(double)), *A12 = (double *) malloc(kk * sizeof(double));
   double *A22 = (double *) malloc(kk * sizeof(double)), *P = (double *) malloc(kk * sizeof(double));
   int i, j, k;
   for (j = 0; j < ne; ++j)
      for (i = 0; i < ne; ++i) {
         A12[i + (size_t)j*ne] = oracle_gcdotc(nfft, c1 + i*n2, c2 + j*n2);
         if (i <= j) {
            A11[i + (size_t)j*ne] = A11[j + (size_t)i*ne] = oracle_gcdotc(nfft, c1 + i*n2, c1 + j*n2);
            A22[i + (size_t)j*ne] = A22[j + (size_t)i*ne] = oracle_gcdotc(nfft, c2 + i*n2, c2 + j*n2);
         }
      }
   for (j = 0; j < ne; ++j)
      for (i = 0; i < ne; ++i) D[i + (size_t)j*ne] = A12[i + (size_t)j*ne] - (i == j ? 1.0 : 0.0);
   /* P = A11 D */
   for (j = 0; j < ne; ++j)
      for (i = 0; i < ne; ++i) {
         double a = 0.0;
         for (k = 0; k < ne; ++k) a += A11[i + (size_t)k*ne] * D[k + (size_t)j*ne];
         P[i + (size_t)j*ne] = a;
      }
   for (j = 0; j < ne; ++j)
      for (i = 0; i < ne; ++i) {
         /* (D^T A12)(i,j), (D^T A12)(j,i) and (D^T P)(i,j) */
         double q = 0.0, qt = 0.0, r = 0.0;
         for (k = 0; k < ne; ++k) {
            q += D[k + (size_t)i*ne] * A12[k + (size_t)j*ne];
            qt += D[k + (size_t)j*ne] * A12[k + (size_t)i*ne];
            r += D[k + (size_t)i*ne] * P[k + (size_t)j*ne];
         }
         const double a22 = A22[i + (size_t)j*ne] - q - qt + r;
         const double a12 = A12[i + (size_t)j*ne] - P[i + (size_t)j*ne];    /* A12'(i,j) */
         const double a21 = A12[j + (size_t)i*ne] - P[j + (size_t)i*ne];    /* A21'(i,j) = A12'(j,i) */
         const double d = (i == j) ? 1.0 : 0.0;
         s22[i + (size_t)j*ne] = (d - a22) * 0.5 / dte;
         s21[i + (size_t)j*ne] = (d - a21) * 0.5;
         s12[i + (size_t)j*ne] = (d - a12) * 0.5;
         s11[i + (size_t)j*ne] = -A11[i + (size_t)j*ne] * 0.5 * dte;
      }
   free(A11); free(A12); free(A22); free(P);
}

/* c2 = c1 + K o (c2 - c1 + dte c1 h): with c2 = c1 + dte cpsi on entry this
 * is c1 + K o dte (cpsi + c1 h), the residual -H c1 + c1 h scaled per G */
static void apply_preconditioner(int nfft3d, size_t n2, int ne, double dte, const double *c1, double *c2)
{
   double *t = (double *) malloc(n2 * ne * sizeof(double));
   size_t k;
   int i, m;
   for (i = 0; i < ne; ++i) {
      double *ti = t + i * n2;
      for (k = 0; k < n2; ++k) ti[k] = 0.0;
      for (m = 0; m < ne; ++m) {
         const double hm = g_precond_h[m + (size_t)i * ne];
         const double *cm = c1 + m * n2;
         for (k = 0; k < n2; ++k) ti[k] += cm[k] * hm;
      }
   }
   for (i = 0; i < ne; ++i) {
      const double *a = c1 + i * n2, *ti = t + i * n2;
      double *b = c2 + i * n2;
      for (k = 0; k < (size_t)nfft3d; ++k) {
         const double f = g_precond_K[k];
         b[2*k]     = a[2*k]     + f * (b[2*k]     - a[2*k]     + dte * ti[2*k]);
         b[2*k + 1] = a[2*k + 1] + f * (b[2*k + 1] - a[2*k + 1] + dte * ti[2*k + 1]);
      }
   }
   free(t);
}

int oracle_cpsd(int inner, int move, int nion, int ispin, const int *ne,
                int nfft, int nfft3d, int n2ft3d, const double *unitg,
                double omega, double dt, double fmass, const int *katm,
                const double *amass, const double *zv, int nsh,
                const double *rcell, double cewald, double rcut,
                const int *masker, const double *G, const double *tg,
                const double *vc, const double *vg, int *lmmax,
                double *vnlnrm, double *vl, double *vnl, double *r1,
                double *r2, double *fion, double *c1, double *c2,
                double *cpsi, double *dn, double *hml, double *lmd,
                double *cpsdwork, double *E, double *deltae, double *deltac,
                double *deltar)
{
   const int nemax = ne[0] + ne[1], n0 = ne[0];
   const size_t n2 = (size_t)n2ft3d, nn = n2 * nemax;
   const double scal1 = 1.0 / ((double)nfft * nfft * nfft);
   const double scal2 = 1.0 / omega;
   const double dv = omega * scal1;
   const double dte = dt / sqrt(fmass);
   double *dng, *vall, *vtmp, *s, *exi, *xcp, *xce, *ex1, *ex2, *ex3;
   double *s11, *s12, *s21, *s22, *st1, *sa1, *sa0;
   double eorbit, eion, exc, pxc, ehartr, eke, elocal, enlocal, eold;
   size_t jr = 0, k;
   int itime, i, j, ii, jj, l;

   double *dti = NULL;
   if (ispin != 1 || ne[1] != 0) {
      fprintf(stderr, "oracle_cpsd: only ispin=1 is restated\n");
      return 1;
   }
   if (nfft3d != (nfft / 2 + 1) * nfft * nfft || n2ft3d != (nfft + 2) * nfft * nfft) {
      fprintf(stderr, "oracle_cpsd: inconsistent grid sizes\n");
      return 2;
   }

   /* ionic time steps first, then the scratch, as in ccpsdv3/cpsd.c:116-165 */
   if (move) {
      dti = cpsdwork + jr; jr += (size_t)nion;
      for (ii = 0; ii < nion; ++ii) dti[ii] = dt / sqrt(amass[katm[ii]]);
   }
   dng  = cpsdwork + jr; jr += n2;
   vall = cpsdwork + jr; jr += n2;
   vtmp = cpsdwork + jr; jr += n2;
   s    = cpsdwork + jr; jr += n2;
   exi  = cpsdwork + jr; jr += n2;
   xcp  = cpsdwork + jr; jr += n2;
   xce  = cpsdwork + jr; jr += n2;
   ex1  = cpsdwork + jr; jr += (size_t)nion * (nfft + 2);
   ex2  = cpsdwork + jr; jr += (size_t)nion * 2 * nfft;
   ex3  = cpsdwork + jr; jr += (size_t)nion * 2 * nfft;
   s11  = cpsdwork + jr; jr += (size_t)n0 * n0;
   s12  = cpsdwork + jr; jr += (size_t)n0 * n0;
   s21  = cpsdwork + jr; jr += (size_t)n0 * n0;
   s22  = cpsdwork + jr; jr += (size_t)n0 * n0;
   st1  = cpsdwork + jr; jr += (size_t)n0 * n0;
   sa1  = cpsdwork + jr; jr += (size_t)n0 * n0;
   sa0  = cpsdwork + jr; jr += (size_t)n0 * n0;

   if (g_precond_K) ensure_precond_D(n0);
   for (itime = 0; itime < inner; ++itime) {
      /* :173-181 */
      memcpy(c1, c2, nn * sizeof(double));
      memcpy(r1, r2, 3 * (size_t)nion * sizeof(double));
      memcpy(cpsi, c1, nn * sizeof(double));
      for (i = 0; i < nemax; ++i) oracle_cr_fft3b(nfft, cpsi + i * n2);

      /* density, :185-194 */
      for (k = 0; k < n2; ++k) dn[k] = 0.0;
      for (i = 0; i < n0; ++i) {
         const double *p = cpsi + i * n2;
         for (k = 0; k < n2; ++k) dn[k] += scal2 * (p[k] * p[k]);
      }
      for (j = 0; j < nfft * nfft; ++j) {
         dn[(size_t)j * (nfft + 2) + nfft] = 0.0;
         dn[(size_t)j * (nfft + 2) + nfft + 1] = 0.0;
      }
      /* :198-205 */
      for (k = 0; k < n2; ++k) dng[k] = (dn[k] + dn[k]) * scal1;
      oracle_rc_fft3f(nfft, dng);
      for (k = 0; k < n2; ++k) dng[k] *= masker[k];

      oracle_phafac(nfft, unitg, nion, r1, ex1, ex2, ex3);

      /* kinetic, :212-221 */
      for (i = 0; i < nemax; ++i)
         for (k = 0; k < (size_t)nfft3d; ++k) {
            c2[i*n2 + 2*k]   = -tg[k] * c1[i*n2 + 2*k];
            c2[i*n2 + 2*k+1] = -tg[k] * c1[i*n2 + 2*k+1];
         }

      /* local and nonlocal pseudopotential, :224-307 */
      for (k = 0; k < n2; ++k) { s[k] = 0.0; vall[k] = 0.0; }
      for (ii = 0; ii < nion; ++ii) {
         int ia = katm[ii];
         const double *rex;
         oracle_strfac(nfft, ex1 + (size_t)ii*(nfft+2), ex2 + (size_t)ii*2*nfft,
                       ex3 + (size_t)ii*2*nfft, exi);
         for (k = 0; k < n2; ++k) exi[k] *= masker[k];
         axpy(n2, zv[ia], exi, s);
         rex = vl + (size_t)ia * nfft3d;
         for (k = 0; k < (size_t)nfft3d; ++k) {
            vtmp[2*k]   = exi[2*k]   * rex[k];
            vtmp[2*k+1] = exi[2*k+1] * rex[k];
            vall[2*k]   += vtmp[2*k];
            vall[2*k+1] += vtmp[2*k+1];
         }
         if (move) {
            /* local force, :246-257 */
            for (k = 0; k < (size_t)nfft3d; ++k)
               xce[k] = dng[2*k+1] * vtmp[2*k] - dng[2*k] * vtmp[2*k+1];
            fion[3*ii]   = oracle_gsdot(nfft, G, xce);
            fion[3*ii+1] = oracle_gsdot(nfft, G + nfft3d, xce);
            fion[3*ii+2] = oracle_gsdot(nfft, G + 2*(size_t)nfft3d, xce);
         }
         for (l = 0; l < lmmax[ia]; ++l) {
            rex = vnl + (size_t)ia * 9 * nfft3d + (size_t)l * nfft3d;
            if (l == 0 || l >= 4) {
               for (k = 0; k < (size_t)nfft3d; ++k) {
                  vtmp[2*k]   = exi[2*k]   * rex[k];
                  vtmp[2*k+1] = exi[2*k+1] * rex[k];
               }
            } else {
               /* p projectors carry a factor i (:277-286) */
               for (k = 0; k < (size_t)nfft3d; ++k) {
                  vtmp[2*k]   = -exi[2*k+1] * rex[k];
                  vtmp[2*k+1] =  exi[2*k]   * rex[k];
               }
            }
            for (i = 0; i < nemax; ++i) {
               double sw1 = -oracle_gcdotc(nfft, vtmp, c1 + i*n2) * scal2
                            / vnlnrm[ia*9 + l];
               axpy(n2, sw1, vtmp, c2 + i*n2);
               if (move) {
                  /* nonlocal force, :291-304 */
                  const double *rex1 = c1 + i * n2;
                  for (k = 0; k < (size_t)nfft3d; ++k)
                     xce[k] = rex1[2*k+1] * vtmp[2*k] - rex1[2*k] * vtmp[2*k+1];
                  sw1 *= 2;   /* ispin == 1 */
                  fion[3*ii]   -= 2 * sw1 * oracle_gsdot(nfft, G, xce);
                  fion[3*ii+1] -= 2 * sw1 * oracle_gsdot(nfft, G + nfft3d, xce);
                  fion[3*ii+2] -= 2 * sw1 * oracle_gsdot(nfft, G + 2*(size_t)nfft3d, xce);
               }
            }
         }
      }

      /* Ewald forces, :310-364 */
      if (move) {
         const double CERFC = 1.128379167;
         const double B1 = 0.0705230784, B2 = 0.0422820123, B3 = 0.0092705272;
         const double B4 = 0.0001520143, B5 = 0.0002765672, B6 = 0.0000430638;
         for (ii = 0; ii < nion; ++ii) {
            int ia = katm[ii];
            oracle_strfac(nfft, ex1 + (size_t)ii*(nfft+2), ex2 + (size_t)ii*2*nfft,
                          ex3 + (size_t)ii*2*nfft, exi);
            for (k = 0; k < n2; ++k) exi[k] *= masker[k];
            /* reciprocal space */
            for (k = 0; k < (size_t)nfft3d; ++k)
               vtmp[k] = (exi[2*k] * s[2*k+1] - exi[2*k+1] * s[2*k]) * vg[k];
            fion[3*ii]   += oracle_gsdot(nfft, G, vtmp) * zv[ia] * scal2;
            fion[3*ii+1] += oracle_gsdot(nfft, G + nfft3d, vtmp) * zv[ia] * scal2;
            fion[3*ii+2] += oracle_gsdot(nfft, G + 2*(size_t)nfft3d, vtmp) * zv[ia] * scal2;
            /* real space, pairs over the nsh images */
            for (jj = ii + 1; jj < nion; ++jj) {
               double dx = r1[3*ii] - r1[3*jj];
               double dy = r1[3*ii+1] - r1[3*jj+1];
               double dz = r1[3*ii+2] - r1[3*jj+2];
               double zz = zv[katm[ii]] * zv[katm[jj]];
               double sw1 = 0.0, sw2 = 0.0, sw3 = 0.0;
               for (l = 0; l < nsh; ++l) {
                  double x = rcell[l] + dx, y = rcell[l+nsh] + dy, z = rcell[l+2*nsh] + dz;
                  double r = sqrt(x*x + y*y + z*z);
                  double w = r / rcut;
                  double aa = 1.0 + w * (B1 + w * (B2 + w * (B3 + w * (B4 + w * (B5 + w * B6)))));
                  double erfc, adiff;
                  aa *= aa * aa * aa;
                  erfc = 1.0 / (aa * aa * aa * aa);
                  adiff = zz * (erfc + CERFC * w * exp(-w * w)) / (r * r * r);
                  sw1 += x * adiff;
                  sw2 += y * adiff;
                  sw3 += z * adiff;
               }
               fion[3*ii]   += sw1; fion[3*ii+1] += sw2; fion[3*ii+2] += sw3;
               fion[3*jj]   -= sw1; fion[3*jj+1] -= sw2; fion[3*jj+2] -= sw3;
            }
         }
      }

      /* LDA, :370-371 */
      oracle_dirac_exchange(n2ft3d, dn, xcp, xce);
      oracle_vosko(n2ft3d, dn, xcp, xce);

      /* total local potential in real space, :373-382 */
      for (k = 0; k < (size_t)nfft3d; ++k) {
         vtmp[2*k]   = vall[2*k]   * scal2 + dng[2*k]   * vc[k];
         vtmp[2*k+1] = vall[2*k+1] * scal2 + dng[2*k+1] * vc[k];
      }
      oracle_cr_fft3b(nfft, vtmp);

      /* -H psi, :384-396 */
      for (i = 0; i < n0; ++i) {
         double *p = cpsi + i * n2;
         for (k = 0; k < n2; ++k) p[k] *= (vtmp[k] + xcp[k]);
         oracle_rc_fft3f(nfft, p);
         for (k = 0; k < n2; ++k)
            p[k] = -scal1 * p[k] * masker[k] + c2[i*n2 + k];
      }

      /* steepest descent, :399-411 */
      memcpy(c2, c1, nn * sizeof(double));
      axpy(nn, dte, cpsi, c2);
      /* preconditioned form (no counterpart in the reference; the checker
       * of the library's preconditioned step): the step of every orbital
       * becomes K(G) times the residual with a given Lagrange matrix h,
       * c2 = c1 + K o (dte cpsi + dte c1 h) */
      if (g_precond_K) apply_preconditioner(nfft3d, n2, n0, dte, c1, c2);
      if (move)
         for (ii = 0; ii < nion; ++ii)
            for (j = 0; j < 3; ++j) r2[3*ii+j] = r1[3*ii+j] + dti[ii] * fion[3*ii+j];

      /* overlap matrices, :415-437 (note: s21 and s12 are filled from one
       * triangle each and mirrored, exactly as the reference does); the
       * preconditioned form first takes the step's component along c1 out
       * of c2 (precond_overlaps) */
      if (g_precond_K) {
         precond_overlaps(nfft, n2, n0, dte, c1, c2, s11, s12, s21, s22, g_precond_D);
      } else {
         const double *C1 = c1, *C2 = c2;
         for (i = 0; i < n0; ++i) {
            s22[i + i*n0] = (1.0 - oracle_gcdotc(nfft, C2 + i*n2, C2 + i*n2)) * 0.5 / dte;
            s21[i + i*n0] = (1.0 - oracle_gcdotc(nfft, C2 + i*n2, C1 + i*n2)) * 0.5;
            s12[i + i*n0] = s21[i + i*n0];
            s11[i + i*n0] = -oracle_gcdotc(nfft, C1 + i*n2, C1 + i*n2) * 0.5 * dte;
            for (j = i + 1; j < n0; ++j) {
               s22[i + j*n0] = -oracle_gcdotc(nfft, C2 + i*n2, C2 + j*n2) * 0.5 / dte;
               s21[i + j*n0] = -oracle_gcdotc(nfft, C2 + i*n2, C1 + j*n2) * 0.5;
               s12[i + j*n0] = -oracle_gcdotc(nfft, C1 + i*n2, C2 + j*n2) * 0.5;
               s11[i + j*n0] = -oracle_gcdotc(nfft, C1 + i*n2, C1 + j*n2) * 0.5 * dte;
               s22[j + i*n0] = s22[i + j*n0];
               s21[j + i*n0] = s21[i + j*n0];
               s12[j + i*n0] = s12[i + j*n0];
               s11[j + i*n0] = s11[i + j*n0];
            }
         }
      }

      /* Lagrange multiplier fixed point, :439-478 (ITERLMD = 30,
       * CONVGLMD = 1e-15 at :19-20) */
      {
         const size_t kk = (size_t)n0 * n0;
         size_t q;
         int it = 0, done = 0;
         double adiff = 0.0;
         memcpy(sa0, s22, kk * sizeof(double));
         while (!done && (it++) < 30) {
            memcpy(sa1, s22, kk * sizeof(double));
            gemm_nn(n0, n0, n0, 1.0, s21, n0, sa0, n0, 1.0, sa1, n0);
            gemm_nn(n0, n0, n0, 1.0, sa0, n0, s12, n0, 1.0, sa1, n0);
            gemm_nn(n0, n0, n0, 1.0, s11, n0, sa0, n0, 0.0, st1, n0);
            gemm_nn(n0, n0, n0, 1.0, sa0, n0, st1, n0, 1.0, sa1, n0);
            adiff = 0.0;
            for (q = 0; q < kk; ++q) {
               double d = fabs(sa1[q] - sa0[q]);
               st1[q] = sa1[q] - sa0[q];
               if (d > adiff) adiff = d;
            }
            if (adiff < 1e-15) done = 1;
            else memcpy(sa0, sa1, kk * sizeof(double));
         }
         g_lag_iters = it;
         if (!done) printf("ierr=10 adiff=%f\n", adiff);
         memcpy(lmd, sa1, kk * sizeof(double));
         /* preconditioned form: c2f = c2' + dte c1 X = c2 + dte c1 (X - D/dte) */
         if (g_precond_K)
            for (q = 0; q < kk; ++q) lmd[q] -= g_precond_D[q] / dte;
      }

      /* constraint, :482-488 */
      gemm_nn((int)n2, n0, n0, dte, c1, n2, lmd, n0, 1.0, c2, n2);
   }

   /* ---- energy tail, :494-672 ---- */
   eorbit = 0.0;
   for (i = 0; i < n0; ++i) {
      hml[i + i*n0] = -oracle_gcdotc(nfft, c1 + i*n2, cpsi + i*n2);
      eorbit += hml[i + i*n0];
      for (j = i + 1; j < n0; ++j) {
         hml[i + j*n0] = -oracle_gcdotc(nfft, c1 + i*n2, cpsi + j*n2);
         hml[j + i*n0] = hml[i + j*n0];
      }
   }
   eorbit *= 2;

   for (k = 0; k < (size_t)nfft3d; ++k)
      vtmp[k] = (s[2*k] * s[2*k] + s[2*k+1] * s[2*k+1]) * vg[k];
   eion = oracle_gssum(nfft, vtmp) * 0.5 * scal2 + cewald;
   for (ii = 0; ii < nion; ++ii)
      for (jj = ii + 1; jj < nion; ++jj) {
         double dx = r1[3*ii] - r1[3*jj];
         double dy = r1[3*ii+1] - r1[3*jj+1];
         double dz = r1[3*ii+2] - r1[3*jj+2];
         double zz = zv[katm[ii]] * zv[katm[jj]];
         for (l = 0; l < nsh; ++l) {
            double x = rcell[l] + dx, y = rcell[l+nsh] + dy, z = rcell[l+2*nsh] + dz;
            double r = sqrt(x*x + y*y + z*z);
            eion += zz * oracle_erfc_as(r / rcut) / r;
         }
      }

   for (k = 0; k < (size_t)nfft3d; ++k)
      vtmp[k] = (dng[2*k] * dng[2*k] + dng[2*k+1] * dng[2*k+1]) * vc[k];
   ehartr = oracle_gssum(nfft, vtmp) * omega * 0.5;

   exc = dot_s(n2ft3d, dn, 1, xce, 1);
   pxc = dot_s(n2ft3d, dn, 1, xcp, 1);
   exc += exc; pxc += pxc;
   exc *= dv; pxc *= dv;

   eke = 0.0;
   for (i = 0; i < nemax; ++i) {
      const double *rex = c1 + i * n2;
      for (k = 0; k < (size_t)nfft3d; ++k) {
         vtmp[2*k]   = rex[2*k]   * tg[k];
         vtmp[2*k+1] = rex[2*k+1] * tg[k];
      }
      eke += oracle_gcdotc(nfft, rex, vtmp);
   }
   eke *= 2;

   elocal = oracle_gcdotc(nfft, vall, dng);

   enlocal = 0.0;
   for (ii = 0; ii < nion; ++ii) {
      int ia = katm[ii];
      oracle_strfac(nfft, ex1 + (size_t)ii*(nfft+2), ex2 + (size_t)ii*2*nfft,
                    ex3 + (size_t)ii*2*nfft, exi);
      for (k = 0; k < n2; ++k) exi[k] *= masker[k];
      for (l = 0; l < lmmax[ia]; ++l) {
         const double *rex = vnl + (size_t)ia * 9 * nfft3d + (size_t)l * nfft3d;
         if (l == 0 || l >= 4) {
            for (k = 0; k < (size_t)nfft3d; ++k) {
               vtmp[2*k]   = exi[2*k]   * rex[k];
               vtmp[2*k+1] = exi[2*k+1] * rex[k];
            }
         } else {
            for (k = 0; k < (size_t)nfft3d; ++k) {
               vtmp[2*k]   = -exi[2*k+1] * rex[k];
               vtmp[2*k+1] =  exi[2*k]   * rex[k];
            }
         }
         for (i = 0; i < nemax; ++i) {
            double sw1 = oracle_gcdotc(nfft, vtmp, c1 + i*n2);
            enlocal += sw1 * sw1 * scal2 / vnlnrm[ia*9 + l];
         }
      }
   }
   enlocal *= 2;

   eold = E[0];
   E[0] = eorbit + eion + exc - ehartr - pxc;
   E[1] = eorbit;
   E[2] = ehartr;
   E[3] = exc;
   E[4] = eion;
   E[5] = eke;
   E[6] = elocal;
   E[7] = enlocal;
   E[8] = 2.0 * ehartr;
   E[9] = pxc;
   *deltae = (E[0] - eold) / (dt * inner);

   for (k = 0; k < nn; ++k) cpsi[k] = c2[k] - c1[k];
   *deltac = 0.0;
   for (i = 0; i < nemax; ++i) {
      double a = oracle_gcdotc(nfft, cpsi + i*n2, cpsi + i*n2);
      if (a > *deltac) *deltac = a;
   }
   *deltac /= dte;
   *deltar = 0.0;
   if (move)
      for (ii = 0; ii < nion; ++ii) {
         double a = sqrt(fion[3*ii] * fion[3*ii] + fion[3*ii+1] * fion[3*ii+1]
                         + fion[3*ii+2] * fion[3*ii+2]);
         if (a > *deltar) *deltar = a;
      }
   return 0;
}

/* ---- epilogue eigen-solver (ccpsdv3/main.c:470-484, eigsrt.c:11-39) --- */

void oracle_syev_sorted(int n, double *a, double *eig)
{
   /* cyclic Jacobi on a copy; a returns the eigenvectors (columns) */
   double *m = (double *) malloc((size_t)n * n * sizeof(double));
   double *v = a;
   int i, j, p, q, sweep;
   memcpy(m, a, (size_t)n * n * sizeof(double));
   for (i = 0; i < n; ++i) for (j = 0; j < n; ++j) v[i + j*n] = (i == j);
   for (sweep = 0; sweep < 100; ++sweep) {
      double off = 0.0;
      for (p = 0; p < n; ++p) for (q = p + 1; q < n; ++q) off += m[p + q*n] * m[p + q*n];
      if (off < 1e-30) break;
      for (p = 0; p < n; ++p)
         for (q = p + 1; q < n; ++q) {
            double apq = m[p + q*n], app, aqq, theta, t, c, s;
            if (fabs(apq) < 1e-300) continue;
            app = m[p + p*n]; aqq = m[q + q*n];
            theta = (aqq - app) / (2.0 * apq);
            t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta*theta + 1.0));
            c = 1.0 / sqrt(t*t + 1.0); s = t * c;
            for (i = 0; i < n; ++i) {
               double mip = m[i + p*n], miq = m[i + q*n];
               m[i + p*n] = c * mip - s * miq;
               m[i + q*n] = s * mip + c * miq;
            }
            for (i = 0; i < n; ++i) {
               double mpi = m[p + i*n], mqi = m[q + i*n];
               m[p + i*n] = c * mpi - s * mqi;
               m[q + i*n] = s * mpi + c * mqi;
            }
            for (i = 0; i < n; ++i) {
               double vip = v[i + p*n], viq = v[i + q*n];
               v[i + p*n] = c * vip - s * viq;
               v[i + q*n] = s * vip + c * viq;
            }
         }
   }
   for (i = 0; i < n; ++i) eig[i] = m[i + i*n];
   /* selection sort, largest first, as eigsrt does */
   for (i = 0; i < n - 1; ++i) {
      int k = i; double pmax = eig[i];
      for (j = i + 1; j < n; ++j) if (eig[j] >= pmax) { k = j; pmax = eig[j]; }
      if (k != i) {
         eig[k] = eig[i]; eig[i] = pmax;
         for (j = 0; j < n; ++j) { double t = v[j + i*n]; v[j + i*n] = v[j + k*n]; v[j + k*n] = t; }
      }
   }
   free(m);
}
