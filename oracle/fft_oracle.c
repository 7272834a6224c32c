/*
 * fft_oracle.c -- test infrastructure only (see oracle.h).
 *
 * Own mixed-radix FFT restating the conventions of the reference's FFT layer:
 *   - 1D transforms: FFTPACK dcfftf/dcfftb (unnormalised, e^{-i}/e^{+i}),
 *     drfftf/drfftb real packing (fftpack/fftdoc, rfftf/rfftb sections);
 *   - 3D wrappers: fftpack3d/fftpack3d.f:22-96 (cr_fft3b: z, y complex
 *     backward passes over kx = 0..N/2, then x complex-to-real with the
 *     imaginary parts of kx = 0 and kx = N/2 ignored and the two pad reals of
 *     every x-row zeroed) and :107-196 (rc_fft3f: x real-to-complex with
 *     Im(kx=0) = Im(kx=N/2) = 0, then y and z complex forward passes).
 * The algorithm (iterative Stockham autosort over radices 4, 2, 3, 5 and
 * generic odd primes, twiddles from cos/sin of exactly reduced angles) is
 * this file's own; only the transform definitions and data layout follow
 * the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "oracle.h"

#define MAXN 4096
#define MAXPASS 16

/* Plan: radices and per-pass twiddle tables for an iterative Stockham
 * autosort transform.  Pass q with radix R and Ns = product of earlier
 * radices maps, for j < n/R (k = j mod Ns):
 *    v_r = in[j + r n/R] * exp(sign 2 pi i r k / (Ns R)),  r < R
 *    V   = DFT_R(v)
 *    out[(j/Ns) Ns R + k + r Ns] = V_r
 * and the output of the last pass is in natural order. */
typedef struct {
   int n, sign, npass;
   int radix[MAXPASS];
   double *tw[MAXPASS];      /* [k][r] complex, k < Ns, r < R */
   double *rtw[MAXPASS];     /* DFT_R roots exp(sign 2 pi i m / R) */
} plan_t;

static plan_t *plans[2][MAXN + 1];

static plan_t *get_plan(int n, int sign)
{
   int s = (sign > 0), q, ns, m, k, r;
   plan_t *p;
   if (plans[s][n]) return plans[s][n];
   p = (plan_t *) calloc(1, sizeof(plan_t));
   p->n = n; p->sign = sign;
   m = n;
   while (m > 1) {
      int f;
      if (m % 4 == 0) f = 4;
      else if (m % 2 == 0) f = 2;
      else if (m % 3 == 0) f = 3;
      else if (m % 5 == 0) f = 5;
      else { for (f = 7; m % f; f += 2) ; }
      p->radix[p->npass++] = f;
      m /= f;
   }
   ns = 1;
   for (q = 0; q < p->npass; ++q) {
      int R = p->radix[q];
      p->tw[q] = (double *) malloc(2 * (size_t)ns * R * sizeof(double));
      for (k = 0; k < ns; ++k)
         for (r = 0; r < R; ++r) {
            /* exact integer reduction of the angle numerator */
            long num = ((long)r * k) % ((long)ns * R);
            double a = 2.0 * M_PI * (double)num / (double)(ns * R);
            p->tw[q][2*(k*R + r)]     = cos(a);
            p->tw[q][2*(k*R + r) + 1] = sign * sin(a);
         }
      p->rtw[q] = (double *) malloc(2 * (size_t)R * sizeof(double));
      for (r = 0; r < R; ++r) {
         double a = 2.0 * M_PI * (double)r / (double)R;
         p->rtw[q][2*r] = cos(a); p->rtw[q][2*r+1] = sign * sin(a);
      }
      ns *= R;
   }
   plans[s][n] = p;
   return p;
}

/* in-register DFT of length R (R <= 32) with roots w[m] = exp(sign 2 pi i m/R) */
static void dft_small(int R, int sign, double *v, const double *w)
{
   double t[64];
   int a, b;
   if (R == 2) {
      double r0 = v[0], i0 = v[1];
      v[0] = r0 + v[2]; v[1] = i0 + v[3];
      v[2] = r0 - v[2]; v[3] = i0 - v[3];
      return;
   }
   if (R == 4) {
      double ar = v[0] + v[4], ai = v[1] + v[5];
      double br = v[0] - v[4], bi = v[1] - v[5];
      double cr = v[2] + v[6], ci = v[3] + v[7];
      double dr = v[2] - v[6], di = v[3] - v[7];
      /* multiply d by sign*i */
      double er = -sign * di, ei = sign * dr;
      v[0] = ar + cr; v[1] = ai + ci;
      v[4] = ar - cr; v[5] = ai - ci;
      v[2] = br + er; v[3] = bi + ei;
      v[6] = br - er; v[7] = bi - ei;
      return;
   }
   for (a = 0; a < R; ++a) {
      double sr = 0.0, si = 0.0;
      for (b = 0; b < R; ++b) {
         int e = (a * b) % R;
         sr += v[2*b] * w[2*e] - v[2*b+1] * w[2*e+1];
         si += v[2*b] * w[2*e+1] + v[2*b+1] * w[2*e];
      }
      t[2*a] = sr; t[2*a+1] = si;
   }
   memcpy(v, t, 2 * R * sizeof(double));
}

void oracle_cfft1d(int n, int sign, double *x)
{
   plan_t *p;
   double buf[2 * MAXN], v[64];
   double *in = x, *out = buf;
   int q, j, r, ns = 1;
   if (n < 1 || n > MAXN) { fprintf(stderr, "oracle_cfft1d: bad n=%d\n", n); abort(); }
   if (n == 1) return;
   p = get_plan(n, sign);
   for (q = 0; q < p->npass; ++q) {
      int R = p->radix[q], nr = n / R;
      const double *tw = p->tw[q];
      if (R > 32) { fprintf(stderr, "oracle_cfft1d: radix %d too large\n", R); abort(); }
      for (j = 0; j < nr; ++j) {
         int k = j % ns;
         int base = (j / ns) * ns * R + k;
         for (r = 0; r < R; ++r) {
            double xr = in[2*(j + r*nr)], xi = in[2*(j + r*nr) + 1];
            double wr = tw[2*(k*R + r)], wi = tw[2*(k*R + r) + 1];
            v[2*r]   = xr * wr - xi * wi;
            v[2*r+1] = xr * wi + xi * wr;
         }
         dft_small(R, sign, v, p->rtw[q]);
         for (r = 0; r < R; ++r) {
            out[2*(base + r*ns)]     = v[2*r];
            out[2*(base + r*ns) + 1] = v[2*r+1];
         }
      }
      ns *= R;
      { double *t = in; in = out; out = t; }
   }
   if (in != x) memcpy(x, in, 2 * (size_t)n * sizeof(double));
}

/* ---- 3D wrappers ------------------------------------------------------ */

void oracle_cr_fft3b(int nfft, double *a)
{
   const int n = nfft, nh = n / 2 + 1, rowr = n + 2;
   int i, j, k;
   double *t = (double *) malloc(2 * n * sizeof(double));

   /* z pass: complex backward over kz for every (kx, ky) column
    * (fftpack3d.f:50-58) */
   for (j = 0; j < n; ++j)
      for (i = 0; i < nh; ++i) {
         for (k = 0; k < n; ++k) {
            size_t idx = 2 * ((size_t)i + (size_t)nh * (j + (size_t)n * k));
            t[2*k] = a[idx]; t[2*k+1] = a[idx+1];
         }
         oracle_cfft1d(n, +1, t);
         for (k = 0; k < n; ++k) {
            size_t idx = 2 * ((size_t)i + (size_t)nh * (j + (size_t)n * k));
            a[idx] = t[2*k]; a[idx+1] = t[2*k+1];
         }
      }
   /* y pass (fftpack3d.f:66-74) */
   for (k = 0; k < n; ++k)
      for (i = 0; i < nh; ++i) {
         for (j = 0; j < n; ++j) {
            size_t idx = 2 * ((size_t)i + (size_t)nh * (j + (size_t)n * k));
            t[2*j] = a[idx]; t[2*j+1] = a[idx+1];
         }
         oracle_cfft1d(n, +1, t);
         for (j = 0; j < n; ++j) {
            size_t idx = 2 * ((size_t)i + (size_t)nh * (j + (size_t)n * k));
            a[idx] = t[2*j]; a[idx+1] = t[2*j+1];
         }
      }
   /* x pass: complex-to-real (fftpack3d.f:81-94).  drfftb uses only
    * Re X_0, (Re,Im) X_1..X_{N/2-1} and Re X_{N/2}; evaluate it as the
    * complex backward transform of the Hermitian extension. */
   for (k = 0; k < n; ++k)
      for (j = 0; j < n; ++j) {
         double *row = a + (size_t)rowr * (j + (size_t)n * k);
         int q;
         t[0] = row[0]; t[1] = 0.0;
         for (q = 1; q < n / 2; ++q) {
            t[2*q] = row[2*q]; t[2*q+1] = row[2*q+1];
            t[2*(n-q)] = row[2*q]; t[2*(n-q)+1] = -row[2*q+1];
         }
         t[n] = row[n]; t[n+1] = 0.0;
         oracle_cfft1d(n, +1, t);
         for (q = 0; q < n; ++q) row[q] = t[2*q];
         row[n] = 0.0; row[n+1] = 0.0;
      }
   free(t);
}

void oracle_rc_fft3f(int nfft, double *a)
{
   const int n = nfft, nh = n / 2 + 1, rowr = n + 2;
   int i, j, k;
   double *t = (double *) malloc(2 * n * sizeof(double));

   /* x pass: real-to-complex, shifted into (re,im) pairs with
    * Im(kx=0) = Im(kx=N/2) = 0 (fftpack3d.f:151-164) */
   for (k = 0; k < n; ++k)
      for (j = 0; j < n; ++j) {
         double *row = a + (size_t)rowr * (j + (size_t)n * k);
         int q;
         for (q = 0; q < n; ++q) { t[2*q] = row[q]; t[2*q+1] = 0.0; }
         oracle_cfft1d(n, -1, t);
         for (q = 0; q <= n / 2; ++q) { row[2*q] = t[2*q]; row[2*q+1] = t[2*q+1]; }
         row[1] = 0.0;
         row[n+1] = 0.0;
      }
   /* y pass (fftpack3d.f:170-178) */
   for (k = 0; k < n; ++k)
      for (i = 0; i < nh; ++i) {
         for (j = 0; j < n; ++j) {
            size_t idx = 2 * ((size_t)i + (size_t)nh * (j + (size_t)n * k));
            t[2*j] = a[idx]; t[2*j+1] = a[idx+1];
         }
         oracle_cfft1d(n, -1, t);
         for (j = 0; j < n; ++j) {
            size_t idx = 2 * ((size_t)i + (size_t)nh * (j + (size_t)n * k));
            a[idx] = t[2*j]; a[idx+1] = t[2*j+1];
         }
      }
   /* z pass (fftpack3d.f:185-193) */
   for (j = 0; j < n; ++j)
      for (i = 0; i < nh; ++i) {
         for (k = 0; k < n; ++k) {
            size_t idx = 2 * ((size_t)i + (size_t)nh * (j + (size_t)n * k));
            t[2*k] = a[idx]; t[2*k+1] = a[idx+1];
         }
         oracle_cfft1d(n, -1, t);
         for (k = 0; k < n; ++k) {
            size_t idx = 2 * ((size_t)i + (size_t)nh * (j + (size_t)n * k));
            a[idx] = t[2*k]; a[idx+1] = t[2*k+1];
         }
      }
   free(t);
}
