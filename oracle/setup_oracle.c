/*
 * setup_oracle.c -- test infrastructure only (see oracle.h).
 *
 * Restatement of the inputs the reference builds before its SD loop:
 * lattice (ccpsdv3/get_cube.c:29-85), G vectors (ccpsdv3/get_g.c:12-35),
 * spherical cutoff mask (ccpsdv3/get_masker.c:12-37), kinetic / Coulomb /
 * Ewald tables, Ewald constant and image cells (ccpsdv3/main.c:145-198,
 * 287-313), the KB pseudopotential transform (ckbppv3/dkbppv3.c:310-444) and
 * the plane-wave starting orbitals (cmakepsi/dmakepsi.c:86-133).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void oracle_get_cube(int icube, double unit, double *omega, double *unita,
                     double *unitg)
{
   static const double sc[9]  = {1, 0, 0,  0, 1, 0,  0, 0, 1};
   static const double fcc[9] = {0.5, 0.5, 0.0,  0.5, 0.0, 0.5,  0.0, 0.5, 0.5};
   static const double bcc[9] = {0.5, 0.5, 0.5, -0.5, 0.5, 0.5,  0.5, 0.5, -0.5};
   const double twopi = 8.0 * atan(1.0);
   const double *base = icube == 2 ? fcc : (icube == 3 ? bcc : sc);
   double v, s;
   int i;
   if (icube < 1 || icube > 3) printf("get_cube: input error -- unknown lattice itype.\n");
   for (i = 0; i < 9; ++i) unita[i] = unit * base[i];
   /* b_i = a_j x a_k */
   unitg[0] = unita[4]*unita[8] - unita[5]*unita[7];
   unitg[1] = unita[5]*unita[6] - unita[3]*unita[8];
   unitg[2] = unita[3]*unita[7] - unita[4]*unita[6];
   unitg[3] = unita[7]*unita[2] - unita[8]*unita[1];
   unitg[4] = unita[8]*unita[0] - unita[6]*unita[2];
   unitg[5] = unita[6]*unita[1] - unita[7]*unita[0];
   unitg[6] = unita[1]*unita[5] - unita[2]*unita[4];
   unitg[7] = unita[2]*unita[3] - unita[0]*unita[5];
   unitg[8] = unita[0]*unita[4] - unita[1]*unita[3];
   v = unita[0]*unitg[0] + unita[1]*unitg[1] + unita[2]*unitg[2];
   s = twopi / v;
   for (i = 0; i < 9; ++i) unitg[i] *= s;
   *omega = fabs(v);
}

static int wrap(int k, int n) { return k < 0 ? k + n : k; }

void oracle_get_g(int nfft, const double *unitg, double *g)
{
   const int nh = nfft / 2, nfft3d = (nh + 1) * nfft * nfft;
   int k1, k2, k3;
   for (k3 = -nh + 1; k3 <= nh; ++k3)
      for (k2 = -nh + 1; k2 <= nh; ++k2)
         for (k1 = 0; k1 <= nh; ++k1) {
            int idx = k1 + wrap(k2, nfft) * (nh + 1) + wrap(k3, nfft) * (nh + 1) * nfft;
            g[idx]              = k1*unitg[0] + k2*unitg[3] + k3*unitg[6];
            g[idx + nfft3d]     = k1*unitg[1] + k2*unitg[4] + k3*unitg[7];
            g[idx + 2*nfft3d]   = k1*unitg[2] + k2*unitg[5] + k3*unitg[8];
         }
}

void oracle_get_masker(int nfft, double ggcut, const double *unitg, int *masker1)
{
   const int nh = nfft / 2, nfft3d = (nh + 1) * nfft * nfft;
   int k1, k2, k3;
   memset(masker1, 0, nfft3d * sizeof(int));
   /* strict bounds: kx < N/2, |ky|,|kz| < N/2 (get_masker.c:21-23) */
   for (k3 = -nh + 1; k3 < nh; ++k3)
      for (k2 = -nh + 1; k2 < nh; ++k2)
         for (k1 = 0; k1 < nh; ++k1) {
            double gx = k1*unitg[0] + k2*unitg[3] + k3*unitg[6];
            double gy = k1*unitg[1] + k2*unitg[4] + k3*unitg[7];
            double gz = k1*unitg[2] + k2*unitg[5] + k3*unitg[8];
            int idx = k1 + wrap(k2, nfft) * (nh + 1) + wrap(k3, nfft) * (nh + 1) * nfft;
            if (gx*gx + gy*gy + gz*gz < ggcut) masker1[idx] = 1;
         }
}

void oracle_setup(int icube, double unit, int nfft, int nion, const int *katm,
                  const double *zv, double *omega, double *unita,
                  double *unitg, double *G, int *masker, double *tg,
                  double *vc, double *vg, double *rcell, double *cewald,
                  double *rcut)
{
   const double pi = 4.0 * atan(1.0), fourpi = 4.0 * pi;
   const int nfft3d = (nfft / 2 + 1) * nfft * nfft;
   double alpha = 0.0, gx, gy, gz, gcut, w, rs, zz = 0.0, z = 0.0;
   int *m1, k, ii, i1, i2, i3, l;

   oracle_get_cube(icube, unit, omega, unita, unitg);
   if (icube == 1) alpha = 1.76011888;
   if (icube == 2) alpha = 1.791753;
   if (icube == 3) alpha = 1.791860;
   *rcut = unit / pi;

   oracle_get_g(nfft, unitg, G);
   gx = unitg[0] * (nfft / 2);
   gy = unitg[1] * (nfft / 2);
   gz = unitg[2] * (nfft / 2);
   gcut = gx*gx + gy*gy + gz*gz;

   w = 0.25 * (*rcut) * (*rcut);
   for (k = 1; k < nfft3d; ++k) {
      double gg = G[k]*G[k] + G[k+nfft3d]*G[k+nfft3d] + G[k+2*nfft3d]*G[k+2*nfft3d];
      tg[k] = 0.5 * gg;
      vc[k] = fourpi / gg;
      vg[k] = vc[k] * exp(-w * gg);
   }
   tg[0] = vc[0] = vg[0] = 0.0;

   m1 = (int *) malloc(nfft3d * sizeof(int));
   oracle_get_masker(nfft, gcut, unitg, m1);
   for (k = 0; k < nfft3d; ++k) {
      masker[2*k] = masker[2*k+1] = m1[k];
      tg[k] *= m1[k]; vc[k] *= m1[k]; vg[k] *= m1[k];
   }
   free(m1);

   /* Ewald constant, main.c:287-298 */
   rs = pow(3.00 * (*omega) / fourpi, 1.00 / 3.00);
   for (ii = 0; ii < nion; ++ii) {
      double q = zv[katm[ii]];
      zz += q * q;
      z += q;
   }
   *cewald = -0.50 * zz * (alpha / rs + oracle_gssum(nfft, vg) / (*omega))
             - 0.50 * (z * z - zz) * (*rcut) * (*rcut) * pi / (*omega);

   /* 27 image cells, main.c:301-313 (ncut = 1) */
   l = 0;
   for (i3 = -1; i3 <= 1; ++i3)
      for (i2 = -1; i2 <= 1; ++i2)
         for (i1 = -1; i1 <= 1; ++i1) {
            rcell[l]      = i1*unita[0] + i2*unita[3] + i3*unita[6];
            rcell[l + 27] = i1*unita[1] + i2*unita[4] + i3*unita[7];
            rcell[l + 54] = i1*unita[2] + i2*unita[5] + i3*unita[8];
            ++l;
         }
}

/* Simpson rule as written at ckbppv3/dkbppv3.c:310-328:
 * h/3 * (2*sum(even) + 4*sum(odd) - y0 - y_{n-1}) */
double oracle_simp(int n, const double *y, double h)
{
   double s2 = 0.0, s4 = 0.0;
   int i;
   for (i = 0; i < n; i += 2) s2 += y[i];
   for (i = 1; i < n; i += 2) s4 += y[i];
   return (2.0 * s2 + 4.0 * s4 - y[0] - y[n-1]) * (h / 3.0);
}

void oracle_kbpp(double zv, int lmax, int nrho, double drho, const double *rho,
                 double *vp, const double *wp, int nfft3d, const double *G,
                 double *vl, double *vnl, double *vnlnrm)
{
   const double pi = 4.0 * atan(1.0), twopi = 2.0 * pi, fourpi = 4.0 * pi;
   const double p0 = sqrt(fourpi), p1 = sqrt(3.0 * fourpi), p2 = sqrt(15.0 * fourpi);
   const int lmmax = lmax * lmax;
   double *f = (double *) malloc(nrho * sizeof(double));
   double *cs = (double *) malloc(nrho * sizeof(double));
   double *sn = (double *) malloc(nrho * sizeof(double));
   int i, l, k;

   /* nonlocal parts relative to the local channel l = lmax (:357-359) */
   for (l = 0; l < lmax; ++l)
      for (i = 0; i < nrho; ++i) vp[i + l*nrho] -= vp[i + lmax*nrho];

   /* normalisation <w|dv|w> per channel (:362-369) */
   for (l = 0; l < lmax; ++l) {
      double a;
      for (i = 0; i < nrho; ++i) f[i] = vp[i + l*nrho] * wp[i + l*nrho] * wp[i + l*nrho];
      a = oracle_simp(nrho, f, drho);
      for (i = l*l; i < (l+1)*(l+1); ++i) vnlnrm[i] = a;
   }

   /* G != 0 (:373-424) */
   for (k = 1; k < nfft3d; ++k) {
      double gx = G[k], gy = G[k + nfft3d], gz = G[k + 2*nfft3d];
      double q = sqrt(gx*gx + gy*gy + gz*gz);
      gx /= q; gy /= q; gz /= q;
      for (i = 0; i < nrho; ++i) { cs[i] = cos(q * rho[i]); sn[i] = sin(q * rho[i]); }
      if (lmax > 2) {
         double d;
         f[0] = 0.0;
         for (i = 1; i < nrho; ++i) {
            double a = 3.0 * (sn[i] / (q * rho[i]) - cs[i]) / (q * rho[i]) - sn[i];
            f[i] = a * wp[i + 2*nrho] * vp[i + 2*nrho];
         }
         d = p2 * oracle_simp(nrho, f, drho) / q;
         vnl[k + 4*(size_t)nfft3d] = d * (3.0*gz*gz - 1.0) / (2.0 * sqrt(3.0));
         vnl[k + 5*(size_t)nfft3d] = d * gx * gy;
         vnl[k + 6*(size_t)nfft3d] = d * gy * gz;
         vnl[k + 7*(size_t)nfft3d] = d * gz * gx;
         vnl[k + 8*(size_t)nfft3d] = d * (gx*gx - gy*gy) / 2.0;
      }
      if (lmax > 1) {
         double p;
         f[0] = 0.0;
         for (i = 1; i < nrho; ++i)
            f[i] = (sn[i] / (q * rho[i]) - cs[i]) * wp[i + nrho] * vp[i + nrho];
         p = p1 * oracle_simp(nrho, f, drho) / q;
         vnl[k + 1*(size_t)nfft3d] = p * gx;
         vnl[k + 2*(size_t)nfft3d] = p * gy;
         vnl[k + 3*(size_t)nfft3d] = p * gz;
      }
      if (lmax > 0) {
         for (i = 0; i < nrho; ++i) f[i] = sn[i] * wp[i] * vp[i];
         vnl[k] = p0 * oracle_simp(nrho, f, drho) / q;
      }
      for (i = 0; i < nrho; ++i) f[i] = rho[i] * vp[i + lmax*nrho] * sn[i];
      vl[k] = oracle_simp(nrho, f, drho) * fourpi / q
            - zv * fourpi * cs[nrho-1] / (q * q);
   }

   /* G = 0 (:428-439) */
   for (i = 0; i < nrho; ++i) f[i] = vp[i + lmax*nrho] * (rho[i] * rho[i]);
   vl[0] = fourpi * oracle_simp(nrho, f, drho) + twopi * zv * rho[nrho-1] * rho[nrho-1];
   for (i = 0; i < nrho; ++i) f[i] = rho[i] * wp[i] * vp[i];
   vnl[0] = p0 * oracle_simp(nrho, f, drho);
   for (l = 1; l < lmmax; ++l) vnl[(size_t)l * nfft3d] = 0.0;

   free(sn); free(cs); free(f);
}

/* cmakepsi/dmakepsi.c:86-133: orbital n is a single plane wave at
 * (k1,k2,k3) with coefficient 1 (l = 0), i/sqrt2 (l < 0) or 1/sqrt2 (l > 0);
 * on the kx = 0 plane its Hermitian partner gets the conjugate. */
void oracle_makepsi(int nfft, int ne, const int *spec, double *psi)
{
   const size_t n2 = (size_t)(nfft + 2) * nfft * nfft;
   const double sc = 1.0 / sqrt(2.0);
   int n;
   memset(psi, 0, (size_t)ne * n2 * sizeof(double));
   for (n = 0; n < ne; ++n) {
      int i = spec[4*n], j = spec[4*n+1], k = spec[4*n+2], l = spec[4*n+3];
      int k2 = (j + nfft) % nfft, k3 = (k + nfft) % nfft;
      double cr, ci;
      double *p = psi + (size_t)n * n2;
      size_t idx;
      if (l < 0)       { cr = 0.0; ci = sc; }
      else if (l == 0) { cr = 1.0; ci = 0.0; }
      else             { cr = sc;  ci = 0.0; }
      idx = 2 * ((size_t)i + (size_t)k2 * (nfft/2 + 1) + (size_t)k3 * (nfft/2 + 1) * nfft);
      p[idx] = cr; p[idx+1] = ci;
      if (i == 0) {
         int kk2 = (nfft - k2) % nfft, kk3 = (nfft - k3) % nfft;
         idx = 2 * ((size_t)kk2 * (nfft/2 + 1) + (size_t)kk3 * (nfft/2 + 1) * nfft);
         p[idx] = cr; p[idx+1] = -ci;
      }
   }
}
