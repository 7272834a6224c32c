/*
 * cpsd_host.c -- host-side inputs of the SD step (see include/cpsd_host.h).
 *
 * Reference counterparts (paths relative to Cpsd-mpi-bamboo/):
 *   lattice            ccpsdv3/get_cube.c:29-85
 *   G vectors          ccpsdv3/get_g.c:12-35
 *   cutoff mask        ccpsdv3/get_masker.c:12-37 (strict |G|^2 < |b1 N/2|^2,
 *                      kx < N/2, |ky|,|kz| < N/2)
 *   tables, Ewald      ccpsdv3/main.c:161-198, 287-313
 *   KB transform       ckbppv3/dkbppv3.c:310-444
 *   plane-wave start   cmakepsi/dmakepsi.c:86-133
 * The KB transform is the expensive part (nfft3d x nrho sin/cos); it runs
 * over G vectors in parallel with OpenMP.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/cpsd_host.h"

static int wrapk(int k, int n) { return k < 0 ? k + n : k; }

static void lattice(int icube, double unit, double *omega, double *a, double *b)
{
   double base[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
   const double twopi = 8.0 * atan(1.0);
   double vol;
   int i;
   if (icube == 2) {
      const double f[9] = {0.5, 0.5, 0.0, 0.5, 0.0, 0.5, 0.0, 0.5, 0.5};
      memcpy(base, f, sizeof f);
   } else if (icube == 3) {
      const double f[9] = {0.5, 0.5, 0.5, -0.5, 0.5, 0.5, 0.5, 0.5, -0.5};
      memcpy(base, f, sizeof f);
   }
   for (i = 0; i < 9; ++i) a[i] = unit * base[i];
   /* reciprocal vectors b_d = 2 pi (a_{d+1} x a_{d+2}) / (a_1 . a_2 x a_3) */
   for (i = 0; i < 3; ++i) {
      const double *u = a + 3 * ((i + 1) % 3), *v = a + 3 * ((i + 2) % 3);
      b[3 * i + 0] = u[1] * v[2] - u[2] * v[1];
      b[3 * i + 1] = u[2] * v[0] - u[0] * v[2];
      b[3 * i + 2] = u[0] * v[1] - u[1] * v[0];
   }
   vol = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
   for (i = 0; i < 9; ++i) b[i] *= twopi / vol;
   *omega = fabs(vol);
}

int cpsdh_setup(int icube, double unit, int nfft, int nion, const int *katm, const double *zv,
                double *omega, double *unita, double *unitg, double *G, int *masker, double *tg,
                double *vc, double *vg, double *rcell, double *cewald, double *rcut)
{
   const double pi = 4.0 * atan(1.0);
   const int n = nfft, nh = n / 2, n3d = (nh + 1) * n * n;
   double madelung, gcut, w, rs, zsum = 0.0, z2sum = 0.0, vgsum_all = 0.0, vgsum_x0 = 0.0;
   int k1, k2, k3, i, l;

   if (icube < 1 || icube > 3 || n < 2 || (n & 1)) return 1;
   lattice(icube, unit, omega, unita, unitg);
   madelung = icube == 1 ? 1.76011888 : (icube == 2 ? 1.791753 : 1.791860);
   *rcut = unit / pi;
   gcut = 0.0;
   for (i = 0; i < 3; ++i) gcut += (unitg[i] * nh) * (unitg[i] * nh);
   w = 0.25 * (*rcut) * (*rcut);

   for (k3 = -nh + 1; k3 <= nh; ++k3)
      for (k2 = -nh + 1; k2 <= nh; ++k2)
         for (k1 = 0; k1 <= nh; ++k1) {
            const int g = k1 + (nh + 1) * (wrapk(k2, n) + n * wrapk(k3, n));
            const double gx = k1 * unitg[0] + k2 * unitg[3] + k3 * unitg[6];
            const double gy = k1 * unitg[1] + k2 * unitg[4] + k3 * unitg[7];
            const double gz = k1 * unitg[2] + k2 * unitg[5] + k3 * unitg[8];
            const double gg = gx * gx + gy * gy + gz * gz;
            const int inside = (k1 < nh && k2 < nh && k3 < nh && gg < gcut);
            G[g] = gx; G[g + n3d] = gy; G[g + 2 * n3d] = gz;
            masker[2 * g] = masker[2 * g + 1] = inside;
            if (g == 0) {
               tg[0] = vc[0] = vg[0] = 0.0;
            } else {
               tg[g] = 0.5 * gg * inside;
               vc[g] = (4.0 * pi / gg) * inside;
               vg[g] = (4.0 * pi / gg) * exp(-w * gg) * inside;
            }
         }
   /* gssum(vg): 2 * sum(all) - sum(kx = 0 plane), summed in storage order */
   for (i = 0; i < n3d; ++i) vgsum_all += vg[i];
   for (i = 0; i < n * n; ++i) vgsum_x0 += vg[i * (nh + 1)];
   for (i = 0; i < nion; ++i) {
      zsum += zv[katm[i]];
      z2sum += zv[katm[i]] * zv[katm[i]];
   }
   rs = pow(3.00 * (*omega) / (4.0 * pi), 1.00 / 3.00);
   *cewald = -0.50 * z2sum * (madelung / rs + (2.0 * vgsum_all - vgsum_x0) / (*omega))
             - 0.50 * (zsum * zsum - z2sum) * (*rcut) * (*rcut) * pi / (*omega);
   l = 0;
   for (k3 = -1; k3 <= 1; ++k3)
      for (k2 = -1; k2 <= 1; ++k2)
         for (k1 = -1; k1 <= 1; ++k1, ++l)
            for (i = 0; i < 3; ++i)
               rcell[l + 27 * i] = k1 * unita[i] + k2 * unita[3 + i] + k3 * unita[6 + i];
   return 0;
}

/* ------------------------------------------------------------- PSP */

static int next_line(FILE *f, char *buf, int len)
{
   return fgets(buf, len, f) != NULL;
}

int cpsdh_read_psp(const char *path, char *atom, double *zv, double *amass, int *lmax, double *rc,
                   int *nrho, double *drho, double *rho, double *vp, double *wp, int nrho_cap)
{
   char line[4096];
   FILE *f = fopen(path, "r");
   int l, i, lm;
   if (!f) return 1;
   if (!next_line(f, line, sizeof line) || sscanf(line, "%2s", atom) != 1) { fclose(f); return 2; }
   /* line 2: zv amass lmax [extra integer, ignored] */
   if (!next_line(f, line, sizeof line) || sscanf(line, "%lf %lf %d", zv, amass, lmax) != 3) {
      fclose(f);
      return 3;
   }
   lm = *lmax;
   if (lm < 0 || lm > 3) { fclose(f); return 4; }
   for (l = 0; l <= lm; ++l)
      if (fscanf(f, "%lf", &rc[l]) != 1) { fclose(f); return 5; }
   if (fscanf(f, "%d %lf", nrho, drho) != 2 || *nrho < 2 || *nrho > nrho_cap) { fclose(f); return 6; }
   for (i = 0; i < *nrho; ++i) {
      if (fscanf(f, "%lf", &rho[i]) != 1) { fclose(f); return 7; }
      for (l = 0; l <= lm; ++l)
         if (fscanf(f, "%lf", &vp[i + l * (*nrho)]) != 1) { fclose(f); return 7; }
   }
   for (i = 0; i < *nrho; ++i) {
      double r;
      if (fscanf(f, "%lf", &r) != 1) { fclose(f); return 8; }
      for (l = 0; l <= lm; ++l)
         if (fscanf(f, "%lf", &wp[i + l * (*nrho)]) != 1) { fclose(f); return 8; }
   }
   fclose(f);
   return 0;
}

/* Weight of point j of an n-point composite Simpson sum over spacing h, as
 * ckbppv3/dkbppv3.c:310-328 forms it, h/3 (2 sum(even) + 4 sum(odd) - y_0 -
 * y_{n-1}): interior even points 2, odd points 4, the ends one less. */
static double simpson_weight(int j, int n, double h)
{
   double w = (j % 2 == 0) ? 2.0 : 4.0;
   if (j == 0) w -= 1.0;
   if (j == n - 1) w -= 1.0;
   return w * h / 3.0;
}

struct qkey {
   double q2;
   int k;
};

static int qkey_cmp(const void *a, const void *b)
{
   const double x = ((const struct qkey *)a)->q2, y = ((const struct qkey *)b)->q2;
   return x < y ? -1 : (x > y ? 1 : 0);
}

/* The Kleinman-Bylander transform of ckbppv3/dkbppv3.c:335-444.
 *
 * With dv_l = V_l - V_lmax for l < lmax (channel lmax is the local part) and
 * the Simpson sums S[.] over the radial table, the transform of G != 0
 * (q = |G|, g = G/q, x = q r) is
 *   s:     vnl_0 = sqrt(4 pi)  S[sin x  wp_0 dv_0] / q
 *   p:     vnl_p = sqrt(12 pi) S[(sin x/x - cos x) wp_1 dv_1] / q  (g_x, g_y, g_z)
 *   d:     vnl_d = sqrt(60 pi) S[(3 (sin x/x - cos x)/x - sin x) wp_2 dv_2] / q
 *          times (3 g_z^2 - 1)/(2 sqrt 3), g_x g_y, g_y g_z, g_z g_x, (g_x^2 - g_y^2)/2
 *   local: vl = 4 pi S[r sin x dv_lmax] / q - 4 pi zv cos(q r_last) / q^2
 * (the p and d sums leave out r_0, whose integrand the reference sets to 0).
 * The four radial sums depend on q only: everything that does not depend on
 * q (Simpson weight times wp_l dv_l, or times r dv_lmax) is folded into one
 * table, the G vectors are grouped by |G|^2 (the lattice repeats each |G|
 * many times), and each distinct q costs one pass over the table with one
 * sin and cos per point.  G = 0 and the normalisations
 * vnlnrm_l = S[dv_l wp_l^2] are direct Simpson sums. */
int cpsdh_kbpp(double zv, int lmax, int nrho, double drho, const double *rho, const double *vp_in,
               const double *wp, int nfft3d, const double *G, double *vl, double *vnl,
               double *vnlnrm)
{
   const double pi = 4.0 * atan(1.0), fourpi = 4.0 * pi;
   const double c0 = sqrt(fourpi), c1 = sqrt(3.0 * fourpi), c2 = sqrt(15.0 * fourpi);
   const int lmmax = lmax * lmax;
   double *tab = (double *)malloc(4 * (size_t)nrho * sizeof(double));   /* s, p, d, local */
   struct qkey *keys = (struct qkey *)malloc((size_t)(nfft3d > 1 ? nfft3d - 1 : 1) * sizeof(struct qkey));
   int *gstart = (int *)malloc((size_t)nfft3d * sizeof(int));
   int j, l, ngroups = 0;
   if (!tab || !keys || !gstart || lmax < 0 || lmax > 3 || nrho < 2) {
      free(tab); free(keys); free(gstart);
      return 1;
   }
   for (j = 0; j < nrho; ++j) {
      const double w = simpson_weight(j, nrho, drho);
      const double vloc = vp_in[j + (size_t)lmax * nrho];
      for (l = 0; l < 3; ++l)
         tab[(size_t)l * nrho + j] = (l < lmax && (l == 0 || j > 0))
                                        ? w * wp[j + (size_t)l * nrho] * (vp_in[j + (size_t)l * nrho] - vloc)
                                        : 0.0;
      tab[3 * (size_t)nrho + j] = w * rho[j] * vloc;
   }
   for (l = 0; l < 9; ++l) vnlnrm[l] = 0.0;
   for (l = 0; l < lmax; ++l) {
      double a = 0.0;
      for (j = 0; j < nrho; ++j) {
         const double dvl = vp_in[j + (size_t)l * nrho] - vp_in[j + (size_t)lmax * nrho];
         a += simpson_weight(j, nrho, drho) * dvl * wp[j + (size_t)l * nrho] * wp[j + (size_t)l * nrho];
      }
      for (j = l * l; j < (l + 1) * (l + 1); ++j) vnlnrm[j] = a;
   }

   /* G != 0 grouped by |G|^2 (equal to rounding) */
   for (j = 1; j < nfft3d; ++j) {
      const double gx = G[j], gy = G[j + nfft3d], gz = G[j + 2 * (size_t)nfft3d];
      keys[j - 1].q2 = gx * gx + gy * gy + gz * gz;
      keys[j - 1].k = j;
   }
   qsort(keys, (size_t)(nfft3d - 1), sizeof(struct qkey), qkey_cmp);
   for (j = 0; j < nfft3d - 1; ++j)
      if (j == 0 || keys[j].q2 > keys[j - 1].q2 * (1.0 + 1e-14)) gstart[ngroups++] = j;
   gstart[ngroups] = nfft3d - 1;

#pragma omp parallel for schedule(dynamic, 16)
   for (int gi = 0; gi < ngroups; ++gi) {
      const double q = sqrt(keys[gstart[gi]].q2);
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, sl = 0.0;
      for (int jj = 0; jj < nrho; ++jj) {
         const double x = q * rho[jj], sn = sin(x);
         s0 += sn * tab[jj];
         sl += sn * tab[3 * (size_t)nrho + jj];
         if (jj > 0 && lmax > 1) {
            const double b1 = sn / x - cos(x);
            s1 += b1 * tab[(size_t)nrho + jj];
            s2 += (3.0 * b1 / x - sn) * tab[2 * (size_t)nrho + jj];
         }
      }
      const double vs = c0 * s0 / q, vp = c1 * s1 / q, vd = c2 * s2 / q;
      const double vloc = fourpi * sl / q - zv * fourpi * cos(q * rho[nrho - 1]) / (q * q);
      for (int e = gstart[gi]; e < gstart[gi + 1]; ++e) {
         const int k = keys[e].k;
         const double qk = sqrt(keys[e].q2);
         const double gx = G[k] / qk, gy = G[k + nfft3d] / qk, gz = G[k + 2 * (size_t)nfft3d] / qk;
         vl[k] = vloc;
         if (lmax > 0) vnl[k] = vs;
         if (lmax > 1) {
            vnl[k + 1 * (size_t)nfft3d] = vp * gx;
            vnl[k + 2 * (size_t)nfft3d] = vp * gy;
            vnl[k + 3 * (size_t)nfft3d] = vp * gz;
         }
         if (lmax > 2) {
            vnl[k + 4 * (size_t)nfft3d] = vd * (3.0 * gz * gz - 1.0) / (2.0 * sqrt(3.0));
            vnl[k + 5 * (size_t)nfft3d] = vd * gx * gy;
            vnl[k + 6 * (size_t)nfft3d] = vd * gy * gz;
            vnl[k + 7 * (size_t)nfft3d] = vd * gz * gx;
            vnl[k + 8 * (size_t)nfft3d] = vd * (gx * gx - gy * gy) / 2.0;
         }
      }
   }

   /* G = 0: vl = 4 pi S[r^2 dv_lmax] + 2 pi zv r_last^2, vnl_s = sqrt(4 pi) S[r wp_0 dv_0].
    * The reference writes vnl_s(G=0) for every lmax (ckbppv3/dkbppv3.c:436-437);
    * with lmax = 0 its channel 0 is the local one, not a difference. */
   {
      double a = 0.0, b = 0.0;
      for (j = 0; j < nrho; ++j) {
         a += tab[3 * (size_t)nrho + j] * rho[j];
         b += lmax > 0 ? tab[j] * rho[j] : simpson_weight(j, nrho, drho) * rho[j] * wp[j] * vp_in[j];
      }
      vl[0] = fourpi * a + 2.0 * pi * zv * rho[nrho - 1] * rho[nrho - 1];
      vnl[0] = c0 * b;
      for (l = 1; l < lmmax; ++l) vnl[(size_t)l * nfft3d] = 0.0;
   }
   free(tab);
   free(keys);
   free(gstart);
   return 0;
}

/* .vpp files, the output of ckbppv3/dkbppv3.c:138-150 and the input of
 * ccpsdv3/main.c:266-283: int icube, int nfft, double unit, char atom[2],
 * double amass, double zv, int lmax, double rc[lmax+1],
 * double vnlnrm[lmax^2], double vl[nfft3d], double vnl[lmax^2 nfft3d],
 * native byte order, no padding (amass is the table's value; the
 * reference's reader scales it by 1822.89 afterwards). */
int cpsdh_write_vpp(const char *path, int icube, int nfft, double unit, const char *atom, double amass, double zv,
                    int lmax, const double *rc, const double *vnlnrm, const double *vl, const double *vnl)
{
   const size_t nfft3d = (size_t)(nfft / 2 + 1) * nfft * nfft, lmmax = (size_t)lmax * lmax;
   char a2[2] = {0, 0};
   FILE *f;
   int ok;
   if (lmax < 0 || lmax > 3 || nfft < 2) return 1;
   a2[0] = atom[0];
   if (atom[0]) a2[1] = atom[1];
   f = fopen(path, "wb");
   if (!f) return 2;
   ok = fwrite(&icube, sizeof(int), 1, f) == 1 && fwrite(&nfft, sizeof(int), 1, f) == 1 &&
        fwrite(&unit, sizeof(double), 1, f) == 1 && fwrite(a2, 1, 2, f) == 2 &&
        fwrite(&amass, sizeof(double), 1, f) == 1 && fwrite(&zv, sizeof(double), 1, f) == 1 &&
        fwrite(&lmax, sizeof(int), 1, f) == 1 && fwrite(rc, sizeof(double), (size_t)lmax + 1, f) == (size_t)lmax + 1 &&
        fwrite(vnlnrm, sizeof(double), lmmax, f) == lmmax && fwrite(vl, sizeof(double), nfft3d, f) == nfft3d &&
        fwrite(vnl, sizeof(double), lmmax * nfft3d, f) == lmmax * nfft3d;
   if (fclose(f) != 0) ok = 0;
   return ok ? 0 : 3;
}

int cpsdh_read_vpp(const char *path, int *icube, int *nfft, double *unit, char *atom, double *amass, double *zv,
                   int *lmax, double *rc, double *vnlnrm, double *vl, double *vnl, long cap_nfft3d)
{
   FILE *f = fopen(path, "rb");
   size_t nfft3d, lmmax;
   int ok;
   if (!f) return 2;
   atom[2] = 0;
   ok = fread(icube, sizeof(int), 1, f) == 1 && fread(nfft, sizeof(int), 1, f) == 1 &&
        fread(unit, sizeof(double), 1, f) == 1 && fread(atom, 1, 2, f) == 2 &&
        fread(amass, sizeof(double), 1, f) == 1 && fread(zv, sizeof(double), 1, f) == 1 &&
        fread(lmax, sizeof(int), 1, f) == 1;
   if (!ok || *lmax < 0 || *lmax > 3 || *nfft < 2 || *nfft > 4096) { fclose(f); return 3; }
   nfft3d = (size_t)(*nfft / 2 + 1) * *nfft * *nfft;
   lmmax = (size_t)*lmax * *lmax;
   if ((long)nfft3d > cap_nfft3d) { fclose(f); return 4; }
   ok = fread(rc, sizeof(double), (size_t)*lmax + 1, f) == (size_t)*lmax + 1 &&
        fread(vnlnrm, sizeof(double), lmmax, f) == lmmax && fread(vl, sizeof(double), nfft3d, f) == nfft3d &&
        fread(vnl, sizeof(double), lmmax * nfft3d, f) == lmmax * nfft3d;
   /* nothing may follow */
   if (ok && fgetc(f) != EOF) ok = 0;
   fclose(f);
   return ok ? 0 : 3;
}

int cpsdh_makepsi(int nfft, int ne, const int *spec, double *psi)
{
   const int nh1 = nfft / 2 + 1;
   const size_t n2 = (size_t)(nfft + 2) * nfft * nfft;
   const double h = 1.0 / sqrt(2.0);
   int n;
   memset(psi, 0, (size_t)ne * n2 * sizeof(double));
   for (n = 0; n < ne; ++n) {
      const int i = spec[4 * n], j = (spec[4 * n + 1] + nfft) % nfft, k = (spec[4 * n + 2] + nfft) % nfft;
      const int l = spec[4 * n + 3];
      const double re = l < 0 ? 0.0 : (l == 0 ? 1.0 : h), im = l < 0 ? h : 0.0;
      double *p = psi + (size_t)n * n2;
      size_t g = (size_t)i + (size_t)nh1 * (j + (size_t)nfft * k);
      p[2 * g] = re; p[2 * g + 1] = im;
      if (i == 0) {
         g = (size_t)nh1 * ((nfft - j) % nfft + (size_t)nfft * ((nfft - k) % nfft));
         p[2 * g] = re; p[2 * g + 1] = -im;
      }
   }
   return 0;
}

double cpsdh_gcdotc(int nfft, const double *a, const double *b)
{
   const size_t n2 = (size_t)(nfft + 2) * nfft * nfft;
   double all = 0.0, x0 = 0.0;
   size_t i;
   for (i = 0; i < n2; ++i) all += a[i] * b[i];
   for (i = 0; i < (size_t)nfft * nfft; ++i) {
      const size_t o = i * (nfft + 2);
      x0 += a[o] * b[o] + a[o + 1] * b[o + 1];
   }
   return 2.0 * all - x0;
}

/* ------------------------------------------------ random start */

static uint64_t splitmix64(uint64_t *s)
{
   uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
   z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
   z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
   return z ^ (z >> 31);
}

static double uniform01(uint64_t *s)
{
   return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

int cpsdh_random_orbitals(int nfft, int ne, const int *masker, const double *G, unsigned long long seed,
                          double decay, double *psi)
{
   const int nh1 = nfft / 2 + 1;
   const size_t n3d = (size_t)nh1 * nfft * nfft, n2 = 2 * n3d;
   size_t ng = 0, g, p;
   size_t *gl;
   double *wt, *cp;
   uint64_t st = (uint64_t)seed;
   int i, j;
   for (g = 0; g < n3d; ++g) ng += masker[2 * g] != 0;
   gl = (size_t *)malloc(ng * sizeof(size_t));
   wt = (double *)malloc(ng * sizeof(double));
   cp = (double *)malloc((size_t)ne * 2 * ng * sizeof(double));
   if (!gl || !wt || !cp) { free(gl); free(wt); free(cp); return 1; }
   p = 0;
   for (g = 0; g < n3d; ++g)
      if (masker[2 * g]) {
         gl[p] = g;
         wt[p] = (g % nh1 == 0) ? 1.0 : 2.0;
         ++p;
      }
   /* Gaussian coefficients, one Box-Muller pair per (orbital, G) */
   for (i = 0; i < ne; ++i)
      for (p = 0; p < ng; ++p) {
         const double u1 = uniform01(&st), u2 = uniform01(&st);
         const double r = sqrt(-2.0 * log(u1));
         double damp = 1.0;
         if (decay > 0.0) {
            const size_t gg = gl[p];
            const double gx = G[gg], gy = G[gg + n3d], gz = G[gg + 2 * n3d];
            damp = exp(-0.5 * decay * (gx * gx + gy * gy + gz * gz));
         }
         cp[(size_t)i * 2 * ng + 2 * p] = damp * r * cos(2.0 * M_PI * u2);
         cp[(size_t)i * 2 * ng + 2 * p + 1] = damp * r * sin(2.0 * M_PI * u2);
      }
   /* Hermitian kx = 0 plane: c(-ky,-kz) = conj c(ky,kz); Im c(0) = 0 */
   {
      size_t *pos = (size_t *)malloc(n3d * sizeof(size_t));
      for (g = 0; g < n3d; ++g) pos[g] = (size_t)-1;
      for (p = 0; p < ng; ++p) pos[gl[p]] = p;
      for (p = 0; p < ng; ++p) {
         const size_t gg = gl[p];
         if (gg % nh1) continue;
         {
            const int ky = (int)((gg / nh1) % nfft), kz = (int)(gg / (nh1 * (size_t)nfft));
            const size_t mate = (size_t)nh1 * ((nfft - ky) % nfft + (size_t)nfft * ((nfft - kz) % nfft));
            const size_t q = pos[mate];
            if (q == (size_t)-1) continue;
            for (i = 0; i < ne; ++i) {
               double *c = cp + (size_t)i * 2 * ng;
               if (q == p) c[2 * p + 1] = 0.0;
               else if (q > p) { c[2 * q] = c[2 * p]; c[2 * q + 1] = -c[2 * p + 1]; }
            }
         }
      }
      free(pos);
   }
   /* modified Gram-Schmidt under sum_p w_p (re re' + im im'); the dot
    * products are summed over 64 fixed chunks, then the chunks in order, so
    * the result does not depend on the thread count */
   {
      enum { NCH = 64 };
      double part[NCH];
      const size_t chunk = (ng + NCH - 1) / NCH;
      for (i = 0; i < ne; ++i) {
         double *ci = cp + (size_t)i * 2 * ng;
         double nrm = 0.0;
         int b;
         for (j = 0; j <= i; ++j) {
            const double *cj = cp + (size_t)j * 2 * ng;
            double d = 0.0;
            long q;
#pragma omp parallel for
            for (b = 0; b < NCH; ++b) {
               size_t q0 = (size_t)b * chunk, q1 = q0 + chunk, r;
               double acc = 0.0;
               if (q1 > ng) q1 = ng;
               for (r = q0; r < q1; ++r) acc += wt[r] * (cj[2 * r] * ci[2 * r] + cj[2 * r + 1] * ci[2 * r + 1]);
               part[b] = acc;
            }
            for (b = 0; b < NCH; ++b) d += part[b];
            if (j == i) { nrm = d; break; }
#pragma omp parallel for
            for (q = 0; q < (long)ng; ++q) { ci[2 * q] -= d * cj[2 * q]; ci[2 * q + 1] -= d * cj[2 * q + 1]; }
         }
         nrm = 1.0 / sqrt(nrm);
         for (p = 0; p < 2 * ng; ++p) ci[p] *= nrm;
      }
   }
   memset(psi, 0, (size_t)ne * n2 * sizeof(double));
   for (i = 0; i < ne; ++i)
      for (p = 0; p < ng; ++p) {
         psi[(size_t)i * n2 + 2 * gl[p]] = cp[(size_t)i * 2 * ng + 2 * p];
         psi[(size_t)i * n2 + 2 * gl[p] + 1] = cp[(size_t)i * 2 * ng + 2 * p + 1];
      }
   free(gl); free(wt); free(cp);
   return 0;
}

/* ------------------------------------------------ restart files */

/* ELCIN / ELCOUT: the fields one after the other, as ccpsdv3/main.c:115-140
 * reads and :539-547 writes them with fread/fwrite */
int cpsdh_write_elc(const char *path, int icube, int nfft, double unit, int ispin, const int *ne,
                    const double *c)
{
   FILE *f = fopen(path, "wb");
   if (!f) return -1;
   const size_t n = (size_t)(ne[0] + ne[1]) * (size_t)(nfft + 2) * nfft * nfft;
   int ok = fwrite(&icube, sizeof(int), 1, f) == 1 && fwrite(&nfft, sizeof(int), 1, f) == 1 &&
            fwrite(&unit, sizeof(double), 1, f) == 1 && fwrite(&ispin, sizeof(int), 1, f) == 1 &&
            fwrite(ne, sizeof(int), 2, f) == 2 && fwrite(c, sizeof(double), n, f) == n;
   if (fclose(f) != 0) ok = 0;
   return ok ? 0 : -1;
}

int cpsdh_read_elc(const char *path, int *icube, int *nfft, double *unit, int *ispin, int *ne,
                   double *c, long cap)
{
   FILE *f = fopen(path, "rb");
   if (!f) return -1;
   int ok = fread(icube, sizeof(int), 1, f) == 1 && fread(nfft, sizeof(int), 1, f) == 1 &&
            fread(unit, sizeof(double), 1, f) == 1 && fread(ispin, sizeof(int), 1, f) == 1 &&
            fread(ne, sizeof(int), 2, f) == 2;
   if (ok && (*nfft < 1 || ne[0] < 0 || ne[1] < 0)) ok = 0;
   if (ok && c) {
      const size_t n = (size_t)(ne[0] + ne[1]) * (size_t)(*nfft + 2) * (*nfft) * (*nfft);
      ok = (long)n <= cap && fread(c, sizeof(double), n, f) == n;
   }
   fclose(f);
   return ok ? 0 : -1;
}

/* IONOUT, ccpsdv3/main.c:551-570.  The reference zeroes r1 just before
 * writing it (:552-553), so the second triple of every ion is 0. */
int cpsdh_write_ionout(const char *path, int nkatm, const char *const *fname, const int *lmax,
                       int nion, const int *katm, const double *r2)
{
   FILE *f = fopen(path, "w");
   if (!f) return -1;
   fprintf(f, "%d\n", nkatm);
   for (int ia = 0; ia < nkatm; ++ia) fprintf(f, "'%s ' %d\n", fname[ia], lmax[ia]);
   fprintf(f, "%d\n", nion);
   for (int ii = 0; ii < nion; ++ii) {
      fprintf(f, "%d ", katm[ii] + 1);
      for (int d = 0; d < 3; ++d) fprintf(f, "%lf ", r2[d + 3 * ii]);
      for (int d = 0; d < 3; ++d) fprintf(f, "%lf ", 0.0);
      fprintf(f, "\n");
   }
   return fclose(f) == 0 ? 0 : -1;
}
