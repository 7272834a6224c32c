/*
 * cpsd_caller.c -- a plain C caller of libcpsd_hip.so, the way the
 * reference's driver calls its SD step: it declares the prototypes itself
 * (as ccpsdv3/main.c:18-56 declares cpsd() and ccpsdv3/cpsd.c calls
 * cr_fft3b_/rc_fft3f_ without prototypes), reads the reference's Example0
 * inputs (CONTROL and IONIN formats of ccpsdv3/main.c:202-251), builds the
 * tables with libcpsd_host.so, and runs `outer` calls of cpsd().
 *
 * usage: cpsd_caller CONTROL IONIN PSPDIR PSI0.DAT outer
 * prints one line per call: "call k E0..E9 deltae deltac deltar", then
 * "fft roundtrip maxerr" for rc_fft3f_ followed by cr_fft3b_ of a real grid.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cpsd_host.h"

/* the drop-in symbols, declared here as a caller of the reference would */
void cpsd(const int inner, const int move, const int nion, const int ispin, const int *ne,
          const int nfft, const int nfft3d, const int n2ft3d, const double *unitg,
          const double omega, const double dt, const double fmass, const int *katm,
          const double *amass, const double *zv, const int nsh, const double *rcell,
          const double cewald, const double rcut, const int *masker, const double *G,
          const double *tg, const double *vc, const double *vg, int *lmmax, double *vnlnrm,
          double *vl, double *vnl, double *r1, double *r2, double *fion, double *c1,
          double *c2, double *cpsi, double *dn, double *hml, double *lmd, double *cpsdwork,
          double *E, double *deltae, double *deltac, double *deltar);
void cr_fft3b_(double *A, int *inc2c, int *inc3c, int *inc2r, int *inc3r, int *nx, int *ny, int *nz,
               double *tmp, int *ntmp);
void rc_fft3f_(double *A, int *inc2c, int *inc3c, int *inc2r, int *inc3r, int *nx, int *ny, int *nz,
               double *tmp, int *ntmp);

static void *xcalloc(size_t n, size_t sz)
{
   void *p = calloc(n ? n : 1, sz);
   if (!p) { fprintf(stderr, "out of memory\n"); exit(2); }
   return p;
}

int main(int argc, char **argv)
{
   if (argc != 6) {
      fprintf(stderr, "usage: %s CONTROL IONIN PSPDIR PSI0.DAT outer\n", argv[0]);
      return 2;
   }
   const int outer = atoi(argv[5]);
   int ispin, imove, icube, inner, outer_ctl;
   double unit, dt, fmass, tole, tolc, tolr;
   FILE *fp = fopen(argv[1], "r");
   if (!fp || fscanf(fp, "%d %d %d %lf %lf %lf %d %d %lf %lf %lf", &ispin, &imove, &icube, &unit, &dt, &fmass,
                     &inner, &outer_ctl, &tole, &tolc, &tolr) != 11) {
      fprintf(stderr, "cannot read CONTROL %s\n", argv[1]);
      return 2;
   }
   fclose(fp);

   /* IONIN: kinds 'name.vpp ' lmax, then katm x y z x' y' z' per ion */
   int nkatm, nion;
   char names[9][16];
   int lmaxk[9];
   fp = fopen(argv[2], "r");
   if (!fp || fscanf(fp, "%d", &nkatm) != 1 || nkatm < 1 || nkatm > 9) { fprintf(stderr, "bad IONIN\n"); return 2; }
   for (int k = 0; k < nkatm; ++k) {
      char q1[2], q2[2];   /* the quotes around "'C01.vpp '", read as ccpsdv3/main.c:232 reads them */
      if (fscanf(fp, "%1s%15s%1s %d", q1, names[k], q2, &lmaxk[k]) != 4) { fprintf(stderr, "bad kind\n"); return 2; }
      char *dot = strchr(names[k], '.');
      if (dot) *dot = 0;   /* "C01.vpp" -> the table C01.dat */
   }
   if (fscanf(fp, "%d", &nion) != 1 || nion < 1) { fprintf(stderr, "bad nion\n"); return 2; }
   int *katm = xcalloc(nion, sizeof(int));
   double *r2 = xcalloc(3 * (size_t)nion, sizeof(double)), *r1 = xcalloc(3 * (size_t)nion, sizeof(double));
   double *fion = xcalloc(3 * (size_t)nion, sizeof(double));
   for (int i = 0; i < nion; ++i) {
      double d[3];
      if (fscanf(fp, "%d %lf %lf %lf %lf %lf %lf", &katm[i], &r2[3 * i], &r2[3 * i + 1], &r2[3 * i + 2], &d[0], &d[1],
                 &d[2]) != 7) {
         fprintf(stderr, "bad ion %d\n", i);
         return 2;
      }
      katm[i] -= 1;
   }
   fclose(fp);

   const int nfft = 32, ne0 = 4;
   const int nfft3d = (nfft / 2 + 1) * nfft * nfft, n2ft3d = (nfft + 2) * nfft * nfft;
   double zv[9], amass[9], vnlnrm[81], rc[10], omega, unita[9], unitg[9], cewald, rcut, rcell[81];
   int lmmax[9];
   double *vl = xcalloc((size_t)nkatm * nfft3d, sizeof(double));
   double *vnl = xcalloc((size_t)nkatm * 9 * nfft3d, sizeof(double));
   double *G = xcalloc(3 * (size_t)nfft3d, sizeof(double));
   int *masker = xcalloc(n2ft3d, sizeof(int));
   double *tg = xcalloc(nfft3d, sizeof(double)), *vc = xcalloc(nfft3d, sizeof(double));
   double *vg = xcalloc(nfft3d, sizeof(double));

   /* the lattice first (for G), then each kind's KB tables on it */
   for (int k = 0; k < nkatm; ++k) zv[k] = 0.0;
   {
      const int cap = 4000;
      double *rho = xcalloc(cap, sizeof(double)), *vp = xcalloc(4 * cap, sizeof(double));
      double *wp = xcalloc(4 * cap, sizeof(double));
      for (int k = 0; k < nkatm; ++k) {
         char path[1024], atom[8];
         int lmax, nrho;
         double drho;
         snprintf(path, sizeof path, "%s/%s.dat", argv[3], names[k]);
         if (cpsdh_read_psp(path, atom, &zv[k], &amass[k], &lmax, rc, &nrho, &drho, rho, vp, wp, cap) != 0) {
            fprintf(stderr, "cannot read %s\n", path);
            return 2;
         }
         amass[k] *= 1822.89;
         if (k == 0 && cpsdh_setup(icube, unit, nfft, nion, katm, zv, &omega, unita, unitg, G, masker, tg, vc, vg,
                                   rcell, &cewald, &rcut) != 0) {
            fprintf(stderr, "setup failed\n");
            return 2;
         }
         if (cpsdh_kbpp(zv[k], lmaxk[k], nrho, drho, rho, vp, wp, nfft3d, G, vl + (size_t)k * nfft3d,
                        vnl + (size_t)k * 9 * nfft3d, vnlnrm + 9 * k) != 0) {
            fprintf(stderr, "kbpp failed\n");
            return 2;
         }
         lmmax[k] = lmaxk[k] * lmaxk[k];
      }
      free(rho); free(vp); free(wp);
      /* the Ewald constant depends on every kind's valence */
      if (cpsdh_setup(icube, unit, nfft, nion, katm, zv, &omega, unita, unitg, G, masker, tg, vc, vg, rcell,
                      &cewald, &rcut) != 0)
         return 2;
   }

   /* plane-wave start from the first ne0 rows of PSI0.DAT */
   int spec[4 * 64];
   fp = fopen(argv[4], "r");
   for (int i = 0; i < ne0; ++i)
      if (!fp || fscanf(fp, "%d %d %d %d", &spec[4 * i], &spec[4 * i + 1], &spec[4 * i + 2], &spec[4 * i + 3]) != 4) {
         fprintf(stderr, "bad PSI0\n");
         return 2;
      }
   fclose(fp);
   const size_t nn = (size_t)ne0 * n2ft3d;
   double *c1 = xcalloc(nn, sizeof(double)), *c2 = xcalloc(nn, sizeof(double)), *cpsi = xcalloc(nn, sizeof(double));
   double *dn = xcalloc(n2ft3d, sizeof(double));
   double *hml = xcalloc(ne0 * ne0, sizeof(double)), *lmd = xcalloc(ne0 * ne0, sizeof(double));
   double *work = xcalloc(7 * (size_t)n2ft3d + (size_t)nion * (5 * nfft + 3) + 7 * ne0 * ne0, sizeof(double));
   if (cpsdh_makepsi(nfft, ne0, spec, c2) != 0) return 2;

   int ne[2] = {ne0, 0};
   double E[10] = {0}, deltae = 0, deltac = 0, deltar = 0;
   for (int call = 0; call < outer; ++call) {
      cpsd(inner, imove, nion, ispin, ne, nfft, nfft3d, n2ft3d, unitg, omega, dt, fmass, katm, amass, zv, 27,
           rcell, cewald, rcut, masker, G, tg, vc, vg, lmmax, vnlnrm, vl, vnl, r1, r2, fion, c1, c2, cpsi, dn, hml,
           lmd, work, E, &deltae, &deltac, &deltar);
      printf("call %d", call + 1);
      for (int i = 0; i < 10; ++i) printf(" %.15e", E[i]);
      printf(" %.15e %.15e %.15e\n", deltae, deltac, deltar);
   }

   /* rc_fft3f_ then cr_fft3b_ of a real field returns N^3 times the field */
   {
      int inc2c = nfft / 2 + 1, inc3c = (nfft / 2 + 1) * nfft, inc2r = nfft + 2, inc3r = (nfft + 2) * nfft;
      int nx = nfft, ny = nfft, nz = nfft, ntmp = 2 * (nfft * nfft + 10) + 40 * nfft;
      double *a = xcalloc(n2ft3d, sizeof(double)), *b = xcalloc(n2ft3d, sizeof(double));
      double *tmp = xcalloc(ntmp, sizeof(double));
      unsigned long long s = 12345;
      for (int z = 0; z < nfft; ++z)
         for (int y = 0; y < nfft; ++y)
            for (int x = 0; x < nfft; ++x) {
               s = s * 6364136223846793005ull + 1442695040888963407ull;
               a[((size_t)z * nfft + y) * (nfft + 2) + x] = (double)(s >> 11) / 9007199254740992.0 - 0.5;
            }
      memcpy(b, a, n2ft3d * sizeof(double));
      rc_fft3f_(b, &inc2c, &inc3c, &inc2r, &inc3r, &nx, &ny, &nz, tmp, &ntmp);
      cr_fft3b_(b, &inc2c, &inc3c, &inc2r, &inc3r, &nx, &ny, &nz, tmp, &ntmp);
      double err = 0.0;
      const double scale = 1.0 / ((double)nfft * nfft * nfft);
      for (size_t k = 0; k < (size_t)n2ft3d; ++k) {
         const double d = fabs(b[k] * scale - a[k]);
         if (d > err) err = d;
      }
      printf("fft roundtrip %.3e\n", err);
      free(a); free(b); free(tmp);
   }
   free(katm); free(r1); free(r2); free(fion);
   free(vl); free(vnl); free(G); free(masker); free(tg); free(vc); free(vg);
   free(c1); free(c2); free(cpsi); free(dn); free(hml); free(lmd); free(work);
   return 0;
}
