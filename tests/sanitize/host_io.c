/*
 * host_io.c -- test infrastructure only: drives the file formats and the
 * orbital builders of libcpsd_host (include/cpsd_host.h) under the
 * sanitizers: a .vpp file written and read back (ckbppv3/dkbppv3.c:138-150,
 * ccpsdv3/main.c:266-283), an ELCIN/ELCOUT file written and read back
 * (ccpsdv3/main.c:115-140, :539-547), an IONOUT file (:551-570), seeded
 * orthonormal orbitals, and the readers' refusals of short and oversized
 * files.  Prints "host io ok" and returns 0 when every round trip is exact.
 *
 * usage: host_io WORKDIR PSPFILE
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cpsd_host.h"

#define FAIL(...) do { fprintf(stderr, __VA_ARGS__); fputc('\n', stderr); return 1; } while (0)

int main(int argc, char **argv)
{
   if (argc != 3) FAIL("usage: %s WORKDIR PSPFILE", argv[0]);
   const int nfft = 16, nfft3d = (nfft / 2 + 1) * nfft * nfft, n2ft3d = (nfft + 2) * nfft * nfft;
   char path[1024];
   const int cap = 4000;
   double *rho = calloc(cap, sizeof(double)), *vp = calloc(4 * cap, sizeof(double));
   double *wp = calloc(4 * cap, sizeof(double));
   char atom[8];
   double zv, amass, drho, rc[10];
   int lmax, nrho;
   if (cpsdh_read_psp(argv[2], atom, &zv, &amass, &lmax, rc, &nrho, &drho, rho, vp, wp, cap)) FAIL("read_psp");
   /* one ion at the origin */
   int katm = 0;
   double omega, unita[9], unitg[9], rcell[81], cewald, rcut;
   double *G = calloc(3 * (size_t)nfft3d, sizeof(double));
   int *masker = calloc(n2ft3d, sizeof(int));
   double *tg = calloc(nfft3d, sizeof(double)), *vc = calloc(nfft3d, sizeof(double));
   double *vg = calloc(nfft3d, sizeof(double));
   if (cpsdh_setup(1, 10.0, nfft, 1, &katm, &zv, &omega, unita, unitg, G, masker, tg, vc, vg, rcell, &cewald,
                   &rcut))
      FAIL("setup");
   double *vl = calloc(nfft3d, sizeof(double)), *vnl = calloc(9 * (size_t)nfft3d, sizeof(double));
   double vnlnrm[9] = {0};
   if (cpsdh_kbpp(zv, lmax, nrho, drho, rho, vp, wp, nfft3d, G, vl, vnl, vnlnrm)) FAIL("kbpp");

   /* .vpp round trip */
   snprintf(path, sizeof path, "%s/X.vpp", argv[1]);
   if (cpsdh_write_vpp(path, 1, nfft, 10.0, atom, amass, zv, lmax, rc, vnlnrm, vl, vnl)) FAIL("write_vpp");
   {
      int ic, nf, lm;
      double un, am, z, rc2[4], nrm2[9];
      char at[3];
      double *vl2 = calloc(nfft3d, sizeof(double)), *vnl2 = calloc(9 * (size_t)nfft3d, sizeof(double));
      if (cpsdh_read_vpp(path, &ic, &nf, &un, at, &am, &z, &lm, rc2, nrm2, vl2, vnl2, nfft3d)) FAIL("read_vpp");
      const int nl = lm * lm;
      if (ic != 1 || nf != nfft || un != 10.0 || am != amass || z != zv || lm != lmax) FAIL("vpp header");
      if (memcmp(vl2, vl, nfft3d * sizeof(double)) || memcmp(vnl2, vnl, (size_t)nl * nfft3d * sizeof(double)) ||
          memcmp(nrm2, vnlnrm, nl * sizeof(double)) || memcmp(rc2, rc, (lm + 1) * sizeof(double)))
         FAIL("vpp data");
      /* a reader with room for a smaller grid refuses the file */
      if (cpsdh_read_vpp(path, &ic, &nf, &un, at, &am, &z, &lm, rc2, nrm2, vl2, vnl2, nfft3d / 2) == 0)
         FAIL("vpp capacity not checked");
      free(vl2); free(vnl2);
   }
   /* a truncated .vpp is refused */
   {
      FILE *f = fopen(path, "rb");
      snprintf(path, sizeof path, "%s/short.vpp", argv[1]);
      FILE *g = fopen(path, "wb");
      if (!f || !g) FAIL("open");
      char buf[4096];
      size_t n = fread(buf, 1, sizeof buf, f);
      fwrite(buf, 1, n / 2, g);
      fclose(f); fclose(g);
      int ic, nf, lm;
      double un, am, z, rc2[4], nrm2[9];
      char at[3];
      double *vl2 = calloc(nfft3d, sizeof(double)), *vnl2 = calloc(9 * (size_t)nfft3d, sizeof(double));
      if (cpsdh_read_vpp(path, &ic, &nf, &un, at, &am, &z, &lm, rc2, nrm2, vl2, vnl2, nfft3d) == 0)
         FAIL("short vpp accepted");
      free(vl2); free(vnl2);
   }

   /* orbitals, ELC round trip */
   const int ne[2] = {6, 0};
   double *psi = calloc(6 * (size_t)n2ft3d, sizeof(double)), *back = calloc(6 * (size_t)n2ft3d, sizeof(double));
   if (cpsdh_random_orbitals(nfft, 6, masker, G, 11ull, 0.05, psi)) FAIL("random_orbitals");
   for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
         const double s = cpsdh_gcdotc(nfft, psi + (size_t)i * n2ft3d, psi + (size_t)j * n2ft3d);
         if ((i == j && (s < 1 - 1e-12 || s > 1 + 1e-12)) || (i != j && (s < -1e-12 || s > 1e-12)))
            FAIL("orbitals not orthonormal (%d %d %g)", i, j, s);
      }
   snprintf(path, sizeof path, "%s/ELCOUT", argv[1]);
   if (cpsdh_write_elc(path, 1, nfft, 10.0, 1, ne, psi)) FAIL("write_elc");
   {
      int ic, nf, isp, ne2[2];
      double un;
      if (cpsdh_read_elc(path, &ic, &nf, &un, &isp, ne2, back, 6L * n2ft3d)) FAIL("read_elc");
      if (ic != 1 || nf != nfft || un != 10.0 || isp != 1 || ne2[0] != 6 || ne2[1] != 0) FAIL("elc header");
      if (memcmp(back, psi, 6 * (size_t)n2ft3d * sizeof(double))) FAIL("elc data");
      if (cpsdh_read_elc(path, &ic, &nf, &un, &isp, ne2, back, 5L * n2ft3d) == 0) FAIL("elc capacity not checked");
   }
   /* IONOUT */
   {
      snprintf(path, sizeof path, "%s/IONOUT", argv[1]);
      const char *names[1] = {"X.vpp"};
      const double r2[3] = {0.5, -0.25, 0.125};
      if (cpsdh_write_ionout(path, 1, names, &lmax, 1, &katm, r2)) FAIL("write_ionout");
   }
   free(rho); free(vp); free(wp); free(G); free(masker); free(tg); free(vc); free(vg); free(vl); free(vnl);
   free(psi); free(back);
   printf("host io ok\n");
   return 0;
}
