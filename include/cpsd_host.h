/*
 * cpsd_host.h -- host-side inputs of the SD step (libcpsd_host.so, plain C).
 *
 * These build what the reference's driver and preprocessors hand to cpsd():
 *   cpsdh_setup        ccpsdv3/main.c:145-198, 287-313 (lattice, G vectors,
 *                      mask, tg/vc/vg, Ewald constant, 27 image cells)
 *   cpsdh_read_psp     the radial table read by ckbppv3/dkbppv3.c:89-114,
 *                      accepting the optional 4th token on line 2 that the
 *                      reference reader misparses
 *   cpsdh_kbpp         ckbppv3/dkbppv3.c:335-444 (KB transform to the grid)
 *   cpsdh_makepsi      cmakepsi/dmakepsi.c:86-133 (plane-wave start)
 *   cpsdh_write_elc, cpsdh_read_elc, cpsdh_write_ionout
 *                      the restart files of ccpsdv3/main.c:115-140, 539-570
 *   cpsdh_write_vpp, cpsdh_read_vpp
 *                      the .vpp files of ckbppv3/dkbppv3.c:138-150 and
 *                      ccpsdv3/main.c:266-283
 * and, with no counterpart in the reference, a seeded Gaussian starting
 * state orthonormalised under the reference's G-space inner product
 * (ccpsdv3/gcdotc.c), used for the large workloads.
 */
#ifndef CPSD_HOST_H
#define CPSD_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* Fills omega, unita[9], unitg[9], G[3*nfft3d], masker[n2ft3d], tg, vc,
 * vg[nfft3d], rcell[3*27], cewald, rcut.  katm 0-based per ion. */
int cpsdh_setup(int icube, double unit, int nfft, int nion, const int *katm, const double *zv,
                double *omega, double *unita, double *unitg, double *G, int *masker, double *tg,
                double *vc, double *vg, double *rcell, double *cewald, double *rcut);

/* Parses a PSP .dat table.  On entry *nrho_cap is the capacity of rho,
 * vp, wp (vp and wp hold (lmax+1)*nrho values, column-major as in the
 * reference).  Returns 0 on success. */
int cpsdh_read_psp(const char *path, char *atom, double *zv, double *amass, int *lmax, double *rc,
                   int *nrho, double *drho, double *rho, double *vp, double *wp, int nrho_cap);

/* KB transform; lmax may be lowered below the table's (channel lmax becomes
 * local).  vp is not modified.  vl[nfft3d], vnl[9*nfft3d], vnlnrm[9]. */
int cpsdh_kbpp(double zv, int lmax, int nrho, double drho, const double *rho, const double *vp,
               const double *wp, int nfft3d, const double *G, double *vl, double *vnl,
               double *vnlnrm);

/* .vpp files (written by ckbppv3/dkbppv3.c:138-150, read by
 * ccpsdv3/main.c:266-283): int icube, int nfft, double unit, char atom[2],
 * double amass, double zv, int lmax, double rc[lmax+1], double
 * vnlnrm[lmax^2], double vl[nfft3d], double vnl[lmax^2 nfft3d], native
 * byte order, no padding; amass unscaled.  The reader takes atom[3] (NUL
 * terminated), rc[4], vnlnrm[9], vl[cap_nfft3d], vnl[9 cap_nfft3d] and
 * refuses a file with more or less data than its header implies.
 * Return 0 on success. */
int cpsdh_write_vpp(const char *path, int icube, int nfft, double unit, const char *atom, double amass,
                    double zv, int lmax, const double *rc, const double *vnlnrm, const double *vl,
                    const double *vnl);
int cpsdh_read_vpp(const char *path, int *icube, int *nfft, double *unit, char *atom, double *amass,
                   double *zv, int *lmax, double *rc, double *vnlnrm, double *vl, double *vnl,
                   long cap_nfft3d);

/* plane-wave orbitals from PSI0.DAT-style rows (i, j, k, l) */
int cpsdh_makepsi(int nfft, int ne, const int *spec, double *psi);

/* Seeded Gaussian orbitals on the masked G vectors (splitmix64 +
 * Box-Muller), kx = 0 plane made Hermitian, Im(G=0) = 0, then
 * orthonormalised (modified Gram-Schmidt under gcdotc).  decay >= 0 damps
 * coefficient magnitudes by exp(-decay*|G|^2/2) (0: plain Gaussian). */
int cpsdh_random_orbitals(int nfft, int ne, const int *masker, const double *G, unsigned long long seed,
                          double decay, double *psi);

/* G-space inner product of ccpsdv3/gcdotc.c */
double cpsdh_gcdotc(int nfft, const double *a, const double *b);

/* ELCIN/ELCOUT restart files (read at ccpsdv3/main.c:115-140, written at
 * :539-547): int icube, int nfft, double unit, int ispin, int ne[2], then
 * (ne[0]+ne[1])*n2ft3d doubles, in native byte order with no padding.
 * cpsdh_read_elc reads the header and, if c is not NULL, the orbitals into
 * c[cap].  Return 0 on success. */
int cpsdh_write_elc(const char *path, int icube, int nfft, double unit, int ispin, const int *ne,
                    const double *c);
int cpsdh_read_elc(const char *path, int *icube, int *nfft, double *unit, int *ispin, int *ne,
                   double *c, long cap);

/* IONOUT (ccpsdv3/main.c:551-570): kinds with their psp names and lmax,
 * then per ion its 1-based kind, r2 and the zeroed r1. */
int cpsdh_write_ionout(const char *path, int nkatm, const char *const *fname, const int *lmax,
                       int nion, const int *katm, const double *r2);

#ifdef __cplusplus
}
#endif
#endif
