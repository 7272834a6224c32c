/*
 * oracle.h -- CPU restatement of Simple-CPSD's electronic steepest-descent
 * step, used ONLY as test infrastructure (the checker).
 *
 * This library is linked or loaded by the tests under tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * else.  It is never the product path: the product is the HIP library built
 * from simple-cpsd_amd/, which fails loudly when its GPU code is missing.
 *
 * Every function restates the reference algorithm in plain C (FP64, serial,
 * sequential sums) and cites the reference file:line it follows.  Paths are
 * relative to the reference's Cpsd-mpi-bamboo/ directory.
 *
 * Scope: ispin = 1 (spin-restricted), fixed (move = 0) or moving (move = 1)
 * ions; ispin = 2 is rejected with an error message and return code.
 */
#ifndef CPSD_ORACLE_H
#define CPSD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1D / 3D FFT (fftpack3d/fftpack3d.f semantics) ------------------- */

/* Unnormalised complex DFT of length n, in place, interleaved (re,im).
 * sign = -1: X_k = sum_j x_j exp(-2 pi i jk/n)   (FFTPACK dcfftf)
 * sign = +1: X_k = sum_j x_j exp(+2 pi i jk/n)   (FFTPACK dcfftb)       */
void oracle_cfft1d(int n, int sign, double *x);

/* cr_fft3b (fftpack3d/fftpack3d.f:22-96): half-complex (N/2+1)xNxN, padded
 * real rows of N+2, unnormalised inverse transform, pads zeroed.          */
void oracle_cr_fft3b(int nfft, double *a);
/* rc_fft3f (fftpack3d/fftpack3d.f:107-196): forward, Im(kx=0)=Im(kx=N/2)=0 */
void oracle_rc_fft3f(int nfft, double *a);

/* ---- G-space reductions (ccpsdv3/gcdotc.c, gsdot.c, gssum.c) ------- */
double oracle_gcdotc(int nfft, const double *a, const double *b);
double oracle_gsdot(int nfft, const double *a, const double *b);
double oracle_gssum(int nfft, const double *a);

/* ---- structure factors (ccpsdv3/strfac.c) ---------------------------- */
void oracle_phafac(int nfft, const double *unitg, int nion, const double *r1,
                   double *ex1, double *ex2, double *ex3);
void oracle_strfac(int nfft, const double *ex1, const double *ex2,
                   const double *ex3, double *exi);

/* ---- LDA (ccpsdv3/dirac_exchange.c, ccpsdv3/vosko.c), ispin = 1 ----- */
void oracle_dirac_exchange(int n, const double *rho, double *vx, double *ex);
void oracle_vosko(int n, const double *dn, double *xcp, double *xce);
/* Abramowitz-Stegun erfc of ccpsdv3/cpsd.c:11-17,541-543 */
double oracle_erfc_as(double w);

/* ---- setup (ccpsdv3/main.c:145-320 and helpers) ---------------------- */
void oracle_get_cube(int icube, double unit, double *omega, double *unita,
                     double *unitg);
void oracle_get_g(int nfft, const double *unitg, double *g);
void oracle_get_masker(int nfft, double ggcut, const double *unitg,
                       int *masker1);
/* Builds G, masker (n2ft3d ints), tg, vc, vg, rcell (27 images), cewald,
 * rcut, omega, unitg, unita.  zv per kind, katm per ion.                    */
void oracle_setup(int icube, double unit, int nfft, int nion, const int *katm,
                  const double *zv, double *omega, double *unita,
                  double *unitg, double *G, int *masker, double *tg,
                  double *vc, double *vg, double *rcell, double *cewald,
                  double *rcut);

/* KB pseudopotential on this cell's grid (ckbppv3/dkbppv3.c:335-444).
 * vp, wp: (lmax+1)*nrho column-major as in the reference (vp is modified). */
void oracle_kbpp(double zv, int lmax, int nrho, double drho, const double *rho,
                 double *vp, const double *wp, int nfft3d, const double *G,
                 double *vl, double *vnl, double *vnlnrm);
/* Simpson rule of ckbppv3/dkbppv3.c:310-328 */
double oracle_simp(int n, const double *y, double h);

/* Plane-wave starting orbitals (cmakepsi/dmakepsi.c:86-133).
 * spec: ne rows of (i, j, k, l) integers as in PSI0.DAT.                   */
void oracle_makepsi(int nfft, int ne, const int *spec, double *psi);

/* ---- the SD step (ccpsdv3/cpsd.c:33-698) ----------------------------- */
/* Same signature and semantics as the reference cpsd(); returns 0 on
 * success, nonzero (with a message on stderr) for unsupported inputs.
 * *lag_iters receives the Lagrange iteration count of the last step
 * (may be NULL).                                                          */
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
                double *deltar);
int oracle_last_lagrange_iterations(void);
/* preconditioned SD step (no counterpart in the reference): with K set,
 * every step becomes c2 = c1 + K o dte (cpsi + c1 h) before the overlaps and
 * the constraint; K[nfft3d] one factor per half-complex G, h[ne*ne]
 * column-major; K = NULL restores the reference's step.  The overlaps are
 * then those of c2 - c1 D, D = <c1|c2> - I (the step's component along c1
 * taken out), and lmd = X - D/dte with X the Lagrange fixed point, so that
 * c2 + dte c1 lmd is orthonormal.  Pointers are kept, not copied. */
void oracle_set_preconditioner(const double *K, const double *h);

/* Symmetric eigen-decomposition used by the epilogue (ccpsdv3/main.c:470-484:
 * dsyev + eigsrt).  Cyclic Jacobi; eigenvalues sorted descending as eigsrt
 * does; V column-major n x n.                                               */
void oracle_syev_sorted(int n, double *a, double *eig);

/* CPU-baseline estimate of one SD step from a bounded sample of each
 * component at full size (sample_timing.c); parts[10]. */
double oracle_step_time_sampled(int nfft, int ne, int nion, int nproj, int lag_iters, int inner,
                                int sample_items, double *parts);

#ifdef __cplusplus
}
#endif
#endif
