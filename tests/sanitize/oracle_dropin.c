/*
 * oracle_dropin.c -- test infrastructure only: the drop-in symbols cpsd(),
 * cr_fft3b_ and rc_fft3f_ (ccpsdv3/cpsd.c:33-71, fftpack3d/fftpack3d.f)
 * over the CPU oracle, so that the plain C caller (tests/c_caller) and the
 * host library run as one CPU program under AddressSanitizer and
 * UndefinedBehaviorSanitizer.
 */
#include "oracle.h"

void cpsd(const int inner, const int move, const int nion, const int ispin, const int *ne,
          const int nfft, const int nfft3d, const int n2ft3d, const double *unitg,
          const double omega, const double dt, const double fmass, const int *katm,
          const double *amass, const double *zv, const int nsh, const double *rcell,
          const double cewald, const double rcut, const int *masker, const double *G,
          const double *tg, const double *vc, const double *vg, int *lmmax, double *vnlnrm,
          double *vl, double *vnl, double *r1, double *r2, double *fion, double *c1,
          double *c2, double *cpsi, double *dn, double *hml, double *lmd, double *cpsdwork,
          double *E, double *deltae, double *deltac, double *deltar)
{
   oracle_cpsd(inner, move, nion, ispin, ne, nfft, nfft3d, n2ft3d, unitg, omega, dt, fmass, katm, amass, zv, nsh,
               rcell, cewald, rcut, masker, G, tg, vc, vg, lmmax, vnlnrm, vl, vnl, r1, r2, fion, c1, c2, cpsi, dn,
               hml, lmd, cpsdwork, E, deltae, deltac, deltar);
}

void cr_fft3b_(double *A, int *inc2c, int *inc3c, int *inc2r, int *inc3r, int *nx, int *ny, int *nz, double *tmp,
               int *ntmp)
{
   (void)inc2c; (void)inc3c; (void)inc2r; (void)inc3r; (void)ny; (void)nz; (void)tmp; (void)ntmp;
   oracle_cr_fft3b(*nx, A);
}

void rc_fft3f_(double *A, int *inc2c, int *inc3c, int *inc2r, int *inc3r, int *nx, int *ny, int *nz, double *tmp,
               int *ntmp)
{
   (void)inc2c; (void)inc3c; (void)inc2r; (void)inc3r; (void)ny; (void)nz; (void)tmp; (void)ntmp;
   oracle_rc_fft3f(*nx, A);
}
