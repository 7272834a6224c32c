/* sd_kernels.h -- argument blocks and launchers of sd_kernels.hip */
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

struct StructArgs {
   int n, ng, ngp, nion, nproj;
   const int *kx, *ky, *kz;          /* grid coordinates of packed G */
   const double2 *ex1, *ex2, *ex3;   /* phase factors per ion */
   const int *katm, *lmmax;
   const double *zv;
   const double *vl;                 /* [kind][ngp] */
   const double *vnl;                /* [kind][9][ngp] */
   double2 *s, *vall, *B;
};

struct LagArgs {
   int ne;
   const double *s11, *s12, *s21, *s22;
   double *X;        /* 2 ne^2 */
   double *st1;      /* ne^2 */
   double *lmd;
   double *wgmax;    /* per workgroup */
   double *info;     /* [iterations, last max|dX|] */
   unsigned *bar;    /* grid-barrier counter, zero at launch (sd_smat clears it) */
   int *err;
   int itmax = 30;   /* the reference's ITERLMD (ccpsdv3/cpsd.c:19) */
};

struct EnergyArgs {
   int ne, nion, nproj, inner;
   const double *hml, *P, *pnrm;
   const double *gsum_part; int ngsum_parts;
   const double *xc_part; int nxc_parts;
   const double *orb_part; int norb_parts;
   const double *ewald_part;
   double scal2, omega, dv, cewald, dt, dte;
   double *E;        /* 12: E[0..9], deltae, deltac; E[0] holds the previous energy on entry */
};

int sd_pack(const double2 *full, size_t fstride, const int *gidx, int ng, int ngp, double2 *packed,
            size_t pstride, int ne, hipStream_t s);
int sd_check_unmasked(const double2 *full, size_t fstride, const int *pidx, int nfft3d, int ne,
                      int *flag, hipStream_t s);
int sd_unpack(const double2 *packed, size_t pstride, const int *pidx, int nfft3d, double2 *full,
              size_t fstride, int ne, hipStream_t s);
int sd_pad_real(const double *dense, int n, double *padded, hipStream_t s);
int sd_structure(const StructArgs &a, hipStream_t s);
int sd_phafac(int n, const double *unitg, int nion, const double *r1, double2 *ex1, double2 *ex2,
              double2 *ex3, hipStream_t s);
int sd_dn_reduce_xc(const double *dnpart, int nparts, size_t n3, double scal2, double *dn, double *xcp,
                    double *xce, hipStream_t s);
int sd_lda(const double *dn, size_t n, double *xcp, double *xce, hipStream_t s);
int sd_gsums(const double2 *sg, const double2 *dng, const double2 *vall, const double *vg, const double *vc,
             const double *w, int ngp, double *part, int nparts, hipStream_t s);
int sd_xcsums(const double *dn, const double *xce, const double *xcp, size_t n3, double *part, int nparts,
              hipStream_t s);
int sd_orbsums(const double2 *c1, const double2 *c2, double2 *cp, const double *tg, const double *w, int ngp,
               int ne, double *part, int nparts, hipStream_t s);
int sd_ewald_real(const double *r1, const int *katm, const double *zv, int nion, const double *rcell,
                  int nsh, double rcut, double *part, hipStream_t s);
int sd_erfc_as(const double *w, int n, double *out, hipStream_t s);
/* s11, s12, s21, s22 of ccpsdv3/cpsd.c:415-437 from the Gram Gm of [C1 C2]
 * (or of [C2 C1], c1_second); with g11 non-null, <c1_i|c1_j> is read from
 * the dense ne x ne g11 instead of Gm.  A non-null bar (the Lagrange
 * kernel's barrier counter) is set to zero in the same launch. */
int sd_smat(const double *Gm, const double *g11, int ne, int c1_second, double dte, double *s11, double *s12,
            double *s21, double *s22, unsigned *bar, hipStream_t s);
/* <c2f_a|c2f_b> of the orbitals after the constraint c2f = c2 + dte c1 lmd,
 * from the step's Gram (G11 from g11 when non-null) and lmd[i + j ne]:
 * G22 + dte (G21 lmd + lmd^T G12) + dte^2 lmd^T G11 lmd, into out (ne x ne);
 * work holds 2 ne x ne */
/* the preconditioned step's s matrices, with the step's component along c1
 * taken out of c2 (DP: 3 ne^2 doubles of work, D first), and lmd <- lmd -
 * D/dte after the Lagrange iteration */
int sd_smat_precond(const double *Gm, const double *g11, int ne, int c1_second, double dte, double *DP, double *s11,
                    double *s12, double *s21, double *s22, unsigned *bar, hipStream_t s);
int sd_precond_lmd(double *lmd, const double *DP, int ne, double dte, hipStream_t s);
int sd_gram_next(const double *Gm, const double *g11, int ne, int c1_second, const double *lmd, double dte,
                 double *work, double *out, hipStream_t s);
int sd_kbcoef(const double *P, int nproj, int ne, const double *pnrm, double scal2, double *X, hipStream_t s);
int sd_lagrange(const LagArgs &a, hipStream_t s);
int sd_energy(const EnergyArgs &a, hipStream_t s);
int sd_hml(const double *M, int ne, double *hml, hipStream_t s);

/* ---- ionic forces and ion update (force_kernels.hip) -------------------- */
struct ForceArgs {
   int n, ng, ngp, nion, nchunk;
   const int *kx, *ky, *kz;          /* grid coordinates of packed G */
   const double2 *ex1, *ex2, *ex3;   /* phase factors per ion */
   const int *katm;
   const double *zv, *vl, *vg, *w;   /* vl [kind][ngp]; w per real (2 per G) */
   const double *Gp;                 /* [3][ngp] Cartesian G of packed entries */
   const double2 *dng, *s;
   double scal2;
   double *part;                     /* [nion][nchunk][3] */
};
struct ForceFinishArgs {
   int nion, nchunk, nsh;
   const int *katm, *lmmax, *poff;   /* poff: first projector of each ion */
   const double *zv, *rcell, *r1, *dti;
   double rcut;
   const double *gpart;              /* [nion][nchunk][3] from sd_force_gspace */
   const double *nlsum;              /* [nproj][3] from sd_force_nlsum */
   double *fion, *r2;
};
int sd_force_gspace(const ForceArgs &a, hipStream_t s);
int sd_force_nlsum(const double *nlpart, int kblocks, int nproj, double *nlsum, hipStream_t s);
int sd_force_finish(const ForceFinishArgs &a, hipStream_t s);
