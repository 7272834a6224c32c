/*
 * cpsd_internal.h -- device data layout and kernel launchers of the MI355X
 * SD step (library-internal).
 *
 * HBM layout (FP64 path):
 *   - orbitals C1, C2, CP: [ne][ngp] complex (interleaved re, im), only the
 *     masked G vectors are stored ("packed"), ordered as the reference's
 *     half-complex grid (kx fastest, then ky, then kz; ccpsdv3/get_g.c:20-34).
 *     Each (ky,kz) row of the grid contributes one contiguous kx interval,
 *     described by rstart/rlen/roff.  ngp = ng rounded up to 64; the tail
 *     is zero and carries weight 0.
 *   - T: z-transformed orbitals [ne][z][ky][kx], kx < N/2+1 (the y/x passes
 *     read one (ky,kx) plane per z).
 *   - real-space fields (dn, V, xcp, xce): dense [z][y][x] of N^3 doubles.
 *   - projectors B: [nproj][ngp] complex, beta_{I,lm}(G) = S_I(G) vnl(G)
 *     (times i for p channels), ions ordered as in the reference loops.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <stddef.h>
#include <stdint.h>

#define CPSD_CHECK(call)                                                       \
   do {                                                                        \
      hipError_t e_ = (call);                                                  \
      if (e_ != hipSuccess) {                                                  \
         cpsd_set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,     \
                        hipGetErrorString(e_));                                \
         return -1;                                                            \
      }                                                                        \
   } while (0)

void cpsd_set_error(const char *fmt, ...);

/* hipFuncAttributeMaxDynamicSharedMemorySize = bytes for kernel k on the
 * current device, set on the first launch per device; done is the launch
 * site's own record (one bit per device ordinal, devices from 64 up are set
 * on every launch) */
static inline int dyn_lds_once(const void *k, int bytes, std::atomic<unsigned long long> &done)
{
   int dev = 0;
   CPSD_CHECK(hipGetDevice(&dev));
   const unsigned long long bit = dev < 64 ? 1ull << dev : 0ull;
   if (bit && (done.load(std::memory_order_acquire) & bit)) return 0;
   CPSD_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
   done.fetch_or(bit, std::memory_order_release);
   return 0;
}

/* ---- FFT ------------------------------------------------------------ */

enum { SRC_PACKED = 0, SRC_VG = 1 };
/* DST_SD: out = c1 + dte (-scal1 V psi - tg c1), the SD update before the
 * nonlocal term (used when -H psi itself is not needed afterwards) */
enum { DST_HPSI = 0, DST_DNG = 1, DST_FULL = 2, DST_SD = 3 };
enum { PM_DENSITY = 0, PM_VTOT = 1, PM_VPSI = 2, PM_C2R_OUT = 3, PM_R2C_DN = 4, PM_R2C_IN = 5 };

struct RowTable {
   const int *rstart, *rlen, *roff;  /* N*N rows each */
   int nhc;                          /* kx columns that can be nonzero */
};

struct ZArgs {
   RowTable rows;
   const double2 *src;   /* packed (or full-grid) input, batch stride sstride */
   size_t sstride;
   double2 *T;           /* [batch][z][ky][NH] */
   const double2 *tw;    /* exp(+2 pi i m/N), m < N */
   /* SRC_VG */
   const double2 *vall, *dng;
   const double *vc;
   double scal2;
   /* r->G outputs */
   double2 *out;         /* packed or full grid, batch stride ostride */
   size_t ostride;
   const double2 *c1;    /* DST_HPSI, DST_SD: kinetic term uses C1 */
   const double *tg;
   double scal1;
   double dte;           /* DST_SD */
   /* set by the launcher: orbitals per workgroup (blockIdx.z covers
    * [z*bper, min(nbatch, (z+1)*bper))) */
   int nbatch, bper;
   /* FP32 transform (SRC_PACKED, DST_HPSI, DST_SD): T holds float2 at the
    * same base pointer; inputs and outputs stay FP64 */
   int fp32;
};

struct PlaneArgs {
   double2 *T;           /* [batch][z][ky][NH] (in place for PM_VPSI) */
   int nhload;           /* kx columns loaded (others are zero) */
   const double2 *tw;
   int b0, nb;           /* PM_DENSITY: orbitals [b0 + blockIdx.y*nb, ...) */
   int accumulate;       /* PM_DENSITY: add into dnpart instead of overwriting */
   int nbatch;
   double *dnpart;       /* PM_DENSITY: [blocks][N^3] */
   const double *V;      /* PM_VPSI */
   double *Vout;         /* PM_VTOT */
   const double *xcp;    /* PM_VTOT */
   const double *dn;     /* PM_R2C_DN: field 2*dn*scal1 */
   double scal1;
   const double *rin;    /* PM_R2C_IN: padded real input [z][y][N+2] */
   double *rout;         /* PM_C2R_OUT: padded real output */
   int fp32;             /* PM_DENSITY, PM_VPSI: float2 planes, FP64 sums and V */
};

int fft_supported(int n);
int launch_zpass_g2r(int n, int src, const ZArgs &a, int nbatch, hipStream_t s);
int launch_zpass_r2g(int n, int dst, const ZArgs &a, int nbatch, hipStream_t s);
int launch_plane(int n, int mode, const PlaneArgs &a, int nblocks, hipStream_t s);
/* workgroups of the density plane pass one CU holds at once */
int plane_density_occupancy(int n, int fp32);
/* the FP64 density and V psi passes at grid n prefetch the next orbital's
 * plane (several orbitals per workgroup pay for it) */
int plane_prefetch_on(int n);

/* ---- GEMM ------------------------------------------------------------- */

/* partial[z][m][n] = sum_{k in chunk z} w[k] A[m][k] B[n][k]
 * (A: M rows of stride lda, B: N rows of stride ldb; K multiple of 32).
 * symm (A == B, M == N): only the 16 x 16 blocks on or above the diagonal
 * are formed.                                                             */
/* out[m][n] = sum_k w[k] A[m][k] B[n][k] directly (no partials) for the
 * small products gemm_tn_small_ok() admits; symm: out is written in full */
int gemm_tn_small_ok(int M, int N);
int launch_gemm_tn_small(const double *A, size_t lda, int M, const double *B, size_t ldb, int N,
                         const double *w, size_t K, int symm, double *out, hipStream_t s);
/* tile list of a symmetric product restricted to some tiles (k_gemm_tn_splitk) */
struct GramTiles {
   int on, noff, nd;
   unsigned char otm[32], otn[32], dt[8];
};
int launch_gemm_tn_splitk(const double *A, size_t lda, int M, const double *B, size_t ldb,
                          int N, const double *w, size_t K, int splits, int symm,
                          double *partial, hipStream_t s);
/* out[m*N + n] = sum_z partial[z][m][n], z < splits (symm: 0, or the block
 * edge below whose diagonal the partials were not formed, all with the same
 * split count) */
int launch_reduce_partials(const double *partial, int splits, int M, int N, double *out,
                           int symm, hipStream_t s);
/* the KB projections' split-K reduction into P (nproj x ne) that also writes
 * the apply coefficients X[m + n nproj] = -P[m][n] scal2 / pnrm[m] */
int launch_reduce_partials_kb(const double *partial, int splits, int nproj, int ne, double *P, const double *pnrm,
                              double scal2, double *X, hipStream_t s);
/* the same for a symmetric launch_gemm_tn_splitk(..., splits, symm = 1, ...),
 * whose diagonal and off-diagonal tiles use different split counts */
int launch_reduce_partials_symm(const double *partial, int M, size_t K, int splits, double *out,
                                hipStream_t s);
/* split counts a symmetric product uses: off-diagonal tiles and diagonal pairs */
void gemm_tn_symm_splits(int M, size_t K, int splits, int *sp_off, int *sp_dd);
/* the same for a product restricted to noff off-diagonal and nd diagonal
 * tiles over about wgs workgroups (diagonal tiles as two half-range splits) */
void gemm_tn_list_splits(int noff, int nd, size_t K, long wgs, int *sp_off, int *sp_dd);
/* the Gram of [C1 C2] (M = 2 ne rows, ne a multiple of 128, C2 from row
 * c2row) without its C1 x C1 block; returns the split counts of the
 * off-diagonal and diagonal tiles for launch_reduce_partials_split */
int launch_gram_c2(const double *A, size_t lda, int M, int c2row, const double *w, size_t K, long wgs,
                   double *partial, size_t partial_cap, int *sp_off, int *sp_dd, hipStream_t s);
/* rows and columns both in [skip0, skip1) are left as they are */
int launch_reduce_partials_split(const double *partial, int sp_off, int sp_dd, int M, int skip0, int skip1,
                                 double *out, hipStream_t s);

/* EPI_STORE: C2[n][k] = result (the rotation c1 = c2 V of the epilogue);
 * EPI_PRECOND: C2 = C1 + Kp o (C2 - C1 + dte result), the preconditioned
 * step with result = C1 h (Kp one factor per complex coefficient) */
enum { EPI_KBAPPLY = 0, EPI_CONSTRAINT = 1, EPI_NLFORCE = 2, EPI_STORE = 3, EPI_PRECOND = 4 };
struct TallArgs {
   const double *A;  size_t lda;   /* A rows (m < Mr), contiguous along k */
   const double *X;  int ldx;      /* X[m*xsm + n*ldx] */
   int xsm = 1;
   int Mr, Nc;                     /* reduction length, output columns */
   size_t K;                       /* output rows (multiple of 64) */
   double *C2; const double *C1; double *CP; size_t ldc;
   double dte; int write_cp;
   /* EPI_NLFORCE: for output row n (a projector), sum over the block's G of
    * w Gp (Z.im beta.re - Z.re beta.im) with beta = Bp[n][G]; one [Nc][3]
    * record per K block in fpart */
   const double2 *Bp; const double *w, *Gp; int ngp; double *fpart;
   const double *Kp = nullptr;     /* EPI_PRECOND: [K / 2] */
   int ntn, per;                   /* set by the launcher */
};
int launch_gemm_tall(int epi, const TallArgs &a, hipStream_t s);

/* FP32 forms of the KB projection and apply (the FP32-transform mode):
 * partial[z][p][n] = sum_{k in split z} w_k Bf[p][k] C[n][k] for `splits`
 * even K ranges (rows of Bf and C of length K), and
 * C2[n][k] += dte sum_p X[p xsm + n ldx] Bf[p][k]; Bf = FP32 projectors */
int launch_kbproj_f32(const float *Bf, int M, const double *C, int N, const double *w, size_t K, int splits,
                      double *partial, size_t partial_cap, hipStream_t s);
int launch_kbapply_f32(const float *Bf, int Mr, const double *X, int xsm, int ldx, int Nc, size_t K,
                       double *C2, size_t ldc, double dte, hipStream_t s);
int launch_to_f32(const double *a, float *b, size_t n, hipStream_t s);
/* the preconditioned step of the FP32 steps: C2 = C1 + Kp o (C2 - C1 + dte
 * C1 H), the product in FP32 (C1, C2: [ne][K] reals, K a multiple of 128) */
int launch_precond_f32(const double *C1, int ne, const double *H, size_t K, double *C2, double dte,
                       const double *Kp, hipStream_t s);
/* the KB apply and the preconditioned update of an FP32 step in one pass:
 * C2 = C1 + Kp o (C2 - C1 + dte (B X + C1 H)) */
int launch_kbapply_precond_f32(const float *Bf, int nproj, const double *X, const double *C1, int ne,
                               const double *H, size_t K, double *C2, double dte, const double *Kp,
                               hipStream_t s);
/* the projection with BF16 projectors Bb and C1 w rounded to BF16, FP32 sums
 * (precision mode 2) */
int launch_kbproj_bf16(const __bf16 *Bb, int M, const double *C, int N, const double *w, size_t K, int splits,
                       double *partial, size_t partial_cap, hipStream_t s);
int launch_to_bf16(const double *a, __bf16 *b, size_t n, hipStream_t s);
/* the overlaps of an FP32 step from its step d = C2 - C1 (split-BF16
 * products, FP32 sums): partial[z][i][n] = sum_k d_i w_k [c1 | d]_n over
 * *splits K ranges (ne 128 or 256: gram_dbf16_ok; ne K 8 bytes below 4 GiB:
 * gram_dbf16_fits), then with the reduced
 * Gx (ne x 2ne) and the carried <c1|c1> the Gram blocks of C2 */
int gram_dbf16_ok(int ne);
int gram_dbf16_fits(int ne, size_t K);
int launch_gram_dbf16(const double *C1, const double *C2, int ne, const double *w, size_t K, double *partial,
                      size_t partial_cap, int *splits, hipStream_t s);
int launch_gram_dfix(const double *Gx, const double *g11, int ne, int c1_second, double *Gm, hipStream_t s);
/* precision mode 3: the split-BF16 projectors (rows BsH + BsL; transposed
 * [K][npad] BtH + BtL), the KB projection on them, and the apply, the
 * preconditioned update and the constraint:
 *   C2 (+)= dte (B^T X + C1^T X2)   (X[p + n ldx], X2[m + n ldx2])
 * or with precond C2 = C1 + Kp o (C2 - C1 + dte v); ne 128 or 256
 * (apply_dbf16_ok); XTh, XTl: ne (npad + Mr2) bf16 each of scratch */
int launch_to_bf16x2(const double *a, __bf16 *BsH, __bf16 *BsL, size_t n, hipStream_t s);
int launch_split_bt(const double *B, int nproj, int npad, size_t K, __bf16 *BtH, __bf16 *BtL, hipStream_t s);
int launch_kbproj_dbf16(const __bf16 *BsH, const __bf16 *BsL, int M, const double *C, int N, const double *w,
                        size_t K, int splits, double *partial, size_t partial_cap, hipStream_t s);
int apply_dbf16_ok(int ne);
int launch_apply_dbf16(const __bf16 *BtH, const __bf16 *BtL, int npad, const double *X, int nproj, int ldx,
                       const double *C1, int Mr2, const double *X2, int ldx2, int ne, size_t K, double *C2,
                       double dte, const double *Kp, int precond, __bf16 *XTh, __bf16 *XTl, hipStream_t s);

/* ---- learned preconditioner (precond_mlp.hip) ----------------------------- */
#define CPSD_MLP_NPARAM 1249
int launch_mlp_predict(const float *X, long n, const float *P, float *y, double *K, double alpha, double kmin,
                       double kmax, hipStream_t s);
int launch_precond_features(const double *tg, const double2 *B, const double *pnrm, int nproj,
                            const double2 *vall, double omega, int ng, int ngp, float *X, hipStream_t s);
int launch_precond_stats(const double2 *CP, const double2 *Rh, const double2 *C1, const double2 *Cinf, int ne,
                         int ngp, double dte, double *A, double *B, hipStream_t s);

/* ---- epilogue (host) ---------------------------------------------------- */
/* eigenpairs of the symmetric n x n column-major matrix a (upper triangle
 * read): a <- eigenvectors in columns, eig <- eigenvalues, both in the order
 * dsyev + eigsrt leave them (descending) */
extern "C" int cpsd_syev_sorted(int n, double *a, double *eig);

/* ---- SD kernels --------------------------------------------------------- */
int sd_kernels_probe(void);
