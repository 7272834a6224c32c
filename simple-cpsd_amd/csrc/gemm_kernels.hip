/*
 * gemm_kernels.hip -- FP64 MFMA contractions over the packed G vectors.
 *
 * The reference forms these as level-1 BLAS loops:
 *   - KB projections  sw = gcdotc(beta, c1_i)      ccpsdv3/cpsd.c:287-290
 *   - overlaps        gcdotc(c_i, c_j) x 4          ccpsdv3/cpsd.c:415-437
 *   - hml             -gcdotc(c1_i, cpsi_j)         ccpsdv3/cpsd.c:497-513
 *   - KB apply        c2_i += sw * beta (daxpy)     ccpsdv3/cpsd.c:290
 *   - constraint      c2 += dte c1 lmd (dgemm)      ccpsdv3/cpsd.c:482-488
 * Here they are two kernels built on v_mfma_f64_16x16x4_f64:
 *   k_gemm_tn_splitk : small M x N output, reduction over the huge K = 2*ngp
 *                      real coefficients, split over blockIdx.z; 128 x 128
 *                      output tiles (4 waves of 64 x 64); the G-space weight
 *                      (1 on the kx = 0 plane, 2 elsewhere: the
 *                      2*sum - sum_{kx=0} of ccpsdv3/gcdotc.c:25-27) is
 *                      applied while staging A.
 *   k_gemm_tall      : tall output [ncol][K] = X^T A with a short reduction
 *                      (projectors or orbitals), 128 x 128 output tiles over
 *                      8 waves, SD epilogues fused.
 * MFMA operand layout (wave64): lane l supplies A[l%16][l/16] and
 * B[l/16][l%16] of the 16x4 and 4x16 blocks; the 16x16 result comes back as
 * four doubles per lane, d[r] = C[l/16 + 4*r][l%16] (measured on gfx950).
 */
#include "cpsd_internal.h"

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma16x16x4(double a, double b, v4d c)
{
   return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

/* ------------------------------------------------------- split-K TN */

#define TT 128  /* output tile edge */
#define KT 16   /* K step */

/* two waves per SIMD: at most 256 VGPR + AGPR per lane */
__global__ void __launch_bounds__(256, 2)
k_gemm_tn_splitk(const double *__restrict__ A, size_t lda, int M,
                 const double *__restrict__ B, size_t ldb, int N,
                 const double *__restrict__ w, size_t K, size_t kchunk, int symm,
                 double *__restrict__ partial)
{
   const int tm = blockIdx.x, tn = blockIdx.y, kz = blockIdx.z;
   if (symm && tm > tn) return;
   const size_t kb = (size_t)kz * kchunk;
   const size_t ke = min(K, kb + kchunk);
   __shared__ double As[2][TT][KT + 1];
   __shared__ double Bs[2][TT][KT + 1];
   __shared__ double Ws[2][KT];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int wm = wave & 1, wn = wave >> 1;
   /* a diagonal tile of a symmetric product reads one operand block */
   const bool same = symm && tm == tn;
   /* loader: row tid/2 of the tile, 8 consecutive k at 8*(tid%2) */
   const int lr = tid >> 1, lq = (tid & 1) * 8;
   const int am = tm * TT + lr, bn = tn * TT + lr;
   const bool aval = am < M, bval = bn < N && !same;
   const double *Ap = A + (size_t)(am < M ? am : 0) * lda;
   const double *Bp = B + (size_t)(bn < N ? bn : 0) * ldb;
   double (*Bt)[TT][KT + 1] = same ? As : Bs;

   v4d acc[4][4];
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

   /* raw prefetch registers: nothing consumes them until the next LDS store,
    * so the loads stay in flight across the MFMA work of the current tile;
    * the K-tile's 16 weights are fetched by 8 threads only */
   double2 ra[4], rb[4], rw;
   auto gload = [&](size_t k0) {
#pragma unroll
      for (int h = 0; h < 4; ++h) {
         ra[h] = make_double2(0.0, 0.0);
         rb[h] = make_double2(0.0, 0.0);
         if (aval) ra[h] = *reinterpret_cast<const double2 *>(Ap + k0 + lq + 2 * h);
         if (bval) rb[h] = *reinterpret_cast<const double2 *>(Bp + k0 + lq + 2 * h);
      }
      if (tid < KT / 2) rw = *reinterpret_cast<const double2 *>(w + k0 + 2 * tid);
   };

   int buf = 0;
   if (kb < ke) gload(kb);
   for (size_t k0 = kb; k0 < ke; k0 += KT) {
#pragma unroll
      for (int h = 0; h < 4; ++h) {
         As[buf][lr][lq + 2 * h] = ra[h].x;
         As[buf][lr][lq + 2 * h + 1] = ra[h].y;
         if (!same) {
            Bs[buf][lr][lq + 2 * h] = rb[h].x;
            Bs[buf][lr][lq + 2 * h + 1] = rb[h].y;
         }
      }
      if (tid < KT / 2) { Ws[buf][2 * tid] = rw.x; Ws[buf][2 * tid + 1] = rw.y; }
      __syncthreads();
      if (k0 + KT < ke) gload(k0 + KT);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
         const int kk = 4 * ks + (lane >> 4);
         const double wk = Ws[buf][kk];
         double a[4], b[4];
#pragma unroll
         for (int i = 0; i < 4; ++i) {
            a[i] = As[buf][64 * wm + 16 * i + (lane & 15)][kk] * wk;
            b[i] = Bt[buf][64 * wn + 16 * i + (lane & 15)][kk];
         }
#pragma unroll
         for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = mfma16x16x4(a[i], b[j], acc[i][j]);
      }
      buf ^= 1;
   }

#pragma unroll
   for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj)
#pragma unroll
         for (int r = 0; r < 4; ++r) {
            const int m = tm * TT + 64 * wm + 16 * bi + (lane >> 4) + 4 * r;
            const int n = tn * TT + 64 * wn + 16 * bj + (lane & 15);
            if (m < M && n < N) partial[((size_t)kz * M + m) * N + n] = acc[bi][bj][r];
         }
}

int gemm_tn_tile(void) { return TT; }

int launch_gemm_tn_splitk(const double *A, size_t lda, int M, const double *B, size_t ldb,
                          int N, const double *w, size_t K, int splits, int symm,
                          double *partial, hipStream_t s)
{
   if (K % KT) { cpsd_set_error("gemm_tn: K=%zu not a multiple of %d", K, KT); return -1; }
   size_t kchunk = (K + splits - 1) / splits;
   kchunk = (kchunk + KT - 1) / KT * KT;
   dim3 grid((M + TT - 1) / TT, (N + TT - 1) / TT, splits);
   hipLaunchKernelGGL(k_gemm_tn_splitk, grid, dim3(256), 0, s, A, lda, M, B, ldb, N, w, K,
                      kchunk, symm, partial);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* out[m][n] = sum_z partial[z][m][n]; with symm only the tiles with
 * tile(m) <= tile(n) exist and the rest is mirrored */
__global__ void k_reduce_partials(const double *__restrict__ partial, int splits, int M, int N,
                                  int symm, double *__restrict__ out)
{
   const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (idx >= (size_t)M * N) return;
   int m = (int)(idx / N), n = (int)(idx % N);
   if (symm && (m / TT) > (n / TT)) { const int t = m; m = n; n = t; }
   double s = 0.0;
   for (int z = 0; z < splits; ++z) s += partial[((size_t)z * M + m) * N + n];
   out[idx] = s;
}

int launch_reduce_partials(const double *partial, int splits, int M, int N, double *out,
                           int symm, hipStream_t s)
{
   const size_t tot = (size_t)M * N;
   hipLaunchKernelGGL(k_reduce_partials, dim3((tot + 255) / 256), dim3(256), 0, s, partial,
                      splits, M, N, symm, out);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ---------------------------------------------------------- tall NN */

#define TK 128  /* output rows (K) per workgroup */

/* waves arranged WK (along K) x WN (along the output columns) = 8 */
template<int EPI, int NT>
__global__ void __launch_bounds__(512) k_gemm_tall(TallArgs a)
{
   constexpr int WN = NT / 64, WK = 8 / WN;
   constexpr int BI = TK / WK / 16;          /* 16-row blocks of K per wave */
   constexpr int BJ = 4;                     /* 16-column blocks per wave */
   __shared__ double As[2][16][TK + 4];
   __shared__ double Xs[2][NT][17];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int wk = wave % WK, wn = wave / WK;
   const size_t k0 = (size_t)blockIdx.x * TK;
   const int n0 = blockIdx.y * NT;
   v4d acc[BI][BJ];
#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

   /* loaders: A chunk 16 x TK (4 doubles per thread), X chunk NT x 16 */
   constexpr int XQ = NT * 16 / 512;
   const int ar = tid / (TK / 4), ac = (tid % (TK / 4)) * 4;
   double ra[4], rx[XQ];
   auto gload = [&](int m0) {
      const int m = m0 + ar;
      if (m < a.Mr) {
         const double *p = a.A + (size_t)m * a.lda + k0 + ac;
         const double2 x = *reinterpret_cast<const double2 *>(p);
         const double2 y = *reinterpret_cast<const double2 *>(p + 2);
         ra[0] = x.x; ra[1] = x.y; ra[2] = y.x; ra[3] = y.y;
      } else {
         ra[0] = ra[1] = ra[2] = ra[3] = 0.0;
      }
#pragma unroll
      for (int q = 0; q < XQ; ++q) {
         const int e = tid + 512 * q;
         const int n = e >> 4, mm = e & 15;
         const int mg = m0 + mm, ng = n0 + n;
         rx[q] = (mg < a.Mr && ng < a.Nc) ? a.X[mg + (size_t)ng * a.ldx] : 0.0;
      }
   };

   int buf = 0;
   gload(0);
   for (int m0 = 0; m0 < a.Mr; m0 += 16) {
#pragma unroll
      for (int q = 0; q < 4; ++q) As[buf][ar][ac + q] = ra[q];
#pragma unroll
      for (int q = 0; q < XQ; ++q) {
         const int e = tid + 512 * q;
         Xs[buf][e >> 4][e & 15] = rx[q];
      }
      __syncthreads();
      if (m0 + 16 < a.Mr) gload(m0 + 16);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
         const int mm = 4 * ks + (lane >> 4);
         double b[BI], x[BJ];
#pragma unroll
         for (int i = 0; i < BI; ++i) b[i] = As[buf][mm][(TK / WK) * wk + 16 * i + (lane & 15)];
#pragma unroll
         for (int j = 0; j < BJ; ++j) x[j] = Xs[buf][64 * wn + 16 * j + (lane & 15)][mm];
#pragma unroll
         for (int i = 0; i < BI; ++i)
#pragma unroll
            for (int j = 0; j < BJ; ++j) acc[i][j] = mfma16x16x4(x[j], b[i], acc[i][j]);
      }
      buf ^= 1;
   }

#pragma unroll
   for (int i = 0; i < BI; ++i) {
      const size_t k = k0 + (TK / WK) * wk + 16 * i + (lane & 15);
#pragma unroll
      for (int j = 0; j < BJ; ++j)
#pragma unroll
         for (int r = 0; r < 4; ++r) {
            const int n = n0 + 64 * wn + 16 * j + (lane >> 4) + 4 * r;
            if (n < a.Nc) {
               const size_t idx = (size_t)n * a.ldc + k;
               if (EPI == EPI_KBAPPLY) {
                  const double cp = a.CP[idx] + acc[i][j][r];
                  a.C2[idx] = a.C1[idx] + a.dte * cp;
                  if (a.write_cp) a.CP[idx] = cp;
               } else {
                  a.C2[idx] += a.dte * acc[i][j][r];
               }
            }
         }
   }
}

int launch_gemm_tall(int epi, const TallArgs &a, hipStream_t s)
{
   if (a.K % TK) { cpsd_set_error("gemm_tall: K=%zu not a multiple of %d", a.K, TK); return -1; }
   /* one output-column tile covering up to 128 orbitals, so the A rows
    * (projectors or C1) stream from HBM once */
   const int NT = a.Nc > 64 ? 128 : 64;
   dim3 grid(a.K / TK, (a.Nc + NT - 1) / NT);
   if (NT == 128) {
      if (epi == EPI_KBAPPLY) hipLaunchKernelGGL((k_gemm_tall<EPI_KBAPPLY, 128>), grid, dim3(512), 0, s, a);
      else hipLaunchKernelGGL((k_gemm_tall<EPI_CONSTRAINT, 128>), grid, dim3(512), 0, s, a);
   } else {
      if (epi == EPI_KBAPPLY) hipLaunchKernelGGL((k_gemm_tall<EPI_KBAPPLY, 64>), grid, dim3(512), 0, s, a);
      else hipLaunchKernelGGL((k_gemm_tall<EPI_CONSTRAINT, 64>), grid, dim3(512), 0, s, a);
   }
   CPSD_CHECK(hipGetLastError());
   return 0;
}
