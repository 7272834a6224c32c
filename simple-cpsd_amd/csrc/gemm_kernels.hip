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
 *                      real coefficients, split over blockIdx.z; the G-space
 *                      weight (1 on the kx = 0 plane, 2 elsewhere, the
 *                      2*sum - sum_{kx=0} of ccpsdv3/gcdotc.c:25-27) is applied
 *                      while staging A.
 *   k_gemm_tall      : tall output [ncol][K] = X^T A with a short reduction
 *                      (projectors or orbitals), SD epilogues fused.
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

#define TT 64   /* output tile edge */
#define KT 16   /* K step */

__global__ void __launch_bounds__(256)
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
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int wm = wave & 1, wn = wave >> 1;
   const int lr = tid >> 2, lq = (tid & 3) * 4;
   const int am = tm * TT + lr, bn = tn * TT + lr;
   const bool aval = am < M, bval = bn < N;
   const double *Ap = A + (size_t)(aval ? am : 0) * lda;
   const double *Bp = B + (size_t)(bval ? bn : 0) * ldb;

   v4d acc[2][2];
#pragma unroll
   for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

   double ra[4], rb[4];
   auto gload = [&](size_t k0) {
      if (aval) {
         const double2 x = *reinterpret_cast<const double2 *>(Ap + k0 + lq);
         const double2 y = *reinterpret_cast<const double2 *>(Ap + k0 + lq + 2);
         const double2 wx = *reinterpret_cast<const double2 *>(w + k0 + lq);
         const double2 wy = *reinterpret_cast<const double2 *>(w + k0 + lq + 2);
         ra[0] = x.x * wx.x; ra[1] = x.y * wx.y; ra[2] = y.x * wy.x; ra[3] = y.y * wy.y;
      } else {
         ra[0] = ra[1] = ra[2] = ra[3] = 0.0;
      }
      if (bval) {
         const double2 x = *reinterpret_cast<const double2 *>(Bp + k0 + lq);
         const double2 y = *reinterpret_cast<const double2 *>(Bp + k0 + lq + 2);
         rb[0] = x.x; rb[1] = x.y; rb[2] = y.x; rb[3] = y.y;
      } else {
         rb[0] = rb[1] = rb[2] = rb[3] = 0.0;
      }
   };

   int buf = 0;
   if (kb < ke) gload(kb);
   for (size_t k0 = kb; k0 < ke; k0 += KT) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
         As[buf][lr][lq + q] = ra[q];
         Bs[buf][lr][lq + q] = rb[q];
      }
      __syncthreads();
      if (k0 + KT < ke) gload(k0 + KT);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
         const int kk = 4 * ks + (lane >> 4);
         const double a0 = As[buf][32 * wm + (lane & 15)][kk];
         const double a1 = As[buf][32 * wm + 16 + (lane & 15)][kk];
         const double b0 = Bs[buf][32 * wn + (lane & 15)][kk];
         const double b1 = Bs[buf][32 * wn + 16 + (lane & 15)][kk];
         acc[0][0] = mfma16x16x4(a0, b0, acc[0][0]);
         acc[0][1] = mfma16x16x4(a0, b1, acc[0][1]);
         acc[1][0] = mfma16x16x4(a1, b0, acc[1][0]);
         acc[1][1] = mfma16x16x4(a1, b1, acc[1][1]);
      }
      buf ^= 1;
   }

#pragma unroll
   for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
         for (int r = 0; r < 4; ++r) {
            const int m = tm * TT + 32 * wm + 16 * bi + (lane >> 4) + 4 * r;
            const int n = tn * TT + 32 * wn + 16 * bj + (lane & 15);
            if (m < M && n < N) partial[((size_t)kz * M + m) * N + n] = acc[bi][bj][r];
         }
}

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

#define TK 64   /* output rows (K) per workgroup */

template<int EPI, int NT>
__global__ void __launch_bounds__(256) k_gemm_tall(TallArgs a)
{
   __shared__ double As[2][16][TK + 4];
   __shared__ double Xs[2][NT][17];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const size_t k0 = (size_t)blockIdx.x * TK;
   const int n0 = blockIdx.y * NT;
   constexpr int NB = NT / 16;
   v4d acc[NB];
#pragma unroll
   for (int j = 0; j < NB; ++j) acc[j] = (v4d){0.0, 0.0, 0.0, 0.0};

   /* loaders: A chunk 16 x 64 (row = tid/16, 4 doubles at col 4*(tid%16));
    * X chunk NT x 16 */
   const int ar = tid >> 4, ac = (tid & 15) * 4;
   double ra[4], rx[NB];
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
      for (int q = 0; q < NB; ++q) {
         const int e = tid + 256 * q;           /* e < NT*16 */
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
      for (int q = 0; q < NB; ++q) {
         const int e = tid + 256 * q;
         Xs[buf][e >> 4][e & 15] = rx[q];
      }
      __syncthreads();
      if (m0 + 16 < a.Mr) gload(m0 + 16);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
         const int mm = 4 * ks + (lane >> 4);
         const double b = As[buf][mm][16 * wave + (lane & 15)];
#pragma unroll
         for (int j = 0; j < NB; ++j) {
            const double x = Xs[buf][16 * j + (lane & 15)][mm];
            acc[j] = mfma16x16x4(x, b, acc[j]);
         }
      }
      buf ^= 1;
   }

   const size_t k = k0 + 16 * wave + (lane & 15);
#pragma unroll
   for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
         const int n = n0 + 16 * j + (lane >> 4) + 4 * r;
         if (n < a.Nc) {
            const size_t idx = (size_t)n * a.ldc + k;
            if (EPI == EPI_KBAPPLY) {
               const double cp = a.CP[idx] + acc[j][r];
               a.C2[idx] = a.C1[idx] + a.dte * cp;
               if (a.write_cp) a.CP[idx] = cp;
            } else {
               a.C2[idx] += a.dte * acc[j][r];
            }
         }
      }
}

int launch_gemm_tall(int epi, const TallArgs &a, hipStream_t s)
{
   if (a.K % TK) { cpsd_set_error("gemm_tall: K=%zu not a multiple of %d", a.K, TK); return -1; }
   const int NT = 64;
   dim3 grid(a.K / TK, (a.Nc + NT - 1) / NT);
   if (epi == EPI_KBAPPLY) hipLaunchKernelGGL((k_gemm_tall<EPI_KBAPPLY, 64>), grid, dim3(256), 0, s, a);
   else hipLaunchKernelGGL((k_gemm_tall<EPI_CONSTRAINT, 64>), grid, dim3(256), 0, s, a);
   CPSD_CHECK(hipGetLastError());
   return 0;
}
