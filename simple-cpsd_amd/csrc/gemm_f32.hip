/*
 * gemm_f32.hip -- FP32 MFMA forms of the Kleinman-Bylander projection and
 * apply (ccpsdv3/cpsd.c:259-306) for the FP32-transform precision mode:
 *   k_kbproj_f32  : partial[z][p][n] = sum_{k in split z} w_k Bf[p][k] C1[n][k]
 *                   (the projections sw before their -1/(Omega vnlnrm)
 *                   scaling), FP32 products summed in FP32 within a split,
 *                   the splits reduced in FP64 by k_reduce_partials;
 *   k_kbapply_f32 : C2[n][k] += dte sum_p X[p][n] Bf[p][k] (the nonlocal
 *                   part of the SD update), the sum in FP32, the update of
 *                   the FP64 orbitals in FP64; 128 reals x 128 orbitals per
 *                   workgroup.
 * Bf is an FP32 copy of the projectors, made once per set of ion positions;
 * the orbitals stay FP64 in HBM and are rounded to FP32 while they are
 * staged in LDS.
 *
 * Precision mode 2 adds the BF16-input projection (k_kbproj_bf16); mode 3
 * runs every product of its FP32 steps in split BF16 (three
 * v_mfma_f32_16x16x32_bf16 per block on hi/lo halves of each FP32 factor):
 *   k_kbproj_dbf16 : the KB projection;
 *   k_apply_dbf16  : the KB apply, the preconditioned update (B X + C1 h in
 *                    one pass) and the constraint (C1 lmd);
 *   k_gram_dbf16   : the overlaps of the step d = c2 - c1, <d|c1> and
 *                    <d|d>, which k_gram_dfix assembles with the carried
 *                    FP64 <c1|c1> (ccpsdv3/cpsd.c:415-437).
 *
 * v_mfma_f32_16x16x4f32: lane l supplies A[l%16][l/16] and B[l/16][l%16];
 * lane l, register r receives D[4(l/16)+r][l%16].  A sum over k may assign
 * its terms to the instruction's 4 k slots in any order, as long as A and B
 * agree: a lane that reads 4 consecutive k with one 16-byte LDS read feeds
 * slot l/16 of 4 successive instructions with them.
 */
#include "cpsd_internal.h"

/* waves per SIMD the FP32 apply kernels are sized for */
#ifndef KBAPPLY_WAVES
#define KBAPPLY_WAVES 6
#endif
/* the fused KB apply + preconditioned update: its second operand source
 * spills at 6 (44 bytes per lane), not at 4 */
#ifndef KBAPPLY_PRE_WAVES
#define KBAPPLY_PRE_WAVES 4
#endif

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma16(float a, float b, f4 c)
{
   return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

/* ------------------------------------------------------ projection */

#define PK 32           /* K step */
#define PP (PK + 4)     /* row pitch (floats): a 16-byte read of 16 rows at one k hits distinct banks */
#define PN 128          /* output columns (orbitals) per tile */

/* 8 waves as 2 (rows) x 4 (columns), each (TM/2) x 32 of the TM x 128 tile;
 * one workgroup per (tile, K split), blocks assigned to the 8 XCDs so that
 * the workgroups one XCD runs together hold consecutive (tile, split)
 * pairs: the tiles of one K range then share that XCD's L2 */
template<int TM>
__global__ void __launch_bounds__(512, 4)
k_kbproj_f32(const float *__restrict__ Bf, size_t ldb, int M, const double *__restrict__ C, size_t ldc,
             int N, const double *__restrict__ w, size_t K, size_t kchunk, int ntm, int ntile, int per,
             double *__restrict__ partial)
{
   constexpr int BI = TM / 32;   /* 16-row blocks per wave, and A loader rows per thread / 32 */
   __shared__ float As[2][TM][PP];
   __shared__ float Bs[2][PN][PP];
   const int bid = blockIdx.x;
   const long g = per > 0 ? (long)(bid & 7) * per + (bid >> 3) : bid;
   const int t = (int)(g % ntile), kz = (int)(g / ntile);
   const int tm = t % ntm, tn = t / ntm;
   const size_t kb = (size_t)kz * kchunk;
   const size_t ke = min(K, kb + kchunk);
   const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int wm = wave & 1, wn = wave >> 1;
   const int q = lane >> 4, rr = lane & 15;
   /* loader: rows lr + 32 h, k pair lq; 16 lanes cover one row's 32 k */
   const int lr = tid >> 4, lq = (tid & 15) * 2;

   f4 acc[BI][2];
#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

   float2 ra[BI];
   double2 rb[4], rw;
   auto gload = [&](size_t k0) {
#pragma unroll
      for (int h = 0; h < BI; ++h) {
         const int am = tm * TM + lr + 32 * h;
         ra[h] = am < M ? *reinterpret_cast<const float2 *>(Bf + (size_t)am * ldb + k0 + lq)
                        : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int h = 0; h < 4; ++h) {
         const int bn = tn * PN + lr + 32 * h;
         rb[h] = bn < N ? *reinterpret_cast<const double2 *>(C + (size_t)bn * ldc + k0 + lq)
                        : make_double2(0.0, 0.0);
      }
      rw = *reinterpret_cast<const double2 *>(w + k0 + lq);
   };
   auto lstore = [&](int bf) {
#pragma unroll
      for (int h = 0; h < BI; ++h) *reinterpret_cast<float2 *>(&As[bf][lr + 32 * h][lq]) = ra[h];
      /* the G-space weight (1 or 2, 0 on the padded tail) is exact in FP32 */
#pragma unroll
      for (int h = 0; h < 4; ++h)
         *reinterpret_cast<float2 *>(&Bs[bf][lr + 32 * h][lq]) =
            make_float2((float)(rb[h].x * rw.x), (float)(rb[h].y * rw.y));
   };
   auto mma = [&](int bf, int hk) {
      f4 a[BI], b[2];
#pragma unroll
      for (int i = 0; i < BI; ++i) a[i] = *reinterpret_cast<const f4 *>(&As[bf][(TM / 2) * wm + 16 * i + rr][16 * hk + 4 * q]);
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const f4 *>(&Bs[bf][32 * wn + 16 * j + rr][16 * hk + 4 * q]);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
         for (int i = 0; i < BI; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma16(a[i][c], b[j][c], acc[i][j]);
   };

   /* K tile k+1 waits in registers while tile k is multiplied and goes to
    * the other LDS buffer halfway through; one barrier per tile */
   if (kb < ke) {
      gload(kb);
      lstore(0);
   }
   __syncthreads();
   if (kb + PK < ke) gload(kb + PK);
   int buf = 0;
   for (size_t k0 = kb; k0 < ke; k0 += PK) {
      mma(buf, 0);
      if (k0 + PK < ke) {
         lstore(buf ^ 1);
         if (k0 + 2 * PK < ke) gload(k0 + 2 * PK);
      }
      mma(buf, 1);
      __syncthreads();
      buf ^= 1;
   }

   /* 16 lanes hold 16 consecutive columns of one row: 128-byte segments */
#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
         const int n = tn * PN + 32 * wn + 16 * j + rr;
#pragma unroll
         for (int r = 0; r < 4; ++r) {
            const int p = tm * TM + (TM / 2) * wm + 16 * i + 4 * q + r;
            if (p < M && n < N) partial[((size_t)kz * M + p) * N + n] = (double)acc[i][j][r];
         }
      }
}

int launch_kbproj_f32(const float *Bf, int M, const double *C, int N, const double *w, size_t K, int splits,
                      double *partial, size_t partial_cap, hipStream_t s)
{
   if (K % PK) { cpsd_set_error("kbproj_f32: K=%zu not a multiple of %d", K, PK); return -1; }
   if (splits < 1 || (size_t)splits * M * N > partial_cap) {
      cpsd_set_error("kbproj_f32: %d splits of %d x %d do not fit the partial buffer", splits, M, N);
      return -1;
   }
   const int pad128 = (M + 127) / 128 * 128 - M, pad96 = (M + 95) / 96 * 96 - M;
   const int tm = pad96 < pad128 ? 96 : 128;
   const int ntm = (M + tm - 1) / tm, ntile = ntm * ((N + PN - 1) / PN);
   size_t kchunk = (K + splits - 1) / splits;
   kchunk = (kchunk + PK - 1) / PK * PK;
   const long pairs = (long)ntile * splits;
   const int per = pairs % 8 == 0 ? (int)(pairs / 8) : 0;
   if (tm == 96)
      hipLaunchKernelGGL((k_kbproj_f32<96>), dim3((unsigned)pairs), dim3(512), 0, s, Bf, K, M, C, K, N, w, K,
                         kchunk, ntm, ntile, per, partial);
   else
      hipLaunchKernelGGL((k_kbproj_f32<128>), dim3((unsigned)pairs), dim3(512), 0, s, Bf, K, M, C, K, N, w, K,
                         kchunk, ntm, ntile, per, partial);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ----------------------------------------------------------- apply */

#define AM 16           /* projectors per step */
#define AN 128          /* output columns (orbitals) per workgroup */
#define XP (AM + 4)     /* Xs row pitch: a 16-byte read of 16 rows hits distinct banks */

/* AKT output reals (k) per workgroup, 8 waves as 2 (columns) x 4 (k), each
 * 64 columns x AKT/4 k = 4 x AKT/64 blocks; the Bs row pitch AKT + 4 puts the
 * 4 rows 4 apart that one read touches 16 banks apart.  WAVES: waves per
 * SIMD the registers are sized for.
 * TA = float: the KB apply, rows of Bf; C2 += dte result.
 * TA = double: the preconditioned step of the FP32 steps (PRE), rows of the
 * FP64 orbitals C1 rounded to FP32 as they are staged, X = hml:
 * C2 = C1 + Kp o (C2 - C1 + dte result), the update in FP64. */
template<int AKT, int WAVES, typename TA, int PRE>
__global__ void __launch_bounds__(512, WAVES)
k_kbapply_f32(const TA *__restrict__ Bf, size_t ldb, int Mr, const double *__restrict__ X, int xsm, int ldx,
              int Nc, size_t K, double *__restrict__ C2, size_t ldc, double dte, int ntn, int per,
              const double *__restrict__ Kp, const double *__restrict__ C1, int Mr2,
              const double *__restrict__ X2, int ldx2)
{
   /* PRE = 2 (TA = float): the KB apply and the preconditioned step of the
    * FP32 steps in one pass.  The reduction runs over the Mr projector rows
    * of Bf (padded to whole AM steps) and then the Mr2 orbital rows of C1
    * (FP64, rounded as they are staged) against hml X2 (X2[m + n ldx2]); the
    * epilogue is PRE's with result = B X + C1 hml. */
   const int Mr1 = PRE == 2 ? (Mr + AM - 1) / AM * AM : Mr;
   const int Mtot = PRE == 2 ? Mr1 + Mr2 : Mr;
   constexpr int BP = AKT + 4, JB = AKT / 64, WKE = AKT / 4, BH = AKT / 128;
   /* main loop: Bs[2][AM][BP] and Xs[2][AN][XP]; epilogue: T[64][BP], one
    * 64-column half of the output tile at a time */
   constexpr int BSZ = 2 * AM * BP, XSZ = 2 * AN * XP, TSZ = 64 * BP;
   __shared__ float lds[BSZ + XSZ > TSZ ? BSZ + XSZ : TSZ];
   float (*Bs)[AM][BP] = reinterpret_cast<float (*)[AM][BP]>(lds);
   float (*Xs)[AN][XP] = reinterpret_cast<float (*)[AN][XP]>(lds + BSZ);
   const int bid = blockIdx.x;
   const int g = per > 0 ? (bid & 7) * per + (bid >> 3) : bid;
   const size_t k0 = (size_t)(g / ntn) * AKT;
   const int n0 = (g % ntn) * AN;
   const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int wn = wave & 1, wk = wave >> 1;
   const int q = lane >> 4, rr = lane & 15;
   /* loaders: Bs row br, BH 16-byte pieces from bc (32 lanes per row); Xs
    * column xn, 4 projectors at xp */
   const int br = tid >> 5, bc = (tid & 31) * 4 * BH;
   const int xn = tid >> 2, xp = (tid & 3) * 4;

   f4 acc[4][JB];   /* [column block][k block] */
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < JB; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

   float4 rb[BH];
   double rx[4];
   auto gload = [&](int m0) {
      if (PRE == 2 && m0 >= Mr1) {
         /* orbital segment: rows of C1, X2 = hml */
         const int m = m0 - Mr1 + br;
#pragma unroll
         for (int h = 0; h < BH; ++h) {
            const size_t k = k0 + bc + 4 * h;
            rb[h] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (m < Mr2 && k < K) {
               const double2 u = *reinterpret_cast<const double2 *>(C1 + (size_t)m * ldc + k);
               const double2 v = *reinterpret_cast<const double2 *>(C1 + (size_t)m * ldc + k + 2);
               rb[h] = make_float4((float)u.x, (float)u.y, (float)v.x, (float)v.y);
            }
         }
         const int n = n0 + xn;
#pragma unroll
         for (int c = 0; c < 4; ++c) {
            const int mm = m0 - Mr1 + xp + c;
            rx[c] = (mm < Mr2 && n < Nc) ? X2[(size_t)mm + (size_t)n * ldx2] : 0.0;
         }
         return;
      }
      const int m = m0 + br;
#pragma unroll
      for (int h = 0; h < BH; ++h) {
         const size_t k = k0 + bc + 4 * h;
         if constexpr (sizeof(TA) == 4) {
            rb[h] = (m < Mr && k < K) ? *reinterpret_cast<const float4 *>(Bf + (size_t)m * ldb + k)
                                      : make_float4(0.f, 0.f, 0.f, 0.f);
         } else {
            rb[h] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (m < Mr && k < K) {
               const double2 u = *reinterpret_cast<const double2 *>(Bf + (size_t)m * ldb + k);
               const double2 v = *reinterpret_cast<const double2 *>(Bf + (size_t)m * ldb + k + 2);
               rb[h] = make_float4((float)u.x, (float)u.y, (float)v.x, (float)v.y);
            }
         }
      }
      const int n = n0 + xn;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
         const int mm = m0 + xp + c;
         rx[c] = (mm < Mr && n < Nc) ? X[(size_t)mm * xsm + (size_t)n * ldx] : 0.0;
      }
   };
   auto lstore = [&](int bf) {
#pragma unroll
      for (int h = 0; h < BH; ++h) *reinterpret_cast<float4 *>(&Bs[bf][br][bc + 4 * h]) = rb[h];
      *reinterpret_cast<float4 *>(&Xs[bf][xn][xp]) =
         make_float4((float)rx[0], (float)rx[1], (float)rx[2], (float)rx[3]);
   };
   /* slot q of instruction c takes projector 4q + c */
   auto mma = [&](int bf) {
      f4 x[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = *reinterpret_cast<const f4 *>(&Xs[bf][64 * wn + 16 * i + rr][4 * q]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
         float b[JB];
#pragma unroll
         for (int j = 0; j < JB; ++j) b[j] = Bs[bf][4 * q + c][WKE * wk + 16 * j + rr];
#pragma unroll
         for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j) acc[i][j] = mfma16(x[i][c], b[j], acc[i][j]);
      }
   };

   gload(0);
   lstore(0);
   __syncthreads();
   if (AM < Mtot) gload(AM);
   int buf = 0;
   for (int m0 = 0; m0 < Mtot; m0 += AM) {
      if (m0 + AM < Mtot) lstore(buf ^ 1);
      if (m0 + 2 * AM < Mtot) gload(m0 + 2 * AM);
      mma(buf);
      __syncthreads();
      buf ^= 1;
   }

   /* epilogue per 64-column half: the half's waves stage their sums in LDS
    * (lane (rr, q), register r holds column 16 i + 4 q + r, k = 16 j + rr
    * of the wave's range), then every wave updates whole rows of C2, AKT/64
    * reals per lane */
   float (*T)[BP] = reinterpret_cast<float (*)[BP]>(lds);
#pragma unroll
   for (int half = 0; half < 2; ++half) {
      if (half > 0) __syncthreads();
      if (wn == half) {
#pragma unroll
         for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j)
#pragma unroll
               for (int r = 0; r < 4; ++r) T[16 * i + 4 * q + r][WKE * wk + 16 * j + rr] = acc[i][j][r];
      }
      __syncthreads();
      const size_t k = k0 + JB * lane;
#pragma unroll 4
      for (int nl = wave; nl < 64; nl += 8) {
         const int n = n0 + 64 * half + nl;
         if (n < Nc && k < K) {
            double2 *p = reinterpret_cast<double2 *>(C2 + (size_t)n * ldc + k);
            if constexpr (JB == 4) {
               const float4 v = *reinterpret_cast<const float4 *>(&T[nl][4 * lane]);
               double2 a = p[0], b = p[1];
               a.x += dte * v.x; a.y += dte * v.y;
               b.x += dte * v.z; b.y += dte * v.w;
               p[0] = a;
               p[1] = b;
            } else if constexpr (PRE) {
               /* one complex coefficient per lane: k = 2 (k / 2) */
               const float2 v = *reinterpret_cast<const float2 *>(&T[nl][2 * lane]);
               const double *c1p = PRE == 2 ? C1 : reinterpret_cast<const double *>(Bf);
               const double2 c1 = *reinterpret_cast<const double2 *>(c1p + (size_t)n * ldc + k);
               const double f = Kp[k >> 1];
               double2 a = p[0];
               a.x = c1.x + f * (a.x - c1.x + dte * v.x);
               a.y = c1.y + f * (a.y - c1.y + dte * v.y);
               p[0] = a;
            } else {
               const float2 v = *reinterpret_cast<const float2 *>(&T[nl][2 * lane]);
               double2 a = p[0];
               a.x += dte * v.x; a.y += dte * v.y;
               p[0] = a;
            }
         }
      }
   }
}

/* 128 reals per workgroup: 32 accumulator registers per lane and 37 KB of
 * LDS let three workgroups share a CU (measured 14% faster than 256 reals
 * at two per CU on the 32-water cell: the C2 update of one workgroup
 * overlaps the products of the others) */
int launch_kbapply_f32(const float *Bf, int Mr, const double *X, int xsm, int ldx, int Nc, size_t K,
                       double *C2, size_t ldc, double dte, hipStream_t s)
{
   constexpr int AKT = 128;
   if (K % 8) { cpsd_set_error("kbapply_f32: K=%zu not a multiple of 8", K); return -1; }
   const int ntn = (Nc + AN - 1) / AN;
   const long blocks = (long)((K + AKT - 1) / AKT) * ntn;
   const int per = blocks % 8 == 0 ? (int)(blocks / 8) : 0;
   hipLaunchKernelGGL((k_kbapply_f32<AKT, KBAPPLY_WAVES, float, 0>), dim3((unsigned)blocks), dim3(512), 0, s, Bf, K,
                      Mr, X, xsm, ldx, Nc, K, C2, ldc, dte, ntn, per, nullptr, nullptr, 0, nullptr, 0);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* the preconditioned step of the FP32 steps: C2 = C1 + Kp o (C2 - C1 + dte
 * C1 H) with the product summed in FP32 over the FP64 orbitals rounded as
 * they are staged (C1, C2: [ne][K] reals; H: ne x ne column-major) */
int launch_precond_f32(const double *C1, int ne, const double *H, size_t K, double *C2, double dte,
                       const double *Kp, hipStream_t s)
{
   constexpr int AKT = 128;
   if (K % AKT) { cpsd_set_error("precond_f32: K=%zu not a multiple of %d", K, AKT); return -1; }
   const int ntn = (ne + AN - 1) / AN;
   const long blocks = (long)(K / AKT) * ntn;
   const int per = blocks % 8 == 0 ? (int)(blocks / 8) : 0;
   hipLaunchKernelGGL((k_kbapply_f32<AKT, KBAPPLY_WAVES, double, 1>), dim3((unsigned)blocks), dim3(512), 0, s, C1,
                      K, ne, H, 1, ne, ne, K, C2, K, dte, ntn, per, Kp, nullptr, 0, nullptr, 0);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* the KB apply and the preconditioned step of the FP32 steps in one pass:
 * C2 = C1 + Kp o (C2 - C1 + dte (B X + C1 H)), both products summed in one
 * FP32 accumulator (B: FP32 projectors [nproj][K]; X: nproj x ne with
 * X[m + n nproj]; C1, C2: [ne][K] reals; H: ne x ne column-major).  One read
 * and one write of C2 instead of two each. */
int launch_kbapply_precond_f32(const float *Bf, int nproj, const double *X, const double *C1, int ne,
                               const double *H, size_t K, double *C2, double dte, const double *Kp,
                               hipStream_t s)
{
   constexpr int AKT = 128;
   if (K % AKT) { cpsd_set_error("kbapply_precond_f32: K=%zu not a multiple of %d", K, AKT); return -1; }
   const int ntn = (ne + AN - 1) / AN;
   const long blocks = (long)(K / AKT) * ntn;
   const int per = blocks % 8 == 0 ? (int)(blocks / 8) : 0;
   hipLaunchKernelGGL((k_kbapply_f32<AKT, KBAPPLY_PRE_WAVES, float, 2>), dim3((unsigned)blocks), dim3(512), 0, s, Bf,
                      K, nproj, X, 1, nproj, ne, K, C2, K, dte, ntn, per, Kp, C1, ne, H, ne);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ------------------------------------------------ BF16-input projection */

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4 mfma_bf16(bf8 a, bf8 b, f4 c)
{
   return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

#define QK 32           /* K step: one 16x16x32 instruction per block */
#define QP (QK + 8)     /* row pitch (bf16): a 16-byte read of 16 rows at one k hits distinct banks */

/* round-to-nearest-even BF16 bits of a finite FP32 value */
__device__ __forceinline__ unsigned bf16_rne(float x)
{
   const unsigned u = __float_as_uint(x);
   return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

/* the split (hi, lo) of two FP32 values x, y: hi = bf16(x), lo = bf16(x -
 * hi), each pair packed as (x, y) in one 32-bit word */
__device__ __forceinline__ void split_pack(float x, float y, unsigned &hi, unsigned &lo)
{
   const unsigned hx = bf16_rne(x), hy = bf16_rne(y);
   const float rx = x - __uint_as_float(hx << 16), ry = y - __uint_as_float(hy << 16);
   hi = hx | (hy << 16);
   lo = bf16_rne(rx) | (bf16_rne(ry) << 16);
}

/* element offset of (row, col) in a tile of 32-bf16 (64-byte) rows whose
 * 16-byte pieces are XOR-swizzled by row: piece c of row r sits at piece
 * c ^ ((r >> 1) & 3).  A 16-byte read of 8 consecutive rows at one piece
 * then touches 8 distinct 16-byte bank groups (two rows share each 128-byte
 * bank line, rows r >> 1 apart are permuted), and the 4-byte writes of two
 * consecutive rows (r even) fill the 32 banks once. */
__device__ __forceinline__ int swz(int row, int col)
{
   return row * 32 + ((((col >> 3) ^ (row >> 1)) & 3) << 3) + (col & 7);
}

/* partial[z][p][n] = sum_{k in split z} w_k Bb[p][k] C1[n][k] with BF16
 * projectors Bb, C1 w rounded to BF16 as it is staged in LDS, FP32 sums;
 * the layout, grid and split reduction of k_kbproj_f32.  Lane l supplies 8
 * consecutive k of row l%16 to A and of column l%16 to B, the same k for
 * both, so the instruction's k order does not matter. */
template<int TM>
__global__ void __launch_bounds__(512, 4)
k_kbproj_bf16(const __bf16 *__restrict__ Bb, size_t ldb, int M, const double *__restrict__ C, size_t ldc,
              int N, const double *__restrict__ w, size_t K, size_t kchunk, int ntm, int ntile, int per,
              double *__restrict__ partial)
{
   constexpr int BI = TM / 32;
   __shared__ __bf16 As[2][TM][QP];
   __shared__ __bf16 Bs[2][PN][QP];
   const int bid = blockIdx.x;
   const long g = per > 0 ? (long)(bid & 7) * per + (bid >> 3) : bid;
   const int t = (int)(g % ntile), kz = (int)(g / ntile);
   const int tm = t % ntm, tn = t / ntm;
   const size_t kb = (size_t)kz * kchunk;
   const size_t ke = min(K, kb + kchunk);
   const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int wm = wave & 1, wn = wave >> 1;
   const int q = lane >> 4, rr = lane & 15;
   /* loaders: A row ar (8 bf16 at aq, 4 threads per 32-k row); B rows
    * lr + 32 h, k pair lq (16 threads per row) */
   const int ar = tid >> 2, aq = (tid & 3) * 8;
   const int lr = tid >> 4, lq = (tid & 15) * 2;

   f4 acc[BI][2];
#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

   uint4 ra;
   double2 rb[4], rw;
   auto gload = [&](size_t k0) {
      const int am = tm * TM + ar;
      ra = make_uint4(0u, 0u, 0u, 0u);
      if (ar < TM && am < M) ra = *reinterpret_cast<const uint4 *>(Bb + (size_t)am * ldb + k0 + aq);
#pragma unroll
      for (int h = 0; h < 4; ++h) {
         const int bn = tn * PN + lr + 32 * h;
         rb[h] = bn < N ? *reinterpret_cast<const double2 *>(C + (size_t)bn * ldc + k0 + lq)
                        : make_double2(0.0, 0.0);
      }
      rw = *reinterpret_cast<const double2 *>(w + k0 + lq);
   };
   auto lstore = [&](int bf) {
      if (ar < TM) *reinterpret_cast<uint4 *>(&As[bf][ar][aq]) = ra;
#pragma unroll
      for (int h = 0; h < 4; ++h) {
         __bf16 *d = &Bs[bf][lr + 32 * h][lq];
         d[0] = (__bf16)(float)(rb[h].x * rw.x);
         d[1] = (__bf16)(float)(rb[h].y * rw.y);
      }
   };
   auto mma = [&](int bf) {
      bf8 a[BI], b[2];
#pragma unroll
      for (int i = 0; i < BI; ++i) a[i] = *reinterpret_cast<const bf8 *>(&As[bf][(TM / 2) * wm + 16 * i + rr][8 * q]);
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf8 *>(&Bs[bf][32 * wn + 16 * j + rr][8 * q]);
#pragma unroll
      for (int i = 0; i < BI; ++i)
#pragma unroll
         for (int j = 0; j < 2; ++j) acc[i][j] = mfma_bf16(a[i], b[j], acc[i][j]);
   };

   if (kb < ke) {
      gload(kb);
      lstore(0);
   }
   __syncthreads();
   if (kb + QK < ke) gload(kb + QK);
   int buf = 0;
   for (size_t k0 = kb; k0 < ke; k0 += QK) {
      if (k0 + QK < ke) lstore(buf ^ 1);
      if (k0 + 2 * QK < ke) gload(k0 + 2 * QK);
      mma(buf);
      __syncthreads();
      buf ^= 1;
   }

#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
         const int n = tn * PN + 32 * wn + 16 * j + rr;
#pragma unroll
         for (int r = 0; r < 4; ++r) {
            const int p = tm * TM + (TM / 2) * wm + 16 * i + 4 * q + r;
            if (p < M && n < N) partial[((size_t)kz * M + p) * N + n] = (double)acc[i][j][r];
         }
      }
}

int launch_kbproj_bf16(const __bf16 *Bb, int M, const double *C, int N, const double *w, size_t K, int splits,
                       double *partial, size_t partial_cap, hipStream_t s)
{
   if (K % QK) { cpsd_set_error("kbproj_bf16: K=%zu not a multiple of %d", K, QK); return -1; }
   if (splits < 1 || (size_t)splits * M * N > partial_cap) {
      cpsd_set_error("kbproj_bf16: %d splits of %d x %d do not fit the partial buffer", splits, M, N);
      return -1;
   }
   const int pad128 = (M + 127) / 128 * 128 - M, pad96 = (M + 95) / 96 * 96 - M;
   const int tm = pad96 < pad128 ? 96 : 128;
   const int ntm = (M + tm - 1) / tm, ntile = ntm * ((N + PN - 1) / PN);
   size_t kchunk = (K + splits - 1) / splits;
   kchunk = (kchunk + QK - 1) / QK * QK;
   const long pairs = (long)ntile * splits;
   const int per = pairs % 8 == 0 ? (int)(pairs / 8) : 0;
   if (tm == 96)
      hipLaunchKernelGGL((k_kbproj_bf16<96>), dim3((unsigned)pairs), dim3(512), 0, s, Bb, K, M, C, K, N, w, K,
                         kchunk, ntm, ntile, per, partial);
   else
      hipLaunchKernelGGL((k_kbproj_bf16<128>), dim3((unsigned)pairs), dim3(512), 0, s, Bb, K, M, C, K, N, w, K,
                         kchunk, ntm, ntile, per, partial);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ------------------------------------ split-BF16 projection (mode 3) */

/* partial[z][p][n] = sum_{k in split z} w_k B[p][k] C1[n][k], both factors
 * split BF16 (B as BsH + BsL, made once per set of projectors; C1 w split as
 * it is staged), three v_mfma_f32_16x16x32_bf16 per block (hi hi, hi lo,
 * lo hi), FP32 sums: the layout, grid and split reduction of
 * k_kbproj_f32 */
template<int TM>
__global__ void __launch_bounds__(512, 4)
k_kbproj_dbf16(const __bf16 *__restrict__ BsH, const __bf16 *__restrict__ BsL, size_t ldb, int M,
               const double *__restrict__ C, size_t ldc, int N, const double *__restrict__ w, size_t K,
               size_t kchunk, int ntm, int ntile, int per, double *__restrict__ partial)
{
   constexpr int BI = TM / 32;
   constexpr int ASZ = TM * 32, BSZ = PN * 32;
   extern __shared__ __bf16 plds[];
   /* [buffer][hi, lo] */
   auto Aat = [&](int bf, int hl) { return plds + (size_t)(2 * bf + hl) * ASZ; };
   auto Bat = [&](int bf, int hl) { return plds + 4 * (size_t)ASZ + (size_t)(2 * bf + hl) * BSZ; };
   const int bid = blockIdx.x;
   const long g = per > 0 ? (long)(bid & 7) * per + (bid >> 3) : bid;
   const int t = (int)(g % ntile), kz = (int)(g / ntile);
   const int tm = t % ntm, tn = t / ntm;
   const size_t kb = (size_t)kz * kchunk;
   const size_t ke = min(K, kb + kchunk);
   const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int wm = wave & 1, wn = wave >> 1;
   const int q = lane >> 4, rr = lane & 15;
   /* loaders: A row ar (8 bf16 at aq, 4 threads per 32-k row); B rows
    * lr + 32 h, k pair lq (16 threads per row) */
   const int ar = tid >> 2, aq = (tid & 3) * 8;
   const int lr = tid >> 4, lq = (tid & 15) * 2;

   f4 acc[BI][2];
#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

   uint4 rah, ral;
   double2 rb[4], rw;
   auto gload = [&](size_t k0) {
      const int am = tm * TM + ar;
      rah = ral = make_uint4(0u, 0u, 0u, 0u);
      if (ar < TM && am < M) {
         rah = *reinterpret_cast<const uint4 *>(BsH + (size_t)am * ldb + k0 + aq);
         ral = *reinterpret_cast<const uint4 *>(BsL + (size_t)am * ldb + k0 + aq);
      }
#pragma unroll
      for (int h = 0; h < 4; ++h) {
         const int bn = tn * PN + lr + 32 * h;
         rb[h] = bn < N ? *reinterpret_cast<const double2 *>(C + (size_t)bn * ldc + k0 + lq)
                        : make_double2(0.0, 0.0);
      }
      rw = *reinterpret_cast<const double2 *>(w + k0 + lq);
   };
   auto lstore = [&](int bf) {
      if (ar < TM) {
         *reinterpret_cast<uint4 *>(Aat(bf, 0) + swz(ar, aq)) = rah;
         *reinterpret_cast<uint4 *>(Aat(bf, 1) + swz(ar, aq)) = ral;
      }
#pragma unroll
      for (int h = 0; h < 4; ++h) {
         /* the G-space weight (1 or 2, 0 on the padded tail) scales exactly */
         unsigned hi, lo;
         split_pack((float)rb[h].x * (float)rw.x, (float)rb[h].y * (float)rw.y, hi, lo);
         const int o = swz(lr + 32 * h, lq);
         *reinterpret_cast<unsigned *>(Bat(bf, 0) + o) = hi;
         *reinterpret_cast<unsigned *>(Bat(bf, 1) + o) = lo;
      }
   };
   auto mma = [&](int bf) {
      bf8 ah[BI], al[BI];
#pragma unroll
      for (int i = 0; i < BI; ++i) {
         const int o = swz((TM / 2) * wm + 16 * i + rr, 8 * q);
         ah[i] = *reinterpret_cast<const bf8 *>(Aat(bf, 0) + o);
         al[i] = *reinterpret_cast<const bf8 *>(Aat(bf, 1) + o);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
         const int o = swz(32 * wn + 16 * j + rr, 8 * q);
         const bf8 bh = *reinterpret_cast<const bf8 *>(Bat(bf, 0) + o);
         const bf8 bl = *reinterpret_cast<const bf8 *>(Bat(bf, 1) + o);
#pragma unroll
         for (int i = 0; i < BI; ++i) {
            acc[i][j] = mfma_bf16(al[i], bh, acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bl, acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bh, acc[i][j]);
         }
      }
   };

   if (kb < ke) {
      gload(kb);
      lstore(0);
   }
   __syncthreads();
   if (kb + QK < ke) gload(kb + QK);
   int buf = 0;
   for (size_t k0 = kb; k0 < ke; k0 += QK) {
      if (k0 + QK < ke) lstore(buf ^ 1);
      if (k0 + 2 * QK < ke) gload(k0 + 2 * QK);
      mma(buf);
      __syncthreads();
      buf ^= 1;
   }

#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
         const int n = tn * PN + 32 * wn + 16 * j + rr;
#pragma unroll
         for (int r = 0; r < 4; ++r) {
            const int p = tm * TM + (TM / 2) * wm + 16 * i + 4 * q + r;
            if (p < M && n < N) partial[((size_t)kz * M + p) * N + n] = (double)acc[i][j][r];
         }
      }
}

int launch_kbproj_dbf16(const __bf16 *BsH, const __bf16 *BsL, int M, const double *C, int N, const double *w,
                        size_t K, int splits, double *partial, size_t partial_cap, hipStream_t s)
{
   if (K % QK) { cpsd_set_error("kbproj_dbf16: K=%zu not a multiple of %d", K, QK); return -1; }
   if (splits < 1 || (size_t)splits * M * N > partial_cap) {
      cpsd_set_error("kbproj_dbf16: %d splits of %d x %d do not fit the partial buffer", splits, M, N);
      return -1;
   }
   const int pad128 = (M + 127) / 128 * 128 - M, pad96 = (M + 95) / 96 * 96 - M;
   const int tm = pad96 < pad128 ? 96 : 128;
   const int ntm = (M + tm - 1) / tm, ntile = ntm * ((N + PN - 1) / PN);
   size_t kchunk = (K + splits - 1) / splits;
   kchunk = (kchunk + QK - 1) / QK * QK;
   const long pairs = (long)ntile * splits;
   const int per = pairs % 8 == 0 ? (int)(pairs / 8) : 0;
   if (tm == 96) {
      constexpr int lds = 4 * (96 + PN) * 32 * (int)sizeof(__bf16);
      static std::atomic<unsigned long long> attr{0};
      if (dyn_lds_once((const void *)k_kbproj_dbf16<96>, lds, attr) != 0) return -1;
      hipLaunchKernelGGL((k_kbproj_dbf16<96>), dim3((unsigned)pairs), dim3(512), lds, s, BsH, BsL, K, M, C, K, N,
                         w, K, kchunk, ntm, ntile, per, partial);
   } else {
      constexpr int lds = 4 * (128 + PN) * 32 * (int)sizeof(__bf16);
      static std::atomic<unsigned long long> attr{0};
      if (dyn_lds_once((const void *)k_kbproj_dbf16<128>, lds, attr) != 0) return -1;
      hipLaunchKernelGGL((k_kbproj_dbf16<128>), dim3((unsigned)pairs), dim3(512), lds, s, BsH, BsL, K, M, C, K,
                         N, w, K, kchunk, ntm, ntile, per, partial);
   }
   CPSD_CHECK(hipGetLastError());
   return 0;
}

__global__ void __launch_bounds__(256) k_to_bf16(const double *__restrict__ a, __bf16 *__restrict__ b, size_t n)
{
   const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (i < n) b[i] = (__bf16)(float)a[i];
}

int launch_to_bf16(const double *a, __bf16 *b, size_t n, hipStream_t s)
{
   if (n == 0) return 0;
   hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ------------------------------------- overlaps of the step (split BF16) */

/* The overlaps of an FP32 step from its step d = c2 - c1 (ccpsdv3/cpsd.c:
 * 415-437 with c2 = c1 + d):
 *   <c2_i|c1_j> = <c1_i|c1_j> + <d_i|c1_j>,
 *   <c2_i|c2_j> = <c1_i|c1_j> + <d_i|c1_j> + <d_j|c1_i> + <d_i|d_j>,
 * with <c1|c1> the FP64 overlap carried from the previous step and the two
 * products of d in reduced precision: d is small (it vanishes at the fixed
 * point), so a relative error of the products is an error of size |d| in
 * the overlaps.  The products are split BF16: an FP32 factor x is x = hi +
 * lo + O(2^-17 |x|) with hi = bf16(x) and lo = bf16(x - hi), and
 *   x y ~ hi_x hi_y + hi_x lo_y + lo_x hi_y
 * (three v_mfma_f32_16x16x32_bf16, FP32 sums), about 2^-16 relative per
 * term at about five times the FP32 matrix rate, so the kernel is bound by
 * its reads of the FP64 orbitals.
 *
 * partial[z][i][n] = sum_{k in split z} d_i(k) w_k V_n(k), i < R, n < 2R,
 * V = [c1 | d]: one workgroup per (CB 128-column blocks, K split), every
 * workgroup holding all R rows of d, so that both operands come from the
 * one read of C1 and C2 rows (the column blocks' rows are among them).
 * Workgroups are assigned to the 8 XCDs so that those of one K range run
 * on one XCD and share its L2.  The d x d products are formed in full; the
 * caller reads one triangle. */
typedef double d2v __attribute__((ext_vector_type(2)));


template<int R, int CB, int NT>
__global__ void __launch_bounds__(NT, NT == 1024 ? 4 : 2)
k_gram_dbf16(const double *__restrict__ C1, const double *__restrict__ C2, size_t ldc,
             const double *__restrict__ w, size_t K, size_t kchunk, int per, double *__restrict__ partial)
{
   constexpr int NW = 2 * R / 128 / CB;            /* workgroups per K range, CB column blocks each */
   constexpr int NWV = NT / 64;                    /* waves */
   constexpr int RG = R / 64, CG = NWV / RG;       /* waves: RG row groups of 64 x CG column groups */
   constexpr int CW = 128 * CB / CG, BJ = CW / 16; /* columns per wave, 16-column blocks per wave */
   constexpr int LR = NT / 16, NP = R / LR;        /* rows per loader pass, passes */
   constexpr int PB = 128 / LR;                    /* loader passes per 128-row block */
   constexpr int ASZ = R * QP, BSZ = 128 * CB * QP;
   extern __shared__ __bf16 dlds[];
   /* [buffer][hi, lo]: A rows of d, B rows of w V */
   auto Aat = [&](int bf, int hl) { return dlds + (size_t)(2 * bf + hl) * ASZ; };
   auto Bat = [&](int bf, int hl) { return dlds + 4 * (size_t)ASZ + (size_t)(2 * bf + hl) * BSZ; };
   const int bid = blockIdx.x;
   const long g = per > 0 ? (long)(bid & 7) * per + (bid >> 3) : bid;
   const int blk0 = (int)(g % NW) * CB, kz = (int)(g / NW);
   const size_t kb = (size_t)kz * kchunk;
   const size_t ke = min(K, kb + kchunk);
   const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int wr = wave % RG, wc = wave / RG;
   const int q = lane >> 4, rr = lane & 15;
   /* loader: rows lr + LR h, k pair lq; 16 lanes cover one row's 32 k */
   const int lr = tid >> 4, lq = (tid & 15) * 2;

   f4 acc[4][BJ];
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

   /* C1 and C2 through buffer loads: one 32-bit byte offset per thread (its
    * first row and k pair) and the part common to the workgroup (pass row
    * and K step) in a scalar offset, so that the 2 NP loads of a K step
    * share one address register.  Every offset stays below 2^32 (R K 8
    * bytes, checked at launch). */
   const __amdgpu_buffer_rsrc_t rc1 = __builtin_amdgcn_make_buffer_rsrc((void *)C1, (short)0, (int)0xffffffff,
                                                                       0x00020000);
   const __amdgpu_buffer_rsrc_t rc2 = __builtin_amdgcn_make_buffer_rsrc((void *)C2, (short)0, (int)0xffffffff,
                                                                       0x00020000);
   const unsigned vb = 8u * ((unsigned)lr * (unsigned)ldc + (unsigned)lq);
   d2v r1[NP], r2[NP];
   double2 rw;
   auto gload = [&](size_t k0) {
#pragma unroll
      for (int h = 0; h < NP; ++h) {
         const unsigned so = (unsigned)(8 * ((size_t)(LR * h) * ldc + k0));
         r1[h] = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rc1, vb, so, 0));
         r2[h] = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rc2, vb, so, 0));
      }
      rw = *reinterpret_cast<const double2 *>(w + k0 + lq);
   };
   auto lstore = [&](int bf) {
#pragma unroll
      for (int h = 0; h < NP; ++h) {
         const int row = lr + LR * h;
         const double dx = r2[h].x - r1[h].x, dy = r2[h].y - r1[h].y;
         const float fx = (float)dx, fy = (float)dy;
         unsigned hi, lo;
         split_pack(fx, fy, hi, lo);
         *reinterpret_cast<unsigned *>(Aat(bf, 0) + row * QP + lq) = hi;
         *reinterpret_cast<unsigned *>(Aat(bf, 1) + row * QP + lq) = lo;
         /* the workgroup's column blocks that hold this orbital: block b is
          * c1 (b < R/128) or d, orbitals 128 (b mod R/128) on */
#pragma unroll
         for (int t = 0; t < CB; ++t) {
            const int b = blk0 + t;
            const bool bd = b >= R / 128;
            const int h0 = (b % (R / 128)) * PB;
            if (h >= h0 && h < h0 + PB) {
               /* the G-space weight (1 or 2, 0 on the padded tail) scales exactly */
               const float vx = (bd ? fx : (float)r1[h].x) * (float)rw.x;
               const float vy = (bd ? fy : (float)r1[h].y) * (float)rw.y;
               const int br = 128 * t + row - LR * h0;
               split_pack(vx, vy, hi, lo);
               *reinterpret_cast<unsigned *>(Bat(bf, 0) + br * QP + lq) = hi;
               *reinterpret_cast<unsigned *>(Bat(bf, 1) + br * QP + lq) = lo;
            }
         }
      }
   };
   auto mma = [&](int bf) {
      bf8 ah[4], al[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
         const int o = (64 * wr + 16 * i + rr) * QP + 8 * q;
         ah[i] = *reinterpret_cast<const bf8 *>(Aat(bf, 0) + o);
         al[i] = *reinterpret_cast<const bf8 *>(Aat(bf, 1) + o);
      }
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
         const int o = (CW * wc + 16 * j + rr) * QP + 8 * q;
         const bf8 bh = *reinterpret_cast<const bf8 *>(Bat(bf, 0) + o);
         const bf8 bl = *reinterpret_cast<const bf8 *>(Bat(bf, 1) + o);
#pragma unroll
         for (int i = 0; i < 4; ++i) {
            acc[i][j] = mfma_bf16(al[i], bh, acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bl, acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bh, acc[i][j]);
         }
      }
   };

   if (kb < ke) {
      gload(kb);
      lstore(0);
   }
   __syncthreads();
   if (kb + QK < ke) gload(kb + QK);
   int buf = 0;
   for (size_t k0 = kb; k0 < ke; k0 += QK) {
      if (k0 + QK < ke) lstore(buf ^ 1);
      if (k0 + 2 * QK < ke) gload(k0 + 2 * QK);
      mma(buf);
      __syncthreads();
      buf ^= 1;
   }

   /* 16 lanes hold 16 consecutive columns of one row: 128-byte segments */
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
         const int n = 128 * blk0 + CW * wc + 16 * j + rr;
#pragma unroll
         for (int r = 0; r < 4; ++r) {
            const int row = 64 * wr + 16 * i + 4 * q + r;
            partial[((size_t)kz * R + row) * (2 * R) + n] = (double)acc[i][j][r];
         }
      }
}

/* workgroups per CU the Gram of d is split into (its partials: splits x ne x
 * 2ne doubles, reduced in FP64 afterwards).  One round of the GPU measured
 * 0.05 ms faster per 64-water step than two (fewer partials to reduce) and
 * 0.2 ms faster than four. */
#ifndef GRAM_DBF16_WPC
#define GRAM_DBF16_WPC 1
#endif

/* dynamic LDS of k_gram_dbf16<R, CB>: two buffers of the hi and lo rows of
 * A (R rows) and B (128 CB rows) */
static constexpr int gram_dbf16_lds(int R, int CB)
{
   return 4 * (R + 128 * CB) * QP * (int)sizeof(__bf16);
}

int gram_dbf16_ok(int ne)
{
   return ne == 128 || ne == 256;
}

/* the buffer loads' 32-bit byte offsets span an orbital block */
int gram_dbf16_fits(int ne, size_t K)
{
   return (double)ne * (double)K * 8.0 < 4294967296.0;
}

/* 8 waves per workgroup (16 waves for the 256 x 256 tiles spilled 65 VGPRs
 * at the 128 a wave has then) */
template<int R, int CB>
static int launch_gram_dbf16_t(const double *C1, const double *C2, const double *w, size_t K, size_t kchunk,
                               long pairs, int per, double *partial, hipStream_t s)
{
   constexpr int lds = gram_dbf16_lds(R, CB);
   constexpr int NT = 512;
   static std::atomic<unsigned long long> attr{0};
   if (dyn_lds_once((const void *)k_gram_dbf16<R, CB, NT>, lds, attr) != 0) return -1;
   hipLaunchKernelGGL((k_gram_dbf16<R, CB, NT>), dim3((unsigned)pairs), dim3(NT), lds, s, C1, C2, K, w, K,
                      kchunk, per, partial);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* partial[z][i][n] of k_gram_dbf16 for the ne orbitals of C1 and C2 (rows of
 * K reals): two 128-column blocks per workgroup where the LDS holds their
 * buffers (every K range read by ne / 128 workgroups), else one;
 * GRAM_DBF16_WPC workgroups per CU over the K splits (long K ranges, few
 * partials); *splits receives their number */
int launch_gram_dbf16(const double *C1, const double *C2, int ne, const double *w, size_t K, double *partial,
                      size_t partial_cap, int *splits, hipStream_t s)
{
   if (!gram_dbf16_ok(ne)) { cpsd_set_error("gram_dbf16: ne=%d is not 128 or 256", ne); return -1; }
   if (K % QK) { cpsd_set_error("gram_dbf16: K=%zu not a multiple of %d", K, QK); return -1; }
   if (!gram_dbf16_fits(ne, K)) {
      cpsd_set_error("gram_dbf16: %d orbitals of %zu reals exceed 4 GiB", ne, K);
      return -1;
   }
   int dev = 0, maxlds = 0, cus = 0;
   CPSD_CHECK(hipGetDevice(&dev));
   CPSD_CHECK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
   CPSD_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
   const long wgs = (long)GRAM_DBF16_WPC * (cus > 0 ? cus : 1);
#ifdef GRAM_DBF16_CB
   const int cb = GRAM_DBF16_CB;   /* build-time override for measurement */
#else
   const int cb = gram_dbf16_lds(ne, 2) <= maxlds ? 2 : 1;
#endif
   const int nw = 2 * ne / 128 / cb;
   long sp = wgs / nw;
   const long maxsp = (long)(K / (QK * 8));
   const long capsp = (long)(partial_cap / (2 * (size_t)ne * ne));
   if (sp > maxsp) sp = maxsp;
   if (sp > capsp) sp = capsp;
   if (sp < 1) {
      cpsd_set_error("gram_dbf16: the partial buffer (%zu) holds no split of %d x %d", partial_cap, ne, 2 * ne);
      return -1;
   }
   size_t kchunk = (K + sp - 1) / sp;
   kchunk = (kchunk + QK - 1) / QK * QK;
   sp = (long)((K + kchunk - 1) / kchunk);   /* no empty split */
   const long pairs = (long)nw * sp;
   const int per = pairs % 8 == 0 ? (int)(pairs / 8) : 0;
   int rc;
   if (ne == 256) rc = cb == 2 ? launch_gram_dbf16_t<256, 2>(C1, C2, w, K, kchunk, pairs, per, partial, s)
                               : launch_gram_dbf16_t<256, 1>(C1, C2, w, K, kchunk, pairs, per, partial, s);
   else rc = cb == 2 ? launch_gram_dbf16_t<128, 2>(C1, C2, w, K, kchunk, pairs, per, partial, s)
                     : launch_gram_dbf16_t<128, 1>(C1, C2, w, K, kchunk, pairs, per, partial, s);
   if (rc != 0) return rc;
   *splits = (int)sp;
   return 0;
}

/* Gx (ne x 2ne, row-major: [<d_i|c1_j> | <d_i|d_j>]) and the carried
 * g11 = <c1|c1> (ne x ne, one triangle read) -> the C2 x C1, C1 x C2 and
 * C2 x C2 blocks of the Gram matrix Gm (2ne x 2ne, row-major, C1 first
 * unless c1_second), exactly symmetric where the overlaps are */
__global__ void __launch_bounds__(256) k_gram_dfix(const double *__restrict__ Gx, const double *__restrict__ g11,
                                                   int ne, int c1_second, double *__restrict__ Gm)
{
   const int idx = blockIdx.x * blockDim.x + threadIdx.x;
   if (idx >= ne * ne) return;
   const int i = idx / ne, j = idx % ne;
   const int lo = min(i, j), hi = max(i, j);
   const int n2 = 2 * ne, o1 = c1_second ? ne : 0, o2 = c1_second ? 0 : ne;
   const double s = g11[(size_t)lo * ne + hi];
   const double aij = Gx[(size_t)i * n2 + j], aji = Gx[(size_t)j * n2 + i];
   const double b = Gx[(size_t)lo * n2 + ne + hi];
   const double g21 = s + aij;
   Gm[(size_t)(o2 + i) * n2 + o1 + j] = g21;     /* <c2_i|c1_j> */
   Gm[(size_t)(o1 + j) * n2 + o2 + i] = g21;     /* <c1_j|c2_i> */
   Gm[(size_t)(o2 + i) * n2 + o2 + j] = s + (aij + aji) + b;
}

int launch_gram_dfix(const double *Gx, const double *g11, int ne, int c1_second, double *Gm, hipStream_t s)
{
   const size_t ee = (size_t)ne * ne;
   hipLaunchKernelGGL(k_gram_dfix, dim3((unsigned)((ee + 255) / 256)), dim3(256), 0, s, Gx, g11, ne, c1_second,
                      Gm);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

__global__ void __launch_bounds__(256) k_to_bf16x2(const double *__restrict__ a, unsigned *__restrict__ hi,
                                                   unsigned *__restrict__ lo, size_t npair)
{
   const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (i < npair) {
      const double2 v = *reinterpret_cast<const double2 *>(a + 2 * i);
      split_pack((float)v.x, (float)v.y, hi[i], lo[i]);
   }
}

/* BsH, BsL: the split-BF16 pair of n values (n even) */
int launch_to_bf16x2(const double *a, __bf16 *BsH, __bf16 *BsL, size_t n, hipStream_t s)
{
   if (n % 2) { cpsd_set_error("to_bf16x2: n=%zu is odd", n); return -1; }
   const size_t np = n / 2;
   if (np == 0) return 0;
   hipLaunchKernelGGL(k_to_bf16x2, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, a,
                      reinterpret_cast<unsigned *>(BsH), reinterpret_cast<unsigned *>(BsL), np);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* --------------------------- split-BF16 apply, update and constraint */

/* C2[n][k] (+)= dte sum_p XT[n][p] V[p][k] over P = npad + Mr2 rows: the
 * projector rows (p < npad: V = B, from BtH/BtL, the split projectors
 * transposed to [K][npad] once per set of projectors) and then the orbital
 * rows (p = npad + m: V = C1 row m, FP64, split and transposed as it is
 * staged); XT = [X^T | X2^T] split (k_split_xt).  Both factors split BF16,
 * three v_mfma_f32_16x16x32_bf16 per block, FP32 sums.
 *   PRE = 0: C2 += dte v (the KB apply, the constraint);
 *   PRE = 1: C2 = C1 + Kp o (C2 - C1 + dte v) (the preconditioned update).
 * One workgroup of 4 NO threads per 128 reals holds all NO = ne orbitals,
 * so that each V row is staged once (16 waves for 256 orbitals: four per
 * SIMD at 96 KB of LDS).  MFMA D[i][j] with i an orbital (A = XT rows, 8
 * consecutive p per lane) and j a real (B = the staged V^T rows, 8
 * consecutive p per lane); lane (rr, q), register r holds orbital 4q + r,
 * real rr of its block.  The sums are staged in LDS for the update of C2
 * (below), which reads and writes whole 1 KB row pieces. */
#define VK 128         /* reals per workgroup */

template<int NO, int PRE>
__global__ void __launch_bounds__(4 * NO, 4)
k_apply_dbf16(const __bf16 *__restrict__ BtH, const __bf16 *__restrict__ BtL, int npad,
              const double *__restrict__ C1, int Mr2, const __bf16 *__restrict__ XTh,
              const __bf16 *__restrict__ XTl, int P, size_t K, double *__restrict__ C2, double dte,
              const double *__restrict__ Kp, int per)
{
   constexpr int NTH = 4 * NO, NWV = NTH / 64;     /* threads, waves */
   constexpr int NG = NO / 64, KG = NWV / NG;      /* waves: NG orbital groups of 64 x KG real groups */
   constexpr int KW = VK / KG, BJ = KW / 16;       /* reals per wave, 16-real blocks per wave */
   constexpr int XSZ = NO * 32, VSZ = VK * 32;
   extern __shared__ __bf16 alds[];
   auto Xat = [&](int bf, int hl) { return alds + (size_t)(2 * bf + hl) * XSZ; };
   auto Vat = [&](int bf, int hl) { return alds + 4 * (size_t)XSZ + (size_t)(2 * bf + hl) * VSZ; };
   const int bid = blockIdx.x;
   const long g = per > 0 ? (long)(bid & 7) * per + (bid >> 3) : bid;
   const size_t kt = (size_t)g * VK;
   const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int wnn = wave % NG, wk = wave / NG;
   const int q = lane >> 4, rr = lane & 15;
   /* orbital-row loader: lane = 4 r + j covers rows m0 + 2r, m0 + 2r + 1 at
    * the k pairs 4 b + j, b = wave + NWV u (64-byte pieces of one row per 4
    * lanes); projector-row loader: piece c of the 1024 (hi, then lo) is
    * real (c % 512) / 4, rows 8 (c % 4) on */
   const int orr = lane >> 2, okp = lane & 3;
   constexpr int CU = 16 / NWV;                    /* orbital-row items per thread */
   constexpr int PU = 1024 / NTH;                  /* projector pieces per thread */

   f4 acc[4][BJ];
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

   uint4 rxh, rxl, rv[PU];
   double2 rc[2 * CU];
   auto gload = [&](int p0) {
      {
         const int n = tid >> 2, pc = (tid & 3) * 8;
         rxh = *reinterpret_cast<const uint4 *>(XTh + (size_t)n * P + p0 + pc);
         rxl = *reinterpret_cast<const uint4 *>(XTl + (size_t)n * P + p0 + pc);
      }
      if (p0 < npad) {
#pragma unroll
         for (int u = 0; u < PU; ++u) {
            const int c = tid + NTH * u, cc = c & 511;
            const __bf16 *src = c < 512 ? BtH : BtL;
            rv[u] = *reinterpret_cast<const uint4 *>(src + (kt + (cc >> 2)) * npad + p0 + (cc & 3) * 8);
         }
      } else {
         const int m = p0 - npad + 2 * orr;
#pragma unroll
         for (int u = 0; u < CU; ++u) {
            const size_t k = kt + 2 * (4 * (wave + NWV * u) + okp);
            rc[2 * u] = *reinterpret_cast<const double2 *>(C1 + (size_t)m * K + k);
            rc[2 * u + 1] = *reinterpret_cast<const double2 *>(C1 + (size_t)(m + 1) * K + k);
         }
      }
   };
   auto lstore = [&](int bf, int p0) {
      {
         const int n = tid >> 2, pc = (tid & 3) * 8;
         *reinterpret_cast<uint4 *>(Xat(bf, 0) + swz(n, pc)) = rxh;
         *reinterpret_cast<uint4 *>(Xat(bf, 1) + swz(n, pc)) = rxl;
      }
      if (p0 < npad) {
#pragma unroll
         for (int u = 0; u < PU; ++u) {
            const int c = tid + NTH * u, cc = c & 511;
            *reinterpret_cast<uint4 *>(Vat(bf, c >> 9) + swz(cc >> 2, (cc & 3) * 8)) = rv[u];
         }
      } else {
#pragma unroll
         for (int u = 0; u < CU; ++u) {
            const int k = 2 * (4 * (wave + NWV * u) + okp);
            unsigned h0, l0, h1, l1;
            /* (row 2r, row 2r + 1) at reals k and k + 1 */
            split_pack((float)rc[2 * u].x, (float)rc[2 * u + 1].x, h0, l0);
            split_pack((float)rc[2 * u].y, (float)rc[2 * u + 1].y, h1, l1);
            *reinterpret_cast<unsigned *>(Vat(bf, 0) + swz(k, 2 * orr)) = h0;
            *reinterpret_cast<unsigned *>(Vat(bf, 1) + swz(k, 2 * orr)) = l0;
            *reinterpret_cast<unsigned *>(Vat(bf, 0) + swz(k + 1, 2 * orr)) = h1;
            *reinterpret_cast<unsigned *>(Vat(bf, 1) + swz(k + 1, 2 * orr)) = l1;
         }
      }
   };
   auto mma = [&](int bf) {
      bf8 ah[4], al[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
         const int o = swz(64 * wnn + 16 * i + rr, 8 * q);
         ah[i] = *reinterpret_cast<const bf8 *>(Xat(bf, 0) + o);
         al[i] = *reinterpret_cast<const bf8 *>(Xat(bf, 1) + o);
      }
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
         const int o = swz(KW * wk + 16 * j + rr, 8 * q);
         const bf8 bh = *reinterpret_cast<const bf8 *>(Vat(bf, 0) + o);
         const bf8 bl = *reinterpret_cast<const bf8 *>(Vat(bf, 1) + o);
#pragma unroll
         for (int i = 0; i < 4; ++i) {
            acc[i][j] = mfma_bf16(al[i], bh, acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bl, acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bh, acc[i][j]);
         }
      }
   };

   gload(0);
   lstore(0, 0);
   __syncthreads();
   if (QK < P) gload(QK);
   int buf = 0;
   for (int p0 = 0; p0 < P; p0 += QK) {
      if (p0 + QK < P) lstore(buf ^ 1, p0 + QK);
      if (p0 + 2 * QK < P) gload(p0 + 2 * QK);
      mma(buf);
      __syncthreads();
      buf ^= 1;
   }

   /* epilogue by halves of NO/2 orbitals: the waves of the half stage their
    * sums in LDS ([n][VK + 4] FP32), then every wave updates whole 1 KB row
    * pieces, lane l the complex coefficient at reals 2l, 2l + 1, with the
    * loads of its HN / NWV rows issued before their stores (C2 is read and
    * written) */
   constexpr int HN = NO / 2, TP = VK + 4, ER = HN / NWV;
   float (*T)[TP] = reinterpret_cast<float (*)[TP]>(alds);
   const size_t kl = kt + 2 * lane;
   double f = 0.0;
   if (PRE) f = Kp[kl >> 1];
#pragma unroll
   for (int half = 0; half < 2; ++half) {
      if (half > 0) __syncthreads();
      if ((64 * wnn) / HN == half) {
#pragma unroll
         for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < BJ; ++j)
#pragma unroll
               for (int r = 0; r < 4; ++r)
                  T[64 * wnn - HN * half + 16 * i + 4 * q + r][KW * wk + 16 * j + rr] = acc[i][j][r];
      }
      __syncthreads();
      double2 c2v[ER], c1v[ER];
#pragma unroll
      for (int u = 0; u < ER; ++u) {
         const size_t o = (size_t)(HN * half + wave + NWV * u) * K + kl;
         c2v[u] = *reinterpret_cast<const double2 *>(C2 + o);
         if (PRE) c1v[u] = *reinterpret_cast<const double2 *>(C1 + o);
      }
#pragma unroll
      for (int u = 0; u < ER; ++u) {
         const int nl = wave + NWV * u;
         const float2 v = *reinterpret_cast<const float2 *>(&T[nl][2 * lane]);
         double2 a = c2v[u];
         if (PRE) {
            a.x = c1v[u].x + f * (a.x - c1v[u].x + dte * (double)v.x);
            a.y = c1v[u].y + f * (a.y - c1v[u].y + dte * (double)v.y);
         } else {
            a.x += dte * (double)v.x;
            a.y += dte * (double)v.y;
         }
         *reinterpret_cast<double2 *>(C2 + (size_t)(HN * half + nl) * K + kl) = a;
      }
   }
}

/* XTh, XTl [ne][P] (P = npad + Mr2): the split of X^T (X[p + n ldx], p <
 * nproj; zero for nproj <= p < npad) followed by X2^T (X2[m + n ldx2],
 * m < Mr2) */
__global__ void __launch_bounds__(256) k_split_xt(const double *__restrict__ X, int nproj, int ldx, int npad,
                                                  const double *__restrict__ X2, int Mr2, int ldx2, int ne,
                                                  unsigned *__restrict__ XTh, unsigned *__restrict__ XTl)
{
   const int P = npad + Mr2, hp = P / 2;
   const long idx = (long)blockIdx.x * 256 + threadIdx.x;
   if (idx >= (long)ne * hp) return;
   const int n = (int)(idx / hp), p = 2 * (int)(idx % hp);
   float v[2];
#pragma unroll
   for (int c = 0; c < 2; ++c) {
      const int pp = p + c;
      v[c] = pp < npad ? (pp < nproj ? (float)X[pp + (size_t)n * ldx] : 0.f)
                       : (float)X2[(pp - npad) + (size_t)n * ldx2];
   }
   split_pack(v[0], v[1], XTh[idx], XTl[idx]);
}

/* BtH, BtL [K][npad]: the split projectors transposed (B: [nproj][K] FP64;
 * zero rows nproj <= p < npad); one thread per (real, projector pair), the
 * reads of B coalesced along the reals */
__global__ void __launch_bounds__(256) k_split_bt(const double *__restrict__ B, int nproj, int npad, size_t K,
                                                  unsigned *__restrict__ BtH, unsigned *__restrict__ BtL)
{
   const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
   const int hp = npad / 2;
   if (idx >= K * hp) return;
   const size_t k = idx % K;
   const int p = 2 * (int)(idx / K);
   const float x = p < nproj ? (float)B[(size_t)p * K + k] : 0.f;
   const float y = p + 1 < nproj ? (float)B[(size_t)(p + 1) * K + k] : 0.f;
   split_pack(x, y, BtH[k * hp + p / 2], BtL[k * hp + p / 2]);
}

int apply_dbf16_ok(int ne)
{
   return ne == 128 || ne == 256;
}

int launch_split_bt(const double *B, int nproj, int npad, size_t K, __bf16 *BtH, __bf16 *BtL, hipStream_t s)
{
   if (npad % 32 || npad < nproj) { cpsd_set_error("split_bt: npad=%d for %d projectors", npad, nproj); return -1; }
   const size_t tot = K * (size_t)(npad / 2);
   if (tot == 0) return 0;
   hipLaunchKernelGGL(k_split_bt, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, B, nproj, npad, K,
                      reinterpret_cast<unsigned *>(BtH), reinterpret_cast<unsigned *>(BtL));
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* C2 (+)= dte (B^T X + C1^T X2) over the ne orbitals (ne = 128 or 256), the
 * products in split BF16: npad projector rows of the transposed split
 * projectors (0 for none; X[p + n ldx], p < nproj) and Mr2 orbital rows of C1
 * (0 for none; X2[m + n ldx2]); precond: C2 = C1 + Kp o (C2 - C1 + dte v).
 * XTh, XTl: ne (npad + Mr2) bf16 each of scratch.  C1, C2: [ne][K] reals,
 * K a multiple of 128. */
int launch_apply_dbf16(const __bf16 *BtH, const __bf16 *BtL, int npad, const double *X, int nproj, int ldx,
                       const double *C1, int Mr2, const double *X2, int ldx2, int ne, size_t K, double *C2,
                       double dte, const double *Kp, int precond, __bf16 *XTh, __bf16 *XTl, hipStream_t s)
{
   if (!apply_dbf16_ok(ne)) { cpsd_set_error("apply_dbf16: ne=%d is not 128 or 256", ne); return -1; }
   if (K % VK) { cpsd_set_error("apply_dbf16: K=%zu not a multiple of %d", K, VK); return -1; }
   if (npad % QK || Mr2 % QK || npad + Mr2 == 0 || npad < nproj) {
      cpsd_set_error("apply_dbf16: %d projector rows (%d padded) and %d orbital rows", nproj, npad, Mr2);
      return -1;
   }
   if (precond && !Kp) { cpsd_set_error("apply_dbf16: the preconditioned update needs Kp"); return -1; }
   const int P = npad + Mr2;
   const long nx = (long)ne * (P / 2);
   hipLaunchKernelGGL(k_split_xt, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, s, X, nproj, ldx, npad, X2,
                      Mr2, ldx2, ne, reinterpret_cast<unsigned *>(XTh), reinterpret_cast<unsigned *>(XTl));
   CPSD_CHECK(hipGetLastError());
   const long blocks = (long)(K / VK);
   const int per = blocks % 8 == 0 ? (int)(blocks / 8) : 0;
#define APPLY_DBF16(AN_, PRE_)                                                                              \
   do {                                                                                                     \
      constexpr int lds = 4 * (AN_ + VK) * 32 * (int)sizeof(__bf16);                                      \
      static std::atomic<unsigned long long> attr{0};                                                       \
      if (dyn_lds_once((const void *)k_apply_dbf16<AN_, PRE_>, lds, attr) != 0) return -1;                  \
      hipLaunchKernelGGL((k_apply_dbf16<AN_, PRE_>), dim3((unsigned)blocks), dim3(4 * AN_), lds, s, BtH,    \
                         BtL, npad, C1, Mr2, XTh, XTl, P, K, C2, dte, Kp, per);                             \
   } while (0)
   if (ne == 256) {
      if (precond) APPLY_DBF16(256, 1); else APPLY_DBF16(256, 0);
   } else {
      if (precond) APPLY_DBF16(128, 1); else APPLY_DBF16(128, 0);
   }
#undef APPLY_DBF16
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ----------------------------------------------------- FP32 projectors */

__global__ void __launch_bounds__(256) k_to_f32(const double *__restrict__ a, float *__restrict__ b, size_t n)
{
   const size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 2;
   if (i + 1 < n) {
      const double2 v = *reinterpret_cast<const double2 *>(a + i);
      *reinterpret_cast<float2 *>(b + i) = make_float2((float)v.x, (float)v.y);
   } else if (i < n) {
      b[i] = (float)a[i];
   }
}

int launch_to_f32(const double *a, float *b, size_t n, hipStream_t s)
{
   if (n == 0) return 0;
   const size_t blocks = (n + 511) / 512;
   hipLaunchKernelGGL(k_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, a, b, n);
   CPSD_CHECK(hipGetLastError());
   return 0;
}
