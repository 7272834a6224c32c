/*
 * gemm_kernels.hip -- FP64 MFMA contractions over the packed G vectors.
 *
 * The reference forms these as level-1 BLAS loops:
 *   - KB projections  sw = gcdotc(beta, c1_i)      ccpsdv3/cpsd.c:287-290
 *   - overlaps        gcdotc(c_i, c_j) x 4          ccpsdv3/cpsd.c:415-437
 *   - hml             -gcdotc(c1_i, cpsi_j)         ccpsdv3/cpsd.c:497-513
 *   - KB apply        c2_i += sw * beta (daxpy)     ccpsdv3/cpsd.c:290
 *   - constraint      c2 += dte c1 lmd (dgemm)      ccpsdv3/cpsd.c:482-488
 * Here they are two kernels built on v_mfma_f64_4x4x4_4b_f64, which on
 * gfx950 sustains ~74 TFLOP/s against ~36-49 for the 16x16x4 FP64 form:
 *   k_gemm_tn_splitk : small M x N output, reduction over the huge K = 2*ngp
 *                      real coefficients, split over blockIdx.z; 128 x 128
 *                      output tiles (8 waves of 64 x 32); the G-space weight
 *                      (1 on the kx = 0 plane, 2 elsewhere: the
 *                      2*sum - sum_{kx=0} of ccpsdv3/gcdotc.c:25-27) is
 *                      applied to the A operand.
 *   k_gemm_tall      : tall output [ncol][K] = X^T A with a short reduction
 *                      (projectors or orbitals), 64-column x 128-row output
 *                      tiles over 8 waves, SD epilogues fused; the nonlocal
 *                      force (ccpsdv3/cpsd.c:291-304) reduces its output
 *                      against the projectors instead of storing it.
 * The instruction layout and the rotated-operand scheme are described in
 * mfma64.h.
 */
#include "cpsd_internal.h"
#include "mfma64.h"
#include <type_traits>
#include <cstring>

/* ------------------------------------------------------- split-K TN */

/* the split-K kernel's rotated B operands: 1 reads each rotation from LDS
 * (the swizzled tiles keep those reads conflict-free), 0 forms them with DPP
 * row rotations of one read.  At the 64-water cell the LDS reads made the
 * carried Gram 4-6% faster; the same change in the tall kernel made it
 * 3-5% slower, so that one keeps the DPP form. */
#ifndef SPLITK_LDS_ROT
#define SPLITK_LDS_ROT 1
#endif
#define TT 128  /* output tile edge */
#define KT 32   /* K step */
#define TP (TT + 2)  /* row pitch of the epilogue staging tile */

/* Diagonal-pair roles (see k_gemm_tn_splitk).  A diagonal tile's 36 upper
 * 16 x 16 blocks are split between four waves: two take half of the upper
 * right quadrant each (block rows 2h, 2h+1 x columns 4-7, 8 blocks), two the
 * upper triangle of a diagonal quadrant (rows and columns o..o+3, o = 0 or
 * 4, 10 blocks).  Role 1 is the quadrant half, role 2 the triangle; block b
 * of a role multiplies A row block dg_ai and B column block dg_bj. */
__device__ constexpr int dg_nblk(int dm) { return dm == 1 ? 8 : 10; }
__device__ constexpr int dg_na(int dm) { return dm == 1 ? 2 : 4; }
__device__ constexpr int dg_ai(int dm, int b)
{
   return dm == 1 ? b >> 2 : (b < 4 ? 0 : b < 7 ? 1 : b < 9 ? 2 : 3);
}
__device__ constexpr int dg_bj(int dm, int b)
{
   return dm == 1 ? b & 3 : (b < 4 ? b : b < 7 ? b - 3 : b < 9 ? b - 5 : 3);
}

/* 8 waves, each a (TMR/2) x 32 block of the TMR x 128 tile (TMR = 128, or 96
 * when that pads the row count less, e.g. 192 projectors).  One workgroup per
 * (tile, K split); the grid is 1D over the valid pairs only.  Workgroups go
 * to the 8 XCDs round-robin by block id, so block b is given pair
 * (b % 8) * per + b / 8: the workgroups one XCD runs together hold
 * consecutive pairs, i.e. all tiles of the same K range, which then share
 * that XCD's L2 (per = pairs / 8; 0 falls back to b itself).
 *
 * A symmetric product (A = B, M = N) forms the tiles above the diagonal and
 * the diagonal tiles two at a time: a diagonal-pair workgroup loads the same
 * 256 rows as the off-diagonal tile (2d, 2d+1) and forms the 2 x 36 blocks on
 * or above the diagonals of tiles 2d and 2d+1 (waves 0-3 the first, 4-7 the
 * second), 18 blocks per SIMD against 16 for an off-diagonal tile, so the
 * workgroups of a launch take about the same time.
 *
 * With a tile list (tl.on), a symmetric product forms only the listed
 * off-diagonal tiles and the listed diagonal tiles, each diagonal tile by a
 * diagonal-pair workgroup whose two halves take the two halves of its K
 * range (rows of the same tile at k and at k + half) and write them as two
 * K splits, 2z and 2z + 1 (SELF instances only). */
template<int TMR, bool SELF>
__global__ void __launch_bounds__(512)
k_gemm_tn_splitk(const double *__restrict__ A, size_t lda, int M,
                 const double *__restrict__ B, size_t ldb, int N,
                 const double *__restrict__ w, size_t K, size_t kchunk, size_t kchunk_dd, int symm,
                 int ntm, int ntile, long poff, int per, GramTiles tl, double *__restrict__ partial)
{
   const int bid = blockIdx.x;
   const long g = per > 0 ? (long)(bid & 7) * per + (bid >> 3) : bid;
   int tm, tn, kz;
   bool dd = false, selfdd = false;
   size_t kc = kchunk;
   if (SELF && symm && tl.on) {
      if (g < poff) {
         const int t = (int)(g % tl.noff);
         kz = (int)(g / tl.noff);
         tm = tl.otm[t];
         tn = tl.otn[t];
      } else {
         const long h = g - poff;
         tm = tn = tl.dt[h % tl.nd];
         kz = (int)(h / tl.nd);
         dd = selfdd = true;
         kc = kchunk_dd;
      }
   } else if (symm) {
      const int noff = ntm * (ntm - 1) / 2, ndd = (ntm + 1) / 2;
      if (g < poff) {   /* tiles strictly above the diagonal, row by row */
         int rem = (int)(g % noff);
         kz = (int)(g / noff);
         tm = 0;
         while (rem >= ntm - 1 - tm) { rem -= ntm - 1 - tm; ++tm; }
         tn = tm + 1 + rem;
      } else {          /* diagonal pair (tm, tm + 1); tile ntm, if named, is empty */
         const long h = g - poff;
         tm = 2 * (int)(h % ndd);
         kz = (int)(h / ndd);
         tn = tm + 1;
         dd = true;
         kc = kchunk_dd;
      }
   } else {
      const int t = (int)(g % ntile);
      kz = (int)(g / ntile);
      tm = t % ntm;
      tn = t / ntm;
   }
   const size_t kb = (size_t)kz * kc;
   const size_t ke = min(K, kb + kc);
   /* a single diagonal tile's second half starts khalf after the first */
   const size_t khalf = (SELF && selfdd && ke > kb) ? ((ke - kb) / KT + 1) / 2 * KT : 0;
   const size_t kend = (SELF && selfdd) ? min(ke, kb + khalf) : ke;
   /* S[buf][0] holds the A tile, S[buf][1] the B tile, rows of KT doubles
    * with the k index XOR-swizzled by the row: element (row, k) sits at
    * column k ^ (row & 15).  An operand read takes 16 rows at one k per
    * 16-lane group; unswizzled (or with any even pad) those rows would share
    * bank pairs, swizzled they cover 16 distinct ones.  The loader's
    * 16-byte stores keep a (k, k+1) pair together: column pair
    * lq ^ (row & 14), its two values swapped when row is odd.  The same
    * memory stages the output tile for the epilogue. */
   __shared__ double Sbuf[2 * 2 * TT * KT > TT * TP ? 2 * 2 * TT * KT : TT * TP];
   double (*const S)[2][TT][KT] = reinterpret_cast<double (*)[2][TT][KT]>(Sbuf);
   __shared__ double Ws[2][SELF ? 2 : 1][KT];   /* weights of the A rows ([0]) and B rows ([1]) */
   /* the wave index is uniform: read it as a scalar */
   const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int wm = wave & 1, wn = wave >> 1;
#if SPLITK_LDS_ROT
   /* the rotated B operands read straight from LDS: rotation r takes column
    * 4 ((c / 4 + r) % 4) + c % 4 of the 16-column block (c = lane % 16) */
   const int s1 = 1;
   int rrow[4];
#pragma unroll
   for (int r = 0; r < 4; ++r) rrow[r] = 4 * ((((lane & 15) >> 2) + r) & 3) + (lane & 3);
#else
   const int s1 = rot_step(lane);
#endif
   /* loader: rows lr + 32 h (h < 4) of the tile, k pair 2 (tid % 16): 16
    * lanes cover one 256-byte row segment, contiguous in global memory and,
    * as 16-byte stores, conflict-free in LDS */
   const int lr = tid >> 4, lq = (tid & 15) * 2;
   constexpr int BI = TMR / 32;   /* 16-row blocks per wave */
   constexpr int BA = BI > 4 ? BI : 4;

   double acc[BA][2][4];   /* [row block][column block][rotation] */
   double accx[2][4];      /* blocks 8, 9 of a triangle role */
#pragma unroll
   for (int i = 0; i < BA; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
         for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
#pragma unroll
   for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) accx[j][r] = 0.0;

   /* K tile k+1 waits in registers while tile k is multiplied; it is moved
    * to the other LDS buffer halfway through tile k's MFMA stream and tile
    * k+2 is requested right after, so loads, LDS stores and MFMAs overlap
    * with one barrier per tile */
   double2 ra[4], rb[4], rw;
   /* the B rows sit khalf further along K (0 except for a single diagonal
    * tile), and are zero past the workgroup's K range */
   auto gload = [&](size_t k0) {
      if constexpr (SELF) {
         const bool bin = k0 + khalf < ke;
#pragma unroll
         for (int h = 0; h < 4; ++h) {
            const int r = lr + 32 * h, am = tm * TMR + r, bn = tn * TT + r;
            ra[h] = make_double2(0.0, 0.0);
            rb[h] = make_double2(0.0, 0.0);
            if (r < TMR && am < M) ra[h] = *reinterpret_cast<const double2 *>(A + (size_t)am * lda + k0 + lq);
            if (bn < N && bin) rb[h] = *reinterpret_cast<const double2 *>(B + (size_t)bn * ldb + k0 + khalf + lq);
         }
         if (tid < KT) {
            const int hb = tid / (KT / 2), kq = 2 * (tid % (KT / 2));
            rw = make_double2(0.0, 0.0);
            if (hb == 0 || bin) rw = *reinterpret_cast<const double2 *>(w + k0 + (hb ? khalf : 0) + kq);
         }
      } else {
#pragma unroll
         for (int h = 0; h < 4; ++h) {
            const int r = lr + 32 * h, am = tm * TMR + r, bn = tn * TT + r;
            ra[h] = make_double2(0.0, 0.0);
            rb[h] = make_double2(0.0, 0.0);
            if (r < TMR && am < M) ra[h] = *reinterpret_cast<const double2 *>(A + (size_t)am * lda + k0 + lq);
            if (bn < N) rb[h] = *reinterpret_cast<const double2 *>(B + (size_t)bn * ldb + k0 + lq);
         }
         if (tid < KT / 2) rw = *reinterpret_cast<const double2 *>(w + k0 + 2 * tid);
      }
   };
   /* rows lr + 32 h all have row & 15 == lr & 15 */
   const int sq = lq ^ (lr & 14);
   const bool swp = lr & 1;
   auto lstore = [&](int bf) {
#pragma unroll
      for (int h = 0; h < 4; ++h) {
         *reinterpret_cast<double2 *>(&S[bf][0][lr + 32 * h][sq]) = swp ? make_double2(ra[h].y, ra[h].x) : ra[h];
         *reinterpret_cast<double2 *>(&S[bf][1][lr + 32 * h][sq]) = swp ? make_double2(rb[h].y, rb[h].x) : rb[h];
      }
      if constexpr (SELF) {
         if (tid < KT) {
            const int hb = tid / (KT / 2), kq = 2 * (tid % (KT / 2));
            Ws[bf][hb][kq] = rw.x;
            Ws[bf][hb][kq + 1] = rw.y;
         }
      } else {
         if (tid < KT / 2) { Ws[bf][0][2 * tid] = rw.x; Ws[bf][0][2 * tid + 1] = rw.y; }
      }
   };

   /* diagonal pair: wave w works on tile half hs = w / 4 (operand S[.][hs]);
    * waves 0, 1, 6, 7 take role 1 (quadrant half 2h, 2h+1 with h = w % 2),
    * waves 2-5 role 2 (triangle o = 4 (w % 2)).  Block b is accumulated in
    * acc[b / 2][b % 2] (b < 8) or accx[b - 8]. */
   const int hs = wave >> 2;
   const int dm = !dd ? 0 : (wave == 0 || wave == 1 || wave == 6 || wave == 7) ? 1 : 2;
   const int drow0 = dm == 1 ? 2 * (wave & 1) : 4 * (wave & 1);   /* first A row block */
   const int dcol0 = dm == 1 ? 4 : drow0;                         /* first B column block */
   auto blk = [&](int b, int r) -> double & { return b < 8 ? acc[(b & 7) >> 1][b & 1][r] : accx[b & 1][r]; };
   auto mma_diag = [&](auto role, int bf, int ks) {
      constexpr int DM = decltype(role)::value;
      const int kk = 4 * ks + (lane >> 4), kx = kk ^ (lane & 15);
      const double wk = Ws[bf][SELF ? hs : 0][kk];
      double a[4], bv[4][4];
#pragma unroll
      for (int i = 0; i < dg_na(DM); ++i) a[i] = S[bf][hs][16 * (drow0 + i) + (lane & 15)][kx] * wk;
#if SPLITK_LDS_ROT
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
         for (int r = 0; r < 4; ++r) bv[j][r] = S[bf][hs][16 * (dcol0 + j) + rrow[r]][kk ^ rrow[r]];
#else
#pragma unroll
      for (int j = 0; j < 4; ++j) rotations(S[bf][hs][16 * (dcol0 + j) + (lane & 15)][kx], bv[j]);
#endif
#pragma unroll
      for (int b = 0; b < dg_nblk(DM); ++b)
#pragma unroll
         for (int r = 0; r < 4; ++r) blk(b, r) = mfma4(a[dg_ai(DM, b)], bv[dg_bj(DM, b)][r], blk(b, r));
   };
   auto mma_full = [&](int bf, int ks) {
      const int kk = 4 * ks + (lane >> 4), kx = kk ^ (lane & 15);
      const double wk = Ws[bf][0][kk];
      double a[BI];
#pragma unroll
      for (int i = 0; i < BI; ++i) a[i] = S[bf][0][(TMR / 2) * wm + 16 * i + (lane & 15)][kx] * wk;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
         double b[4];
#if SPLITK_LDS_ROT
#pragma unroll
         for (int r = 0; r < 4; ++r) b[r] = S[bf][1][32 * wn + 16 * j + rrow[r]][kk ^ rrow[r]];
#else
         rotations(S[bf][1][32 * wn + 16 * j + (lane & 15)][kx], b);
#endif
#pragma unroll
         for (int i = 0; i < BI; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = mfma4(a[i], b[r], acc[i][j][r]);
      }
   };

   /* the K loop, specialised per wave role so that no branch sits between
    * the unrolled MFMA steps */
   auto kloop = [&](auto role) {
      constexpr int DM = decltype(role)::value;
      if (kb < kend) {
         gload(kb);
         lstore(0);
      }
      __syncthreads();
      if (kb + KT < kend) gload(kb + KT);
      int buf = 0;
      for (size_t k0 = kb; k0 < kend; k0 += KT) {
#pragma unroll
         for (int ks = 0; ks < KT / 8; ++ks) DM ? mma_diag(role, buf, ks) : mma_full(buf, ks);
         if (k0 + KT < kend) {
            lstore(buf ^ 1);
            if (k0 + 2 * KT < kend) gload(k0 + 2 * KT);
         }
#pragma unroll
         for (int ks = KT / 8; ks < KT / 4; ++ks) DM ? mma_diag(role, buf, ks) : mma_full(buf, ks);
         __syncthreads();
         buf ^= 1;
      }
   };
   if (TMR != TT || dm == 0) kloop(std::integral_constant<int, 0>{});
   else if (dm == 1) kloop(std::integral_constant<int, 1>{});
   else kloop(std::integral_constant<int, 2>{});

   /* stage the tile in LDS, then write whole 1 KB rows of the partial */
   double (*T)[TP] = reinterpret_cast<double (*)[TP]>(Sbuf);
   const int q = lane >> 4, ab = (lane >> 2) & 3, bb = lane & 3;
   auto put = [&](int m, int n, int ml, int zz) {
      double *dst = partial + ((size_t)zz * M + m) * N + n;
      if (n + 1 < N && !(N & 1)) {
         *reinterpret_cast<double2 *>(dst) = make_double2(T[ml][2 * lane], T[ml][2 * lane + 1]);
      } else {
         dst[0] = T[ml][2 * lane];
         if (n + 1 < N) dst[1] = T[ml][2 * lane + 1];
      }
   };
   if (TMR == TT && dd) {
      auto stage = [&](auto role) {
         constexpr int DM = decltype(role)::value;
#pragma unroll
         for (int b = 0; b < dg_nblk(DM); ++b) {
            const int I = drow0 + dg_ai(DM, b), J = dcol0 + dg_bj(DM, b);
#pragma unroll
            for (int r = 0; r < 4; ++r) T[16 * I + 4 * ab + q][16 * J + 4 * ((ab + s1 * r) & 3) + bb] = blk(b, r);
         }
      };
      /* one diagonal tile at a time; only its upper 16 x 16 blocks exist */
#pragma unroll
      for (int h = 0; h < 2; ++h) {
         if (hs == h) {
            if (dm == 1) stage(std::integral_constant<int, 1>{});
            else stage(std::integral_constant<int, 2>{});
         }
         __syncthreads();
         const int t0 = (tm + ((SELF && selfdd) ? 0 : h)) * TT, n = t0 + 2 * lane;
         if (n < N) {
            for (int ml = wave; ml < TT; ml += 8) {
               const int m = t0 + ml;
               if (m >= M) break;
               if (2 * lane < 16 * (ml >> 4)) continue;
               put(m, n, ml, (SELF && selfdd) ? 2 * kz + h : kz);
            }
         }
         __syncthreads();
      }
      return;
   }
#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
         for (int r = 0; r < 4; ++r)
            T[(TMR / 2) * wm + 16 * i + 4 * ab + q][32 * wn + 16 * j + 4 * ((ab + s1 * r) & 3) + bb] =
               acc[i][j][r];
   __syncthreads();
   const int n = tn * TT + 2 * lane;
   if (n < N) {
#pragma unroll 4
      for (int ml = wave; ml < TMR; ml += 8) {
         const int m = tm * TMR + ml;
         if (m >= M) break;
         put(m, n, ml, kz);
      }
   }
}

/* K splits of a symmetric product: the tiles above the diagonal take sp_off
 * and the diagonal pairs sp_dd, in the ratio of their work per SIMD (16 : 18
 * blocks of 16 x 16), so that both kinds of workgroup take about as long;
 * the total is about what `splits` for every tile would give */
void gemm_tn_symm_splits(int M, size_t K, int splits, int *sp_off, int *sp_dd)
{
   const int ntm = (M + TT - 1) / TT;
   const int noff = ntm * (ntm - 1) / 2, ndd = (ntm + 1) / 2;
   const int kmax = (int)(K / KT) > 0 ? (int)(K / KT) : 1;
   if (noff == 0) {
      *sp_off = 0;
      *sp_dd = splits < kmax ? splits : kmax;
      return;
   }
   const long W = (long)(noff + ndd) * splits;
   int so = (int)(W * 16 / (16L * noff + 18L * ndd));
   if (so < 1) so = 1;
   int sd = (int)((W - (long)noff * so + ndd - 1) / ndd);
   if (sd < 1) sd = 1;
   *sp_off = so < kmax ? so : kmax;
   *sp_dd = sd < kmax ? sd : kmax;
}

/* tiles of one K split: all of them, or for a symmetric product those above
 * the diagonal plus the diagonal pairs */
int gemm_tn_tiles(int M, int N, int symm)
{
   if (symm) {
      const int ntm = (M + TT - 1) / TT;
      return ntm * (ntm - 1) / 2 + (ntm + 1) / 2;
   }
   const int pad128 = (M + 127) / 128 * 128 - M, pad96 = (M + 95) / 96 * 96 - M;
   const int tmr = pad96 < pad128 ? 96 : TT;
   return ((M + tmr - 1) / tmr) * ((N + TT - 1) / TT);
}


int launch_gemm_tn_splitk(const double *A, size_t lda, int M, const double *B, size_t ldb,
                          int N, const double *w, size_t K, int splits, int symm,
                          double *partial, hipStream_t s)
{
   if (K % KT) { cpsd_set_error("gemm_tn: K=%zu not a multiple of %d", K, KT); return -1; }
   if (symm && M != N) { cpsd_set_error("gemm_tn: symmetric product needs M == N"); return -1; }
   auto chunk = [&](int sp) {
      size_t kc = (K + sp - 1) / sp;
      return (kc + KT - 1) / KT * KT;
   };
   /* row tile: 128, or 96 when that leaves fewer padding rows (a symmetric
    * product keeps square tiles) */
   const int pad128 = (M + 127) / 128 * 128 - M, pad96 = (M + 95) / 96 * 96 - M;
   const int tmr = (!symm && pad96 < pad128) ? 96 : TT;
   const int ntm = (M + tmr - 1) / tmr;
   const int ntile = gemm_tn_tiles(M, N, symm);
   long pairs, poff = 0;
   size_t kchunk, kchunk_dd = 0;
   if (symm) {
      int so, sd;
      gemm_tn_symm_splits(M, K, splits, &so, &sd);
      const int noff = ntm * (ntm - 1) / 2, ndd = (ntm + 1) / 2;
      poff = (long)noff * so;
      pairs = poff + (long)ndd * sd;
      kchunk = so > 0 ? chunk(so) : 0;
      kchunk_dd = chunk(sd);
   } else {
      pairs = (long)ntile * splits;
      kchunk = chunk(splits);
   }
   const int per = pairs % 8 == 0 ? (int)(pairs / 8) : 0;
   GramTiles tl;
   memset(&tl, 0, sizeof tl);
   if (tmr == 96)
      hipLaunchKernelGGL((k_gemm_tn_splitk<96, false>), dim3((unsigned)pairs), dim3(512), 0, s, A, lda, M, B,
                         ldb, N, w, K, kchunk, kchunk_dd, symm, ntm, ntile, poff, per, tl, partial);
   else
      hipLaunchKernelGGL((k_gemm_tn_splitk<TT, false>), dim3((unsigned)pairs), dim3(512), 0, s, A, lda, M, B,
                         ldb, N, w, K, kchunk, kchunk_dd, symm, ntm, ntile, poff, per, tl, partial);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* K splits of a symmetric product restricted to the tiles of a list: the
 * off-diagonal tiles take sp_off splits and the diagonal tiles sp_dd, each of
 * the latter formed as two half-range splits (2 sp_dd partials).  A diagonal
 * workgroup does 18 blocks per SIMD over half its range against 16 over the
 * whole range for an off-diagonal one, so sp_dd = 9/16 sp_off balances them */
void gemm_tn_list_splits(int noff, int nd, size_t K, long wgs, int *sp_off, int *sp_dd)
{
   const int kmax = (int)(K / KT) > 0 ? (int)(K / KT) : 1;
   const int kmax2 = kmax / 2 > 0 ? kmax / 2 : 1;
   int so = noff > 0 ? (int)(16 * wgs / (16L * noff + 9L * nd)) : 0;
   if (noff > 0 && so < 1) so = 1;
   if (so > kmax) so = kmax;
   int sd = nd > 0 ? (int)((wgs - (long)noff * so + nd - 1) / nd) : 0;
   if (nd > 0 && sd < 1) sd = 1;
   if (sd > kmax2) sd = kmax2;
   *sp_off = so;
   *sp_dd = sd;
}

/* the Gram of [C1 C2] (M = 2 ne rows of stride lda, ne a multiple of 128)
 * without the C1 x C1 block: every 128 x 128 tile with a row or column in the
 * C2 block (rows [c2row, c2row + ne)) */
int launch_gram_c2(const double *A, size_t lda, int M, int c2row, const double *w, size_t K, long wgs,
                   double *partial, size_t partial_cap, int *sp_off, int *sp_dd, hipStream_t s)
{
   if (K % KT) { cpsd_set_error("gram: K=%zu not a multiple of %d", K, KT); return -1; }
   if (M % TT || c2row % TT || M > 8 * TT) {
      cpsd_set_error("gram: %d rows (C2 from row %d) are not whole 128-row tiles (at most 8)", M, c2row);
      return -1;
   }
   const int ntm = M / TT, c0 = c2row / TT, c1 = c0 + M / 2 / TT;
   auto in_c2 = [&](int t) { return t >= c0 && t < c1; };
   GramTiles tl;
   memset(&tl, 0, sizeof tl);
   tl.on = 1;
   for (int i = 0; i < ntm; ++i)
      for (int j = i + 1; j < ntm; ++j)
         if (in_c2(i) || in_c2(j)) { tl.otm[tl.noff] = (unsigned char)i; tl.otn[tl.noff] = (unsigned char)j; ++tl.noff; }
   for (int i = 0; i < ntm; ++i)
      if (in_c2(i)) tl.dt[tl.nd++] = (unsigned char)i;
   int so, sd;
   gemm_tn_list_splits(tl.noff, tl.nd, K, wgs, &so, &sd);
   const int zmax = so > 2 * sd ? so : 2 * sd;
   if ((size_t)zmax * M * M > partial_cap) {
      cpsd_set_error("gram: %d splits of %d x %d do not fit the partial buffer", zmax, M, M);
      return -1;
   }
   auto chunk = [&](int sp) {
      size_t kc = (K + sp - 1) / sp;
      return (kc + KT - 1) / KT * KT;
   };
   const long poff = (long)tl.noff * so;
   const long pairs = poff + (long)tl.nd * sd;
   const int per = pairs % 8 == 0 ? (int)(pairs / 8) : 0;
   const size_t kchunk = so > 0 ? chunk(so) : 0, kchunk_dd = chunk(sd);
   hipLaunchKernelGGL((k_gemm_tn_splitk<TT, true>), dim3((unsigned)pairs), dim3(512), 0, s, A, lda, M, A, lda, M,
                      w, K, kchunk, kchunk_dd, 1, ntm, 0, poff, per, tl, partial);
   CPSD_CHECK(hipGetLastError());
   *sp_off = so;
   *sp_dd = 2 * sd;
   return 0;
}

/* out[m][n] = sum_z partial[z][m][n] over the splits that formed (m, n):
 * `splits`, or for a symmetric product (symm = block edge of the formed
 * blocks) sp_dd inside the 128 x 128 diagonal tiles and `splits` elsewhere;
 * only blocks with block(m) <= block(n) exist and the rest is mirrored.  A
 * workgroup covers 64 consecutive outputs with four thread groups: group g
 * sums the splits z = g mod 4 in two interleaved chains, and the four group
 * sums are added in g order, so the result does not depend on scheduling. */
#define RW 16   /* waves per reduction workgroup, each over every RW-th split */

/* with X non-null (KB projections: out = P, nproj x ne), also the apply
 * coefficients X[m + n M] = -P[m][n] scal2 / pnrm[m] (ccpsdv3/cpsd.c:289) */
__global__ void __launch_bounds__(64 * RW) k_reduce_partials(const double *__restrict__ partial, int splits,
                                                         int sp_dd, int M, int N, int symm, int skip0,
                                                         int skip1, double *__restrict__ out,
                                                         const double *__restrict__ pnrm, double scal2,
                                                         double *__restrict__ X)
{
   __shared__ double red[RW][64];
   const int g = threadIdx.x >> 6, l = threadIdx.x & 63;
   const size_t idx = (size_t)blockIdx.x * 64 + l;
   const int m = (int)(idx / N), n = (int)(idx % N);
   /* symmetric: the blocks below the diagonal are written from their mirror;
    * rows and columns both in [skip0, skip1) were not formed */
   const bool valid = idx < (size_t)M * N && !(symm && (m / symm) > (n / symm)) &&
                      !(m >= skip0 && m < skip1 && n >= skip0 && n < skip1);
   const int nz = (symm && m / TT == n / TT) ? sp_dd : splits;
   /* wave g sums splits g, g + RW, ... into eight accumulators, so that eight
    * loads are in flight per lane, then adds them in a fixed tree */
   double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
   if (valid) {
      const size_t stride = (size_t)M * N;
      const double *p = partial + (size_t)m * N + n;
      int z = g;
      for (; z + 7 * RW < nz; z += 8 * RW) {
#pragma unroll
         for (int u = 0; u < 8; ++u) acc[u] += p[(size_t)(z + RW * u) * stride];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
         if (z + RW * u < nz) acc[u] += p[(size_t)(z + RW * u) * stride];
   }
   red[g][l] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
   __syncthreads();
   if (g == 0 && valid) {
      double v = red[0][l];
#pragma unroll
      for (int w = 1; w < RW; ++w) v += red[w][l];
      out[idx] = v;
      if (symm && (m / symm) < (n / symm)) out[(size_t)n * N + m] = v;
      if (X) X[m + (size_t)n * M] = -v * scal2 / pnrm[m];
   }
}

int launch_reduce_partials(const double *partial, int splits, int M, int N, double *out,
                           int symm, hipStream_t s)
{
   const size_t tot = (size_t)M * N;
   hipLaunchKernelGGL(k_reduce_partials, dim3((tot + 63) / 64), dim3(64 * RW), 0, s, partial,
                      splits, splits, M, N, symm, 0, 0, out, nullptr, 0.0, nullptr);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int launch_reduce_partials_kb(const double *partial, int splits, int nproj, int ne, double *P, const double *pnrm,
                              double scal2, double *X, hipStream_t s)
{
   const size_t tot = (size_t)nproj * ne;
   hipLaunchKernelGGL(k_reduce_partials, dim3((tot + 63) / 64), dim3(64 * RW), 0, s, partial,
                      splits, splits, nproj, ne, 0, 0, 0, P, pnrm, scal2, X);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int launch_reduce_partials_split(const double *partial, int sp_off, int sp_dd, int M, int skip0, int skip1,
                                 double *out, hipStream_t s)
{
   const size_t tot = (size_t)M * M;
   hipLaunchKernelGGL(k_reduce_partials, dim3((tot + 63) / 64), dim3(64 * RW), 0, s, partial,
                      sp_off, sp_dd, M, M, 16, skip0, skip1, out, nullptr, 0.0, nullptr);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int launch_reduce_partials_symm(const double *partial, int M, size_t K, int splits, double *out,
                                hipStream_t s)
{
   int so, sd;
   gemm_tn_symm_splits(M, K, splits, &so, &sd);
   const size_t tot = (size_t)M * M;
   hipLaunchKernelGGL(k_reduce_partials, dim3((tot + 63) / 64), dim3(64 * RW), 0, s, partial,
                      so, sd, M, M, 16, 0, 0, out, nullptr, 0.0, nullptr);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ------------------------------------------------- small M x N (TN) */

/* sum over the 64 lanes of a wave, same tree in every wave */
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
   for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
   return v;
}

/* out[m][n] = sum_k w[k] A[m][k] B[n][k] for small M, N: one 1024-thread
 * workgroup per output (per upper-triangle pair m <= n when symm, mirrored),
 * its 16 waves taking interleaved 16-byte pieces of K so that many loads are
 * in flight; wave sums by a fixed shuffle tree, then the 16 wave sums in
 * order: no split-K partials and no reduction launch */
__global__ void __launch_bounds__(1024) k_gemm_tn_small(const double *__restrict__ A, size_t lda, int M,
                                                        const double *__restrict__ B, size_t ldb, int N,
                                                        const double *__restrict__ w, size_t K, int symm,
                                                        double *__restrict__ out)
{
   __shared__ double ws[16];
   const int o = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
   int m, n;
   if (symm) {
      int r = o;
      m = 0;
      while (r >= M - m) { r -= M - m; ++m; }
      n = m + r;
   } else {
      m = o / N;
      n = o % N;
   }
   const double *a = A + (size_t)m * lda, *b = B + (size_t)n * ldb;
   double s0 = 0.0, s1 = 0.0;
#pragma unroll 2
   for (size_t k = 2 * (size_t)t; k < K; k += 2048) {
      const double2 av = *reinterpret_cast<const double2 *>(a + k);
      const double2 bv = *reinterpret_cast<const double2 *>(b + k);
      const double2 wk = *reinterpret_cast<const double2 *>(w + k);
      s0 += (wk.x * av.x) * bv.x;
      s1 += (wk.y * av.y) * bv.y;
   }
   const double v = wave_sum(s0 + s1);
   if (lane == 0) ws[wv] = v;
   __syncthreads();
   if (t == 0) {
      double tot = 0.0;
      for (int i = 0; i < 16; ++i) tot += ws[i];
      out[(size_t)m * N + n] = tot;
      if (symm) out[(size_t)n * N + m] = tot;
   }
}

int gemm_tn_small_ok(int M, int N) { return M <= 16 && N <= 16; }

int launch_gemm_tn_small(const double *A, size_t lda, int M, const double *B, size_t ldb, int N,
                         const double *w, size_t K, int symm, double *out, hipStream_t s)
{
   if (K % 2) { cpsd_set_error("gemm_tn_small: K=%zu is odd", K); return -1; }
   if (symm && M != N) { cpsd_set_error("gemm_tn_small: symmetric product needs M == N"); return -1; }
   const long outs = symm ? (long)M * (M + 1) / 2 : (long)M * N;
   hipLaunchKernelGGL(k_gemm_tn_small, dim3((unsigned)outs), dim3(1024), 0, s, A, lda, M, B, ldb, N, w, K, symm,
                      out);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ---------------------------------------------------------- tall NN */

#define TK 128  /* output rows (K) per workgroup */
#define MC 16   /* reduction step */


/* waves arranged WK (along K) x WN (along the output columns) = 8; at most
 * 128 VGPRs and 68 KB of LDS, so that two workgroups share a CU and one's
 * epilogue traffic overlaps the other's MFMA stream */
template<int EPI, int NT>
__global__ void __launch_bounds__(512, 4) k_gemm_tall(TallArgs a)
{
   constexpr int WN = NT / 64, WK = 8 / WN;
   constexpr int BI = TK / WK / 16;          /* 16-row blocks of K per wave */
   constexpr int BJ = 4;                     /* 16-column blocks per wave */
   constexpr int AP = TK + 16;   /* A panel pitch: a 64-bit read of 16 rows x 2 k is conflict-free */
   constexpr int AS = 2 * MC * AP, XS = 2 * NT * (MC + 1), TS = 64 * (TK + 2);
   constexpr int LD = AS + XS > TS ? AS + XS : TS;
   /* main loop: As[2][MC][AP] and Xs[2][NT][MC+1]; epilogue: T[64][TK+2],
    * one 64-column half of the output tile at a time */
   __shared__ double lds[LD];
   double (*As)[MC][AP] = reinterpret_cast<double (*)[MC][AP]>(lds);
   double (*Xs)[NT][MC + 1] = reinterpret_cast<double (*)[NT][MC + 1]>(lds + AS);
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int q = lane >> 4, ab = (lane >> 2) & 3, bb = lane & 3;
   const int wk = wave % WK, wn = wave / WK;
   const int s1 = rot_step(lane);
   /* block -> (K block, column tile) with the column tile fastest, remapped
    * so that the workgroups one XCD runs together (block id mod 8) hold
    * consecutive pairs: the column tiles of one K block then read its A
    * rows through the same L2 */
   const int bid = blockIdx.x;
   const int g = a.per > 0 ? (bid & 7) * a.per + (bid >> 3) : bid;
   const size_t k0 = (size_t)(g / a.ntn) * TK;
   const int n0 = (g % a.ntn) * NT;
   double acc[BI][BJ][4];   /* [K block][column block][rotation] */
#pragma unroll
   for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j)
#pragma unroll
         for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;

   /* loaders: A chunk MC x TK, rows ar + 8 h with 64 lanes per 1 KB row (two
    * 16-byte pieces per thread); X chunk NT x MC */
   constexpr int XQ = NT * MC / 512;
   static_assert(TK == 128 && MC == 16, "A loader covers 16 rows of 128");
   const int ar = tid >> 6, ac = (tid & 63) * 2;
   double2 ra[2];
   double rx[XQ];
   auto gload = [&](int m0) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
         const int m = m0 + ar + 8 * h;
         ra[h] = make_double2(0.0, 0.0);
         if (m < a.Mr) ra[h] = *reinterpret_cast<const double2 *>(a.A + (size_t)m * a.lda + k0 + ac);
      }
#pragma unroll
      for (int u = 0; u < XQ; ++u) {
         const int e = tid + 512 * u;
         const int n = e / MC, mm = e % MC;
         const int mg = m0 + mm, ng = n0 + n;
         rx[u] = (mg < a.Mr && ng < a.Nc) ? a.X[(size_t)mg * a.xsm + (size_t)ng * a.ldx] : 0.0;
      }
   };
   auto lstore = [&](int bf) {
#pragma unroll
      for (int h = 0; h < 2; ++h) *reinterpret_cast<double2 *>(&As[bf][ar + 8 * h][ac]) = ra[h];
#pragma unroll
      for (int u = 0; u < XQ; ++u) {
         const int e = tid + 512 * u;
         Xs[bf][e / MC][e % MC] = rx[u];
      }
   };
   auto mma = [&](int bf, int ks) {
      const int mm = 4 * ks + (lane >> 4);
      double x[BJ];
#pragma unroll
      for (int j = 0; j < BJ; ++j) x[j] = Xs[bf][64 * wn + 16 * j + (lane & 15)][mm];
#pragma unroll
      for (int i = 0; i < BI; ++i) {
         double b[4];
         rotations(As[bf][mm][(TK / WK) * wk + 16 * i + (lane & 15)], b);
#pragma unroll
         for (int j = 0; j < BJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = mfma4(x[j], b[r], acc[i][j][r]);
      }
   };

   /* chunk c+1 waits in registers while chunk c is multiplied and is moved
    * to the other LDS buffer halfway through, chunk c+2 requested right
    * after: one barrier per chunk */
   gload(0);
   lstore(0);
   __syncthreads();
   if (MC < a.Mr) gload(MC);
   int buf = 0;
   for (int m0 = 0; m0 < a.Mr; m0 += MC) {
#pragma unroll
      for (int ks = 0; ks < MC / 8; ++ks) mma(buf, ks);
      if (m0 + MC < a.Mr) {
         lstore(buf ^ 1);
         if (m0 + 2 * MC < a.Mr) gload(m0 + 2 * MC);
      }
#pragma unroll
      for (int ks = MC / 8; ks < MC / 4; ++ks) mma(buf, ks);
      __syncthreads();
      buf ^= 1;
   }

   /* epilogue per 64-column half: the half's waves stage their tile in LDS,
    * then all waves apply the update along whole 1 KB rows */
   double (*T)[TK + 2] = reinterpret_cast<double (*)[TK + 2]>(lds);
#pragma unroll
   for (int half = 0; half < WN; ++half) {
      if (half > 0) __syncthreads();
      if (wn == half) {
#pragma unroll
         for (int i = 0; i < BI; ++i)
#pragma unroll
            for (int j = 0; j < BJ; ++j)
#pragma unroll
               for (int r = 0; r < 4; ++r)
                  T[16 * j + 4 * ab + q][(TK / WK) * wk + 16 * i + 4 * ((ab + s1 * r) & 3) + bb] =
                     acc[i][j][r];
      }
      __syncthreads();
      constexpr int EU = EPI == EPI_NLFORCE ? 1 : 2;   /* the force reduction does not unroll */
#pragma unroll EU
      for (int nl = wave; nl < 64; nl += 8) {
         const int n = n0 + 64 * half + nl;
         if (n >= a.Nc) break;
         const size_t idx = (size_t)n * a.ldc + k0 + 2 * lane;
         const double2 v = make_double2(T[nl][2 * lane], T[nl][2 * lane + 1]);
         if (EPI == EPI_KBAPPLY) {
            double2 cp = *reinterpret_cast<const double2 *>(a.CP + idx);
            const double2 c1 = *reinterpret_cast<const double2 *>(a.C1 + idx);
            cp.x += v.x; cp.y += v.y;
            *reinterpret_cast<double2 *>(a.C2 + idx) = make_double2(c1.x + a.dte * cp.x, c1.y + a.dte * cp.y);
            if (a.write_cp) *reinterpret_cast<double2 *>(a.CP + idx) = cp;
         } else if (EPI == EPI_CONSTRAINT) {
            double2 c2 = *reinterpret_cast<const double2 *>(a.C2 + idx);
            c2.x += a.dte * v.x; c2.y += a.dte * v.y;
            *reinterpret_cast<double2 *>(a.C2 + idx) = c2;
         } else if (EPI == EPI_STORE) {
            *reinterpret_cast<double2 *>(a.C2 + idx) = v;
         } else if (EPI == EPI_PRECOND) {
            const double2 c1 = *reinterpret_cast<const double2 *>(a.C1 + idx);
            double2 c2 = *reinterpret_cast<const double2 *>(a.C2 + idx);
            const double f = a.Kp[(k0 >> 1) + lane];
            c2.x = c1.x + f * (c2.x - c1.x + a.dte * v.x);
            c2.y = c1.y + f * (c2.y - c1.y + a.dte * v.y);
            *reinterpret_cast<double2 *>(a.C2 + idx) = c2;
         } else {
            /* lane holds Z_n(G) for G = k0/2 + lane (ccpsdv3/cpsd.c:296-302 with
             * the orbital sum moved inside) */
            const size_t g = (k0 >> 1) + lane;
            const double2 b = a.Bp[(size_t)n * a.ngp + g];
            const double t = a.w[k0 + 2 * lane] * (v.y * b.x - v.x * b.y);
            const double f0 = wave_sum(a.Gp[g] * t);
            const double f1 = wave_sum(a.Gp[(size_t)a.ngp + g] * t);
            const double f2 = wave_sum(a.Gp[2 * (size_t)a.ngp + g] * t);
            if (lane == 0) {
               double *o = a.fpart + ((size_t)(k0 / TK) * a.Nc + n) * 3;
               o[0] = f0; o[1] = f1; o[2] = f2;
            }
         }
      }
   }
}

int launch_gemm_tall(int epi, const TallArgs &args, hipStream_t s)
{
   if (args.K % TK) { cpsd_set_error("gemm_tall: K=%zu not a multiple of %d", args.K, TK); return -1; }
   TallArgs a = args;
   a.ntn = (a.Nc + 63) / 64;
   const long blocks = (long)(a.K / TK) * a.ntn;
   a.per = blocks % 8 == 0 ? (int)(blocks / 8) : 0;
   if (epi == EPI_KBAPPLY) hipLaunchKernelGGL((k_gemm_tall<EPI_KBAPPLY, 64>), dim3((unsigned)blocks), dim3(512), 0, s, a);
   else if (epi == EPI_CONSTRAINT) hipLaunchKernelGGL((k_gemm_tall<EPI_CONSTRAINT, 64>), dim3((unsigned)blocks), dim3(512), 0, s, a);
   else if (epi == EPI_STORE) hipLaunchKernelGGL((k_gemm_tall<EPI_STORE, 64>), dim3((unsigned)blocks), dim3(512), 0, s, a);
   else if (epi == EPI_PRECOND) hipLaunchKernelGGL((k_gemm_tall<EPI_PRECOND, 64>), dim3((unsigned)blocks), dim3(512), 0, s, a);
   else hipLaunchKernelGGL((k_gemm_tall<EPI_NLFORCE, 64>), dim3((unsigned)blocks), dim3(512), 0, s, a);
   CPSD_CHECK(hipGetLastError());
   return 0;
}
