/*
 * fft_kernels.hip -- batched 3D real<->half-complex FFTs of the SD step on
 * gfx950, FP64, with FP32 variants of the orbital passes (the transforms
 * run on float2 in LDS and T; inputs, the density sum and the outputs stay
 * FP64).
 *
 * The 3D transform of cr_fft3b / rc_fft3f (fftpack3d/fftpack3d.f:22-96,
 * 107-196) is split into
 *   z pass   : one workgroup per (kx block, ky, orbital); the kz column
 *              block is staged in LDS.  G->r reads the packed masked
 *              coefficients directly (zero fill outside the sphere); r->G
 *              writes only masked coefficients, with the SD epilogue fused.
 *   y+x pass : one workgroup per (z plane, orbital [block]); the whole
 *              (ky, kx) plane of N x (N/2+1) complex lives in LDS (133 KB at
 *              N = 128), the y transforms and the real x transform run back
 *              to back, and the real-space work (density, V*psi) is fused
 *              between the inverse and forward x transforms.
 * The y and z passes run over kx = 0..N/2 before the x pass, as the
 * reference does, so the imaginary parts of kx = 0 and kx = N/2 are dropped
 * at the same point (the complex-to-real x transform); rc_fft3f's zeroing of
 * Im(kx=0) and Im(kx=N/2) is reproduced after the forward x transform.
 */
#include "cpsd_internal.h"
#include "fft_core.h"

/* waves per SIMD (the second __launch_bounds__ argument) the FP32 orbital
 * kernels are compiled for (the plane kernels at N >= 80 only); their
 * float2 tiles let LDS hold twice the FP64 kernels' workgroups per CU.  At
 * N = 96, 6 waves (3 workgroups per CU) speed up the V psi pass by 6%; the
 * density and z passes spill at 6 and run slower there. */
#ifndef DENS_WAVES_F32
#define DENS_WAVES_F32 4
#endif
#ifndef VPSI_WAVES_F32
#define VPSI_WAVES_F32 6
#endif
#ifndef ZPASS_WAVES_F32
#define ZPASS_WAVES_F32 4
#endif
/* waves per SIMD the FP64 G -> r z pass is sized for: at N = 128 two (one
 * workgroup per CU) without spills ran 11% faster than four with 84 bytes
 * per lane spilled */
#ifndef ZPASS_G2R_WAVES_F64
#define ZPASS_G2R_WAVES_F64(N) ((N) >= 128 ? 2 : 4)
#endif

/* ------------------------------------------------------------ z passes */

/* row-table entries (start, length, offset) of the N rows (ky, kz = 0..N-1)
 * a z-pass workgroup touches, staged once in LDS */
template<int N>
__device__ __forceinline__ void stage_rows(const RowTable &rows, int ky, int (*rt)[N])
{
   for (int kz = threadIdx.x; kz < N; kz += blockDim.x) {
      const int r = ky + N * kz;
      rt[0][kz] = rows.rstart[r];
      rt[1][kz] = rows.rlen[r];
      rt[2][kz] = rows.roff[r];
   }
   __syncthreads();
}

/* G -> r along kz for bper orbitals of one (kx block, ky) row: the gather of
 * orbital b+1 is issued into the registers it is loaded into while orbital
 * b is transformed and written, so each workgroup keeps one column block of
 * loads in flight */
/* The z passes' logical workgroup (x, y, z) = (kx block, ky row, orbitals).
 * The hardware hands workgroups to the 8 XCDs round-robin in linear order
 * (x fastest), so neighbouring kx blocks and ky rows, which share cache
 * lines of T and of the packed orbitals at their edges, would each fetch
 * them into another L2.  Logical ids are regrouped so that every XCD runs
 * a contiguous range of them; a grid whose size is not a multiple of 8
 * keeps the hardware order. */
__device__ __forceinline__ void zpass_block(int &x, int &y, int &z)
{
   const unsigned long nx = gridDim.x, ny = gridDim.y;
   const unsigned long tot = nx * ny * gridDim.z;
   unsigned long l = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
   if (tot % 8 == 0) l = (l & 7) * (tot / 8) + (l >> 3);
   x = (int)(l % nx);
   l /= nx;
   y = (int)(l % ny);
   z = (int)(l / ny);
}

template<int N, int SRC, typename V>
__global__ void __launch_bounds__(GridCfg<N>::ZT, sizeof(V) == 8 ? ZPASS_WAVES_F32 : ZPASS_G2R_WAVES_F64(N)) k_zpass_g2r(ZArgs a)
{
   using C = GridCfg<N>;
   constexpr int NH = N / 2 + 1, KXB = C::KXB, ZT = C::ZT;
   constexpr int EPT = (N * KXB + ZT - 1) / ZT;   /* elements per thread */
   __shared__ V sm[N * KXB];
   __shared__ int rt[3][N];
   int bx, ky, bz;
   zpass_block(bx, ky, bz);
   const int kx0 = bx * KXB;
   const int b0 = bz * a.bper, b1 = min(a.nbatch, b0 + a.bper);
   const int ncol = min(KXB, a.rows.nhc - kx0);
   /* twiddles are read from LDS (the barrier in stage_rows publishes them) */
   __shared__ V tws[N];
   for (int m = threadIdx.x; m < N; m += blockDim.x) tws[m] = cvt<V>(a.tw[m]);
   stage_rows<N>(a.rows, ky, rt);

   /* gather positions, the same for every orbital of the row */
   int pe[EPT];
#pragma unroll
   for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + ZT * q;
      const int kz = e / KXB, kxl = e % KXB, kx = kx0 + kxl;
      pe[q] = -1;
      if (e < N * KXB && kxl < ncol) {
         const int rs = rt[0][kz], rl = rt[1][kz];
         if (kx >= rs && kx < rs + rl) pe[q] = rt[2][kz] + kx - rs;
      }
   }
   V v[EPT];
   auto gather = [&](int b) {
      if (SRC == SRC_PACKED) {
         const double2 *src = a.src + (size_t)b * a.sstride;
#pragma unroll
         for (int q = 0; q < EPT; ++q) v[q] = cvt<V>(pe[q] >= 0 ? src[pe[q]] : make_double2(0.0, 0.0));
      } else {
#pragma unroll
         for (int q = 0; q < EPT; ++q) {
            double2 f = make_double2(0.0, 0.0);
            if (pe[q] >= 0) {
               const double2 vl = a.vall[pe[q]], dg = a.dng[pe[q]];
               const double c = a.vc[pe[q]];
               f = make_double2(vl.x * a.scal2 + dg.x * c, vl.y * a.scal2 + dg.y * c);
            }
            v[q] = cvt<V>(f);
         }
      }
   };
   if (b0 < b1) gather(b0);
   for (int b = b0; b < b1; ++b) {
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
         const int e = threadIdx.x + ZT * q;
         if (e < N * KXB) sm[e] = v[q];
      }
      __syncthreads();
      if (SRC == SRC_PACKED && b + 1 < b1) gather(b + 1);   /* the field pass runs one row */
      lds_fft_nat2tr<C::NY1, C::NY2, +1>(sm, ncol, 1, KXB, tws, 1);
      V *T = reinterpret_cast<V *>(a.T) + (size_t)b * N * N * NH;
      for (int e = threadIdx.x; e < N * KXB; e += blockDim.x) {
         const int s = e / KXB, kxl = e % KXB;
         if (kxl < ncol) {
            const int z = tr_index<C::NY1, C::NY2>(s);
            T[((size_t)z * N + ky) * NH + kx0 + kxl] = sm[e];
         }
      }
      __syncthreads();
   }
}

template<int N, int DST, typename V>
__global__ void __launch_bounds__(GridCfg<N>::ZT, sizeof(V) == 8 ? ZPASS_WAVES_F32 : 1) k_zpass_r2g(ZArgs a)
{
   using C = GridCfg<N>;
   constexpr int NH = N / 2 + 1, KXB = C::KXB, ZT = C::ZT;
   constexpr int EPT = (N * KXB + ZT - 1) / ZT;
   __shared__ V sm[N * KXB];
   __shared__ int rt[3][N];
   int bx, ky, b;
   zpass_block(bx, ky, b);
   const int kx0 = bx * KXB;
   const int ncol = min(KXB, a.rows.nhc - kx0);
   const V *T = reinterpret_cast<const V *>(a.T) + (size_t)b * N * N * NH;
   __shared__ V tws[N];
   for (int m = threadIdx.x; m < N; m += blockDim.x) tws[m] = cvt<V>(a.tw[m]);
   stage_rows<N>(a.rows, ky, rt);

   {
      V v[EPT];
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
         const int e = threadIdx.x + ZT * q;
         const int s = e / KXB, kxl = e % KXB;
         v[q] = cvt<V>(make_double2(0.0, 0.0));
         if (e < N * KXB && kxl < ncol) {
            const int z = tr_index<C::NY1, C::NY2>(s);
            v[q] = T[((size_t)z * N + ky) * NH + kx0 + kxl];
         }
      }
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
         const int e = threadIdx.x + ZT * q;
         if (e < N * KXB) sm[e] = v[q];
      }
   }
   __syncthreads();
   lds_fft_tr2nat<C::NY1, C::NY2, -1>(sm, ncol, 1, KXB, tws, 1);
   double2 *out = a.out + (size_t)b * a.ostride;
   int pe[EPT];
#pragma unroll
   for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + ZT * q;
      const int kz = e / KXB, kxl = e % KXB, kx = kx0 + kxl;
      pe[q] = -1;
      if (e < N * KXB && kxl < ncol) {
         const int rs = rt[0][kz], rl = rt[1][kz];
         if (kx >= rs && kx < rs + rl) pe[q] = rt[2][kz] + kx - rs;
      }
   }
   if (DST == DST_HPSI || DST == DST_SD) {
      /* kinetic term inputs, all loads issued first */
      double2 c[EPT];
      double t[EPT];
      const double2 *c1 = a.c1 + (size_t)b * a.ostride;
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
         c[q] = pe[q] >= 0 ? c1[pe[q]] : make_double2(0.0, 0.0);
         t[q] = pe[q] >= 0 ? a.tg[pe[q]] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
         if (pe[q] >= 0) {
            const double2 v = to_d2(sm[threadIdx.x + ZT * q]);
            const double2 h = make_double2(-a.scal1 * v.x - t[q] * c[q].x, -a.scal1 * v.y - t[q] * c[q].y);
            out[pe[q]] = DST == DST_SD ? make_double2(c[q].x + a.dte * h.x, c[q].y + a.dte * h.y) : h;
         }
      }
   } else {
#pragma unroll
      for (int q = 0; q < EPT; ++q)
         if (pe[q] >= 0) out[pe[q]] = to_d2(sm[threadIdx.x + ZT * q]);
   }
}

/* ------------------------------------------------------ plane helpers */

/* Z_k = (X_k + conj X_{M-k}) + i W^k (X_k - conj X_{M-k}), W = e^{2 pi i/N};
 * only Re X_0 and Re X_M are used (drfftb ignores their imaginary parts).
 * The pairs (k, M-k) with 0 < k < M/2 run in one loop and the self-paired
 * columns k = 0 and k = M/2 in a second one (they touch other elements, so
 * no barrier is needed between them): every wave then takes one path. */
template<int N, typename V>
__device__ __forceinline__ void c2r_pre(V *sm, const V *__restrict__ tw)
{
   using R = typename Cx<V>::R;
   constexpr int NH = N / 2 + 1, M = N / 2, KG = M / 2 - 1;
   static_assert(M % 2 == 0, "N/2 must be even");
   for (int it = threadIdx.x; it < N * KG; it += blockDim.x) {
      const int s = it / KG, k = 1 + it % KG;
      V *row = sm + s * NH;
      const V a = row[k], b = row[M - k];
      const V w = tw[k];
      const V wm = Cx<V>::mk(-w.x, w.y);  /* W^{M-k} = -conj(W^k) */
      row[k] = cadd(cadd(a, cconj(b)), cmuli(cmul(w, csub(a, cconj(b)))));
      row[M - k] = cadd(cadd(b, cconj(a)), cmuli(cmul(wm, csub(b, cconj(a)))));
   }
   for (int it = threadIdx.x; it < 2 * N; it += blockDim.x) {
      V *row = sm + (it >> 1) * NH;
      if (it & 1) {
         const V a = row[M / 2];
         const V w = tw[M / 2];
         row[M / 2] = cadd(cadd(a, cconj(a)), cmuli(cmul(w, csub(a, cconj(a)))));
      } else {
         const R x0 = row[0].x, xm = row[M].x;
         row[0] = Cx<V>::mk(x0 + xm, x0 - xm);
      }
   }
}

/* X_k = (Z_k + conj Z_{M-k})/2 - (i/2) W^{-k} (Z_k - conj Z_{M-k}), k = 0..M,
 * with Im X_0 = Im X_M = 0 exactly (fftpack3d.f:160-161); same split of the
 * work as c2r_pre. */
template<int N, typename V>
__device__ __forceinline__ void r2c_post(V *sm, const V *__restrict__ tw)
{
   using R = typename Cx<V>::R;
   constexpr int NH = N / 2 + 1, M = N / 2, KG = M / 2 - 1;
   for (int it = threadIdx.x; it < N * KG; it += blockDim.x) {
      const int s = it / KG, k = 1 + it % KG;
      V *row = sm + s * NH;
      const V za = row[k], zb = row[M - k];
      const V w = cconj(tw[k]);                        /* W^{-k} */
      const V wm = Cx<V>::mk(-tw[k].x, -tw[k].y);      /* W^{-(M-k)} = -W^k */
      const V sa = cadd(za, cconj(zb)), da = csub(za, cconj(zb));
      const V sb = cadd(zb, cconj(za)), db = csub(zb, cconj(za));
      const V ta = cmuli(cmul(w, da)), tb = cmuli(cmul(wm, db));
      row[k] = Cx<V>::mk(R(0.5) * (sa.x - ta.x), R(0.5) * (sa.y - ta.y));
      row[M - k] = Cx<V>::mk(R(0.5) * (sb.x - tb.x), R(0.5) * (sb.y - tb.y));
   }
   for (int it = threadIdx.x; it < 2 * N; it += blockDim.x) {
      V *row = sm + (it >> 1) * NH;
      if (it & 1) {
         const V z = row[M / 2];
         const V w = cconj(tw[M / 2]);
         const V sa = cadd(z, cconj(z)), ta = cmuli(cmul(w, csub(z, cconj(z))));
         row[M / 2] = Cx<V>::mk(R(0.5) * (sa.x - ta.x), R(0.5) * (sa.y - ta.y));
      } else {
         const V z0 = row[0];
         row[0] = Cx<V>::mk(z0.x + z0.y, R(0));
         row[M] = Cx<V>::mk(z0.x - z0.y, R(0));
      }
   }
}

template<int N, typename V>
__device__ __forceinline__ void load_plane(V *sm, const V *__restrict__ Tp, int nhload)
{
   constexpr int NH = N / 2 + 1, NT = GridCfg<N>::PLANE_THREADS;
   constexpr int EPT = (N * NH + NT - 1) / NT;
   V v[EPT];
#pragma unroll
   for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + NT * q;
      v[q] = (e < N * NH && e % NH < nhload) ? Tp[e] : Cx<V>::mk(0, 0);
   }
#pragma unroll
   for (int q = 0; q < EPT; ++q) {
      const int e = threadIdx.x + NT * q;
      if (e < N * NH) sm[e] = v[q];
   }
   __syncthreads();
}

/* y inverse (natural ky -> transposed y), then x complex-to-real */
template<int N, typename V>
__device__ __forceinline__ void inverse_yx(V *sm, const V *__restrict__ tw)
{
   using C = GridCfg<N>;
   constexpr int NH = N / 2 + 1;
   lds_fft_nat2tr<C::NY1, C::NY2, +1>(sm, NH, 1, NH, tw, 1);
   c2r_pre<N>(sm, tw);
   __syncthreads();
   lds_fft_nat2tr<C::NX1, C::NX2, +1>(sm, N, NH, 1, tw, 2);
}

/* x real-to-complex then y forward (transposed y -> natural ky) */
template<int N, typename V>
__device__ __forceinline__ void forward_xy(V *sm, const V *__restrict__ tw)
{
   using C = GridCfg<N>;
   constexpr int NH = N / 2 + 1;
   lds_fft_tr2nat<C::NX1, C::NX2, -1>(sm, N, NH, 1, tw, 2);
   r2c_post<N>(sm, tw);
   __syncthreads();
   lds_fft_tr2nat<C::NY1, C::NY2, -1>(sm, NH, 1, NH, tw, 1);
}

/* real-space position of complex slot m of row slot s (x of the pair's first
 * real); row slots and complex slots are in transposed order after the
 * inverse passes */
template<int N>
__device__ __forceinline__ void slot_xy(int s, int m, int &y, int &x0)
{
   using C = GridCfg<N>;
   y = tr_index<C::NY1, C::NY2>(s);
   x0 = 2 * tr_index<C::NX1, C::NX2>(m);
}

/* ------------------------------------------------------- plane kernel */

/* waves per SIMD the LDS lets a plane kernel hold: whole workgroups of the
 * plane plus its twiddles in 160 KB */
template<int N, typename V>
constexpr int plane_lds_waves()
{
   constexpr long bytes = ((long)N * (N / 2 + 1) + N) * (long)sizeof(V);
   constexpr int wg = (int)(160L * 1024 / bytes);
   return wg * (GridCfg<N>::PLANE_THREADS / 64) / 4;
}

/* waves per SIMD the plane kernel's registers are sized for: 4 for the
 * FP64 density pass at N = 96, DENS_WAVES_F32 / VPSI_WAVES_F32 for the FP32
 * orbital passes at N >= 80, but never more than the LDS admits (at
 * N = 128 the FP32 planes fit two workgroups, 4 waves per SIMD: sizing the
 * registers for 6 only spilled them), and 2 for an FP32 density pass that
 * prefetches (plane_prefetch_mode) */
/* the FP64 plane passes (density, V psi) fetch the next orbital's plane
 * into registers while the current one is transformed when a plane is
 * larger than PLANE_PF_KB KB: such planes leave one workgroup per CU, whose
 * HBM reads would otherwise wait on its LDS transforms.  The 75 KB planes
 * of 96^3 fit two workgroups per CU, which overlap each other: prefetching
 * there (one workgroup per CU for the registers) measured slower, V psi
 * 0.51 -> 0.69 ms on 32 waters. */
#ifndef PLANE_PF_KB
#define PLANE_PF_KB 80
#endif
template<int N, typename V>
constexpr bool plane_prefetch()
{
   return sizeof(V) == 16 && (size_t)N * (N / 2 + 1) * sizeof(V) > (size_t)PLANE_PF_KB * 1024;
}
/* the FP32 density pass at N >= 112: one workgroup per CU with the next
 * orbital's plane prefetched, as in FP64, instead of two workgroups whose
 * 128 registers per lane spill the FP64 density sums (17 VGPRs spilled at
 * N = 128).  At 128^3 and 256 orbitals: 1.11 -> 0.99 ms per pass.  The
 * FP32 V psi pass keeps two workgroups per CU: prefetching it there
 * measured 1.61 -> 2.27 ms. */
#ifndef DENS_PF_F32
#define DENS_PF_F32 1
#endif
template<int N, int MODE, typename V>
constexpr bool plane_prefetch_mode()
{
   return plane_prefetch<N, V>() || (DENS_PF_F32 && sizeof(V) == 8 && N >= 112 && MODE == PM_DENSITY);
}

template<int N, int MODE, typename V>
constexpr int plane_min_waves()
{
   constexpr int lw = plane_lds_waves<N, V>() > 0 ? plane_lds_waves<N, V>() : 1;
   constexpr int want = (sizeof(V) == 8 && plane_prefetch_mode<N, MODE, V>()) ? 2
                      : (sizeof(V) == 8 && N >= 80 && MODE == PM_DENSITY) ? DENS_WAVES_F32
                      : (sizeof(V) == 8 && N >= 80 && MODE == PM_VPSI) ? VPSI_WAVES_F32
                      : (N == 96 && MODE == PM_DENSITY && !plane_prefetch<N, V>()) ? 4 : 1;
   return want < lw ? want : lw;
}

extern __shared__ __align__(16) unsigned char plane_lds[];

template<int N, int MODE, typename V>
__global__ void __launch_bounds__(GridCfg<N>::PLANE_THREADS, (plane_min_waves<N, MODE, V>())) k_plane(PlaneArgs a)
{
   using C = GridCfg<N>;
   using R = typename Cx<V>::R;
   constexpr int NH = N / 2 + 1, M = N / 2, NT = C::PLANE_THREADS;
   V *sm = reinterpret_cast<V *>(plane_lds);
   /* the V psi pass runs the orbital groups of one z plane in consecutive
    * workgroups (blockIdx.x the group), which the 8 XCDs take round-robin,
    * so the plane's potential is fetched by few L2s at about the same time;
    * the other modes keep blockIdx.x = z */
   const int z = MODE == PM_VPSI ? blockIdx.y : blockIdx.x;
   const size_t plane = (size_t)N * NH;
   /* twiddles live in LDS after the plane; every mode passes a barrier
    * before its first transform */
   V *tws = sm + plane;
   for (int m = threadIdx.x; m < N; m += blockDim.x) tws[m] = cvt<V>(a.tw[m]);
   V *Tb = reinterpret_cast<V *>(a.T);

   if constexpr (MODE == PM_DENSITY) {
      constexpr int ACC = (N * N + NT - 1) / NT;
      double acc[ACC];
#pragma unroll
      for (int q = 0; q < ACC; ++q) acc[q] = 0.0;
      const int bb0 = a.b0 + blockIdx.y * a.nb;
      const int bb1 = min(bb0 + a.nb, a.nbatch);
      auto accumulate = [&]() {
         const R *smr = reinterpret_cast<const R *>(sm);
#pragma unroll
         for (int q = 0; q < ACC; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < N * N) {
               const int s = p / N, j = p % N;
               const double v = smr[(size_t)s * 2 * NH + j];
               acc[q] += v * v;
            }
         }
         __syncthreads();
      };
      /* as in the V psi pass: an FP64 plane that fills more than half the
       * LDS leaves one workgroup per CU, so the next orbital's plane is
       * fetched into registers while the current one is transformed */
      constexpr bool PF = plane_prefetch_mode<N, MODE, V>();
      if constexpr (PF) {
         constexpr int EPT = (N * NH + NT - 1) / NT;
         V nx[EPT];
         auto fetch = [&](int b) {
            const V *Tq = Tb + ((size_t)b * N + z) * plane;
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
               const int e = threadIdx.x + NT * q;
               nx[q] = (e < N * NH && e % NH < a.nhload) ? Tq[e] : Cx<V>::mk(0, 0);
            }
         };
         if (bb0 < bb1) fetch(bb0);
         for (int b = bb0; b < bb1; ++b) {
            /* the previous orbital's sums have read the LDS plane */
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
               const int e = threadIdx.x + NT * q;
               if (e < N * NH) sm[e] = nx[q];
            }
            __syncthreads();
            if (b + 1 < bb1) fetch(b + 1);
            inverse_yx<N>(sm, tws);
            accumulate();
         }
      } else {
         for (int b = bb0; b < bb1; ++b) {
            load_plane<N>(sm, Tb + ((size_t)b * N + z) * plane, a.nhload);
            inverse_yx<N>(sm, tws);
            accumulate();
         }
      }
      double *dst = a.dnpart + ((size_t)blockIdx.y * N + z) * N * N;
#pragma unroll
      for (int q = 0; q < ACC; ++q) {
         const int p = threadIdx.x + q * NT;
         if (p < N * N) {
            const int s = p / N, j = p % N;
            int y, x0;
            slot_xy<N>(s, j >> 1, y, x0);
            double *d = dst + y * N + x0 + (j & 1);
            *d = a.accumulate ? *d + acc[q] : acc[q];
         }
      }
   } else if constexpr (MODE == PM_VTOT || MODE == PM_C2R_OUT) {
      const int b = blockIdx.y;
      load_plane<N>(sm, Tb + ((size_t)b * N + z) * plane, a.nhload);
      inverse_yx<N>(sm, tws);
      for (int it = threadIdx.x; it < N * M; it += blockDim.x) {
         const int s = it / M, m = it % M;
         int y, x0;
         slot_xy<N>(s, m, y, x0);
         const double2 v = to_d2(sm[s * NH + m]);
         if (MODE == PM_VTOT) {
            const size_t g = ((size_t)z * N + y) * N + x0;
            a.Vout[g] = v.x + a.xcp[g];
            a.Vout[g + 1] = v.y + a.xcp[g + 1];
         } else {
            double *row = a.rout + (size_t)b * N * N * (N + 2) + ((size_t)z * N + y) * (N + 2);
            row[x0] = v.x;
            row[x0 + 1] = v.y;
            if (m == 0) { row[N] = 0.0; row[N + 1] = 0.0; }
         }
      }
   } else if constexpr (MODE == PM_VPSI) {
      /* orbitals [bb0, bb1) of plane z; the potential is loaded once for
       * all of them.  When an FP64 plane fills more than half the LDS (one
       * workgroup per CU, N >= 112), the next orbital's plane is fetched
       * into registers while the current one is transformed: its HBM reads
       * overlap the LDS transforms that would otherwise leave the CU idle
       * between planes.  Smaller planes leave room for two workgroups per
       * CU, which overlap each other; they take one orbital each. */
      constexpr bool PF = plane_prefetch<N, V>();
      constexpr int EPT = (N * NH + NT - 1) / NT;
      constexpr int EPV = (N * M + NT - 1) / NT;
      const int nb = PF && a.nb > 0 ? a.nb : 1;
      const int bb0 = blockIdx.x * nb, bb1 = min(bb0 + nb, a.nbatch);
      V vv[EPV];
#pragma unroll
      for (int q = 0; q < EPV; ++q) {
         const int it = threadIdx.x + NT * q;
         vv[q] = Cx<V>::mk(1, 1);
         if (it < N * M) {
            int y, x0;
            slot_xy<N>(it / M, it % M, y, x0);
            vv[q] = cvt<V>(*reinterpret_cast<const double2 *>(a.V + ((size_t)z * N + y) * N + x0));
         }
      }
      auto apply_v = [&]() {
#pragma unroll
         for (int q = 0; q < EPV; ++q) {
            const int it = threadIdx.x + NT * q;
            if (it < N * M) {
               const int s = it / M, m = it % M;
               V v = sm[s * NH + m];
               v.x *= vv[q].x;
               v.y *= vv[q].y;
               sm[s * NH + m] = v;
            }
         }
         __syncthreads();
      };
      if constexpr (PF) {
         V nx[EPT];
         auto fetch = [&](int b) {
            const V *Tq = Tb + ((size_t)b * N + z) * plane;
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
               const int e = threadIdx.x + NT * q;
               nx[q] = (e < N * NH && e % NH < a.nhload) ? Tq[e] : Cx<V>::mk(0, 0);
            }
         };
         if (bb0 < bb1) fetch(bb0);
         for (int b = bb0; b < bb1; ++b) {
            V *Tp = Tb + ((size_t)b * N + z) * plane;
            /* the previous orbital's write-out has read the LDS plane */
            __syncthreads();
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
               const int e = threadIdx.x + NT * q;
               if (e < N * NH) sm[e] = nx[q];
            }
            __syncthreads();
            if (b + 1 < bb1) fetch(b + 1);
            inverse_yx<N>(sm, tws);
            apply_v();
            forward_xy<N>(sm, tws);
            for (int e = threadIdx.x; e < N * NH; e += blockDim.x) Tp[e] = sm[e];
         }
      } else {
         V *Tp = Tb + ((size_t)bb0 * N + z) * plane;
         load_plane<N>(sm, Tp, a.nhload);
         inverse_yx<N>(sm, tws);
         apply_v();
         forward_xy<N>(sm, tws);
         for (int e = threadIdx.x; e < N * NH; e += blockDim.x) Tp[e] = sm[e];
      }
   } else {  /* PM_R2C_DN, PM_R2C_IN */
      const int b = blockIdx.y;
      for (int it = threadIdx.x; it < N * M; it += blockDim.x) {
         const int s = it / M, m = it % M;
         int y, x0;
         slot_xy<N>(s, m, y, x0);
         double2 v;
         if (MODE == PM_R2C_DN) {
            const size_t g = ((size_t)z * N + y) * N + x0;
            const double d0 = a.dn[g], d1 = a.dn[g + 1];
            v = make_double2((d0 + d0) * a.scal1, (d1 + d1) * a.scal1);
         } else {
            const double *row = a.rin + (size_t)b * N * N * (N + 2) + ((size_t)z * N + y) * (N + 2);
            v = make_double2(row[x0], row[x0 + 1]);
         }
         sm[s * NH + m] = cvt<V>(v);
      }
      __syncthreads();
      forward_xy<N>(sm, tws);
      V *Tp = Tb + ((size_t)b * N + z) * plane;
      for (int e = threadIdx.x; e < N * NH; e += blockDim.x) Tp[e] = sm[e];
   }
}

/* ------------------------------------------------------------ launchers */

int fft_supported(int n)
{
   switch (n) {
#define CASE(nn) case nn: return 1;
      CPSD_FOR_EACH_GRID(CASE)
#undef CASE
   default: return 0;
   }
}

template<int N>
static int zg2r(int src, const ZArgs &args, int nbatch, hipStream_t s)
{
   constexpr int KXB = GridCfg<N>::KXB;
   ZArgs a = args;
   const int nkb = (a.rows.nhc + KXB - 1) / KXB;
   /* several orbitals per workgroup while the grid still has >= 8 workgroups
    * per CU to spread */
   a.nbatch = nbatch;
   a.bper = 1;
   while (a.bper < 4 && (long)nkb * N * ((nbatch + 2 * a.bper - 1) / (2 * a.bper)) >= 2048) a.bper *= 2;
   dim3 grid(nkb, N, (nbatch + a.bper - 1) / a.bper);
   constexpr int ZT = GridCfg<N>::ZT;
   if (a.fp32) {
      if (src != SRC_PACKED) {
         cpsd_set_error("zpass g2r: FP32 only for packed orbitals");
         return -1;
      }
      hipLaunchKernelGGL((k_zpass_g2r<N, SRC_PACKED, float2>), grid, dim3(ZT), 0, s, a);
   } else if (src == SRC_PACKED) {
      hipLaunchKernelGGL((k_zpass_g2r<N, SRC_PACKED, double2>), grid, dim3(ZT), 0, s, a);
   } else {
      hipLaunchKernelGGL((k_zpass_g2r<N, SRC_VG, double2>), grid, dim3(ZT), 0, s, a);
   }
   CPSD_CHECK(hipGetLastError());
   return 0;
}

template<int N>
static int zr2g(int dst, const ZArgs &a, int nbatch, hipStream_t s)
{
   constexpr int KXB = GridCfg<N>::KXB;
   dim3 grid((a.rows.nhc + KXB - 1) / KXB, N, nbatch);
   constexpr int ZT = GridCfg<N>::ZT;
   if (a.fp32) {
      if (dst == DST_HPSI) hipLaunchKernelGGL((k_zpass_r2g<N, DST_HPSI, float2>), grid, dim3(ZT), 0, s, a);
      else if (dst == DST_SD) hipLaunchKernelGGL((k_zpass_r2g<N, DST_SD, float2>), grid, dim3(ZT), 0, s, a);
      else {
         cpsd_set_error("zpass r2g: FP32 only for the orbital outputs");
         return -1;
      }
   } else if (dst == DST_HPSI) hipLaunchKernelGGL((k_zpass_r2g<N, DST_HPSI, double2>), grid, dim3(ZT), 0, s, a);
   else if (dst == DST_SD) hipLaunchKernelGGL((k_zpass_r2g<N, DST_SD, double2>), grid, dim3(ZT), 0, s, a);
   else if (dst == DST_DNG) hipLaunchKernelGGL((k_zpass_r2g<N, DST_DNG, double2>), grid, dim3(ZT), 0, s, a);
   else hipLaunchKernelGGL((k_zpass_r2g<N, DST_FULL, double2>), grid, dim3(ZT), 0, s, a);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

template<int N, typename V>
static constexpr size_t plane_lds_bytes()
{
   return ((size_t)N * (N / 2 + 1) + N) * sizeof(V);   /* plane + twiddles */
}

template<int N, int MODE, typename V>
static int plane_launch(const PlaneArgs &a, int nblocks, hipStream_t s)
{
   const size_t lds = plane_lds_bytes<N, V>();
   static std::atomic<unsigned long long> attr_set{0};
   if (dyn_lds_once((const void *)k_plane<N, MODE, V>, (int)lds, attr_set) != 0) return -1;
   const dim3 grid = MODE == PM_VPSI ? dim3(nblocks, N) : dim3(N, nblocks);
   hipLaunchKernelGGL((k_plane<N, MODE, V>), grid, dim3(GridCfg<N>::PLANE_THREADS), lds, s, a);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

template<int N>
static int plane_dispatch(int mode, const PlaneArgs &a, int nblocks, hipStream_t s)
{
   if (a.fp32) {
      switch (mode) {
      case PM_DENSITY: return plane_launch<N, PM_DENSITY, float2>(a, nblocks, s);
      case PM_VPSI:    return plane_launch<N, PM_VPSI, float2>(a, nblocks, s);
      }
      cpsd_set_error("plane: mode %d has no FP32 form", mode);
      return -1;
   }
   switch (mode) {
   case PM_DENSITY: return plane_launch<N, PM_DENSITY, double2>(a, nblocks, s);
   case PM_VTOT:    return plane_launch<N, PM_VTOT, double2>(a, nblocks, s);
   case PM_VPSI:    return plane_launch<N, PM_VPSI, double2>(a, nblocks, s);
   case PM_C2R_OUT: return plane_launch<N, PM_C2R_OUT, double2>(a, nblocks, s);
   case PM_R2C_DN:  return plane_launch<N, PM_R2C_DN, double2>(a, nblocks, s);
   case PM_R2C_IN:  return plane_launch<N, PM_R2C_IN, double2>(a, nblocks, s);
   }
   cpsd_set_error("plane: bad mode %d", mode);
   return -1;
}

int launch_zpass_g2r(int n, int src, const ZArgs &a, int nbatch, hipStream_t s)
{
   switch (n) {
#define CASE(nn) case nn: return zg2r<nn>(src, a, nbatch, s);
      CPSD_FOR_EACH_GRID(CASE)
#undef CASE
   }
   cpsd_set_error("FFT grid %d not supported", n);
   return -1;
}

int launch_zpass_r2g(int n, int dst, const ZArgs &a, int nbatch, hipStream_t s)
{
   switch (n) {
#define CASE(nn) case nn: return zr2g<nn>(dst, a, nbatch, s);
      CPSD_FOR_EACH_GRID(CASE)
#undef CASE
   }
   cpsd_set_error("FFT grid %d not supported", n);
   return -1;
}

template<int N, typename V>
static int density_blocks_per_cu()
{
   const size_t lds = plane_lds_bytes<N, V>();
   const void *k = (const void *)k_plane<N, PM_DENSITY, V>;
   int nb = 0;
   if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess ||
       hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, GridCfg<N>::PLANE_THREADS, lds) != hipSuccess)
      return 1;
   return nb > 0 ? nb : 1;
}

int plane_prefetch_on(int n)
{
   return (size_t)n * (n / 2 + 1) * 16 > (size_t)PLANE_PF_KB * 1024;
}

int plane_density_occupancy(int n, int fp32)
{
   switch (n) {
#define CASE(nn) case nn: return fp32 ? density_blocks_per_cu<nn, float2>() : density_blocks_per_cu<nn, double2>();
      CPSD_FOR_EACH_GRID(CASE)
#undef CASE
   }
   return 1;
}

int launch_plane(int n, int mode, const PlaneArgs &a, int nblocks, hipStream_t s)
{
   switch (n) {
#define CASE(nn) case nn: return plane_dispatch<nn>(mode, a, nblocks, s);
      CPSD_FOR_EACH_GRID(CASE)
#undef CASE
   }
   cpsd_set_error("FFT grid %d not supported", n);
   return -1;
}
