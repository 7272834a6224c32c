/*
 * fft_core.h -- in-register and LDS-resident DFT building blocks, FP64
 * (double2) or FP32 (float2).
 *
 * Conventions (those of fftpack3d/fftpack3d.f and FFTPACK):
 *   DIR = -1 : X_k = sum_n x_n exp(-2 pi i nk/N)   (forward, rc_fft3f)
 *   DIR = +1 : X_k = sum_n x_n exp(+2 pi i nk/N)   (backward, cr_fft3b)
 * Transforms are unnormalised.
 *
 * A length-N transform with N = N1*N2 runs in place in LDS as two register
 * passes separated by one barrier.  Slot s = N2*a + b holds logical index
 * a + N1*b in the "transposed" order:
 *   lds_fft_nat2tr: natural input  -> transposed output
 *   lds_fft_tr2nat: transposed input -> natural output
 * Every item reads and writes only its own slots inside a pass, so no extra
 * scratch is needed; callers un-permute indices when they move data to or
 * from global memory (tr_index below).
 */
#pragma once
#include <hip/hip_runtime.h>
#include "small_twiddles.h"

/* complex helpers for the FP64 (double2) and FP32 (float2) transforms */
template<typename V> struct Cx;
template<> struct Cx<double2> {
   using R = double;
   static __device__ __forceinline__ double2 mk(double x, double y) { return make_double2(x, y); }
};
template<> struct Cx<float2> {
   using R = float;
   static __device__ __forceinline__ float2 mk(float x, float y) { return make_float2(x, y); }
};

/* storage conversions: FP64 tables and orbitals in, FP64 results out */
template<typename V> __device__ __forceinline__ V cvt(double2 a);
template<> __device__ __forceinline__ double2 cvt<double2>(double2 a) { return a; }
template<> __device__ __forceinline__ float2 cvt<float2>(double2 a)
{
   return make_float2((float)a.x, (float)a.y);
}
__device__ __forceinline__ double2 to_d2(double2 a) { return a; }
__device__ __forceinline__ double2 to_d2(float2 a) { return make_double2(a.x, a.y); }

template<typename V> __device__ __forceinline__ V cmul(V a, V b)
{
   return Cx<V>::mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template<typename V> __device__ __forceinline__ V cadd(V a, V b) { return Cx<V>::mk(a.x + b.x, a.y + b.y); }
template<typename V> __device__ __forceinline__ V csub(V a, V b) { return Cx<V>::mk(a.x - b.x, a.y - b.y); }
template<typename V> __device__ __forceinline__ V cconj(V a) { return Cx<V>::mk(a.x, -a.y); }
/* i * a */
template<typename V> __device__ __forceinline__ V cmuli(V a) { return Cx<V>::mk(-a.y, a.x); }
template<typename V> __device__ __forceinline__ V cscale(V a, typename Cx<V>::R s) { return Cx<V>::mk(a.x * s, a.y * s); }

/* ---------------------------------------------------------------- DFTs */

template<int L, int DIR> struct DFT;

template<int DIR> struct DFT<1, DIR> {
   template<typename V> static __device__ __forceinline__ void run(V *) {}
};

template<int DIR> struct DFT<2, DIR> {
   template<typename V> static __device__ __forceinline__ void run(V *v)
   {
      V a = v[0], b = v[1];
      v[0] = cadd(a, b);
      v[1] = csub(a, b);
   }
};

template<int DIR> struct DFT<3, DIR> {
   template<typename V> static __device__ __forceinline__ void run(V *v)
   {
      using R = typename Cx<V>::R;
      const R s3 = R(0.8660254037844386) * DIR;
      V a = v[0], t1 = cadd(v[1], v[2]);
      V t2 = Cx<V>::mk(a.x - R(0.5) * t1.x, a.y - R(0.5) * t1.y);
      V t3 = cscale(csub(v[1], v[2]), s3);
      v[0] = cadd(a, t1);
      v[1] = cadd(t2, cmuli(t3));
      v[2] = csub(t2, cmuli(t3));
   }
};

template<int DIR> struct DFT<4, DIR> {
   template<typename V> static __device__ __forceinline__ void run(V *v)
   {
      V ac = cadd(v[0], v[2]), amc = csub(v[0], v[2]);
      V bd = cadd(v[1], v[3]), bmd = csub(v[1], v[3]);
      V j = cmuli(bmd);
      if (DIR < 0) j = Cx<V>::mk(-j.x, -j.y);
      v[0] = cadd(ac, bd);
      v[2] = csub(ac, bd);
      v[1] = cadd(amc, j);
      v[3] = csub(amc, j);
   }
};

template<int DIR> struct DFT<5, DIR> {
   template<typename V> static __device__ __forceinline__ void run(V *v)
   {
      using R = typename Cx<V>::R;
      const R c1 = R(0.30901699437494745), c2 = R(-0.8090169943749475);
      const R s1 = R(0.9510565162951535) * DIR, s2 = R(0.5877852522924731) * DIR;
      V a = v[0];
      V be = cadd(v[1], v[4]), bme = csub(v[1], v[4]);
      V cd = cadd(v[2], v[3]), cmd = csub(v[2], v[3]);
      V r1 = Cx<V>::mk(a.x + c1 * be.x + c2 * cd.x, a.y + c1 * be.y + c2 * cd.y);
      V r2 = Cx<V>::mk(a.x + c2 * be.x + c1 * cd.x, a.y + c2 * be.y + c1 * cd.y);
      V i1 = cmuli(Cx<V>::mk(s1 * bme.x + s2 * cmd.x, s1 * bme.y + s2 * cmd.y));
      V i2 = cmuli(Cx<V>::mk(s2 * bme.x - s1 * cmd.x, s2 * bme.y - s1 * cmd.y));
      v[0] = cadd(a, cadd(be, cd));
      v[1] = cadd(r1, i1);
      v[4] = csub(r1, i1);
      v[2] = cadd(r2, i2);
      v[3] = csub(r2, i2);
   }
};

/* composite L = A*B in registers: n = B*a + b, k = ka + A*kb */
template<int A, int B, int DIR, typename V>
__device__ __forceinline__ void dft_composite(V *v)
{
   using R = typename Cx<V>::R;
   constexpr int L = A * B;
   V u[L];
#pragma unroll
   for (int b = 0; b < B; ++b) {
      V t[A];
#pragma unroll
      for (int a = 0; a < A; ++a) t[a] = v[B * a + b];
      DFT<A, DIR>::run(t);
#pragma unroll
      for (int ka = 0; ka < A; ++ka) {
         if (b == 0 || ka == 0) {
            u[b * A + ka] = t[ka];
         } else {
            const int m = (b * ka) % L;
            V w = Cx<V>::mk(R(SmallTw<L>::re(m)), R(DIR * SmallTw<L>::im(m)));
            u[b * A + ka] = cmul(t[ka], w);
         }
      }
   }
#pragma unroll
   for (int ka = 0; ka < A; ++ka) {
      V t[B];
#pragma unroll
      for (int b = 0; b < B; ++b) t[b] = u[b * A + ka];
      DFT<B, DIR>::run(t);
#pragma unroll
      for (int kb = 0; kb < B; ++kb) v[ka + A * kb] = t[kb];
   }
}

template<int DIR> struct DFT<6, DIR>  { template<typename V> static __device__ __forceinline__ void run(V *v) { dft_composite<3, 2, DIR>(v); } };
template<int DIR> struct DFT<8, DIR>  { template<typename V> static __device__ __forceinline__ void run(V *v) { dft_composite<4, 2, DIR>(v); } };
template<int DIR> struct DFT<10, DIR> { template<typename V> static __device__ __forceinline__ void run(V *v) { dft_composite<5, 2, DIR>(v); } };
template<int DIR> struct DFT<12, DIR> { template<typename V> static __device__ __forceinline__ void run(V *v) { dft_composite<4, 3, DIR>(v); } };
template<int DIR> struct DFT<16, DIR> { template<typename V> static __device__ __forceinline__ void run(V *v) { dft_composite<4, 4, DIR>(v); } };

/* twiddle exp(DIR 2 pi i m / N) from a table of exp(+2 pi i m / N), m < N */
template<int DIR, typename V>
__device__ __forceinline__ V twiddle(const V *__restrict__ tw, int m)
{
   V w = tw[m];
   if (DIR < 0) w.y = -w.y;
   return w;
}

/* logical index held by slot s in the transposed order of N = N1*N2 */
template<int N1, int N2>
__device__ __forceinline__ int tr_index(int s) { return s / N2 + N1 * (s % N2); }

/* slot holding logical index j in the transposed order */
template<int N1, int N2>
__device__ __forceinline__ int tr_slot(int j) { return N2 * (j % N1) + j / N1; }

/* count transforms; element j of transform f sits at buf[f*fs + j*es].
 * tw: table of exp(+2 pi i m / (N*tws)), twiddle W_N^m read at m*tws.  */
template<int N1, int N2, int DIR, typename V>
__device__ __forceinline__ void lds_fft_nat2tr(V *buf, int count, int fs, int es,
                                               const V *__restrict__ tw, int tws)
{
   constexpr int N = N1 * N2;
   for (int it = threadIdx.x; it < count * N2; it += blockDim.x) {
      const int f = it % count, n2 = it / count;
      V *p = buf + f * fs;
      V v[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) v[n1] = p[(N2 * n1 + n2) * es];
      DFT<N1, DIR>::run(v);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twiddle<DIR>(tw, ((n2 * k1) % N) * tws));
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) p[(N2 * k1 + n2) * es] = v[k1];
   }
   __syncthreads();
   for (int it = threadIdx.x; it < count * N1; it += blockDim.x) {
      const int f = it % count, k1 = it / count;
      V *p = buf + f * fs;
      V v[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) v[n2] = p[(N2 * k1 + n2) * es];
      DFT<N2, DIR>::run(v);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) p[(N2 * k1 + k2) * es] = v[k2];
   }
   __syncthreads();
}

template<int N1, int N2, int DIR, typename V>
__device__ __forceinline__ void lds_fft_tr2nat(V *buf, int count, int fs, int es,
                                               const V *__restrict__ tw, int tws)
{
   constexpr int N = N1 * N2;
   for (int it = threadIdx.x; it < count * N1; it += blockDim.x) {
      const int f = it % count, n1 = it / count;
      V *p = buf + f * fs;
      V v[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) v[n2] = p[(N2 * n1 + n2) * es];
      DFT<N2, DIR>::run(v);
#pragma unroll
      for (int k2 = 1; k2 < N2; ++k2) v[k2] = cmul(v[k2], twiddle<DIR>(tw, ((n1 * k2) % N) * tws));
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) p[(N2 * n1 + k2) * es] = v[k2];
   }
   __syncthreads();
   for (int it = threadIdx.x; it < count * N2; it += blockDim.x) {
      const int f = it % count, k2 = it / count;
      V *p = buf + f * fs;
      V v[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) v[n1] = p[(N2 * n1 + k2) * es];
      DFT<N1, DIR>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) p[(N2 * k1 + k2) * es] = v[k1];
   }
   __syncthreads();
}

/* ------------------------------------------------ grid configurations */

/* N = NY1*NY2 for the y and z passes, N/2 = NX1*NX2 for the real x pass */
template<int N> struct GridCfg;
/* z pass: KXB kx columns per workgroup of ZT threads; whole kx rows up to
 * N = 96 (no row is split between workgroups), two balanced halves at 128 */
#define GRID_CFG(n, y1, y2, x1, x2, pt, kxb, zt)                               \
   template<> struct GridCfg<n> {                                              \
      static constexpr int NY1 = y1, NY2 = y2, NX1 = x1, NX2 = x2;             \
      static constexpr int PLANE_THREADS = pt, KXB = kxb, ZT = zt;             \
   };
GRID_CFG(8,   4,  2, 2, 2, 256, 5,  256)
GRID_CFG(12,  4,  3, 3, 2, 256, 7,  256)
GRID_CFG(16,  4,  4, 4, 2, 256, 9,  256)
GRID_CFG(20,  5,  4, 5, 2, 256, 11, 256)
GRID_CFG(24,  6,  4, 4, 3, 256, 13, 256)
GRID_CFG(32,  8,  4, 4, 4, 256, 17, 256)
GRID_CFG(40,  8,  5, 5, 4, 256, 21, 256)
GRID_CFG(48,  8,  6, 6, 4, 256, 25, 256)
GRID_CFG(64,  8,  8, 8, 4, 256, 33, 256)
GRID_CFG(80,  10, 8, 8, 5, 512, 41, 512)
GRID_CFG(96,  8, 12, 8, 6, 512, 49, 512)
GRID_CFG(128, 8, 16, 8, 8, 512, 33, 512)
#undef GRID_CFG

#define CPSD_FOR_EACH_GRID(X) X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128)
