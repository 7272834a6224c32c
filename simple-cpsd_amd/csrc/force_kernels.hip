/*
 * force_kernels.hip -- ionic forces and the ion update of a moving-ion SD
 * step (FP64), ccpsdv3/cpsd.c:
 *   local force      :246-257   F_I += gsdot(G, Im-pair(dng, S_I vl))
 *   nonlocal force   :291-304   F_I -= 4 sw gsdot(G, Im-pair(c1_i, beta_Il))
 *   Ewald, G space   :322-332   F_I += gsdot(G, Im-pair(S_I, s) vg) zv scal2
 *   Ewald, real space:334-362   pair sums over the nsh image cells
 *   ion update       :401-411   r2 = r1 + dt/sqrt(M_I) F_I
 * The G-space sums run over the packed masked G with the gsdot weights
 * (1 on the kx = 0 plane, 2 elsewhere).  The nonlocal force comes from the
 * tall MFMA kernel (gemm_kernels.hip, EPI_NLFORCE), which forms
 * Z_p = sum_i sw_pi c1_i and reduces w G Im-pair(Z_p, beta_p) per projector;
 * this file adds the other terms and sums everything in a fixed order, so a
 * run is deterministic.
 */
#include "cpsd_internal.h"
#include "sd_kernels.h"

template<int NT>
__device__ __forceinline__ double fsum(double v, double *red)
{
   red[threadIdx.x] = v;
   __syncthreads();
#pragma unroll
   for (int s = NT / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
   }
   const double r = red[0];
   __syncthreads();
   return r;
}

/* block (chunk, ion): partial G-space local + Ewald force of that ion */
__global__ void __launch_bounds__(256) k_force_gspace(ForceArgs a)
{
   __shared__ double red[256];
   const int ch = blockIdx.x, ii = blockIdx.y;
   const int ia = a.katm[ii];
   const double zsc = a.zv[ia] * a.scal2;
   const int per = (a.ng + a.nchunk - 1) / a.nchunk;
   const int p0 = ch * per, p1 = min(a.ng, p0 + per);
   const double2 *e1t = a.ex1 + (size_t)ii * (a.n / 2 + 1);
   const double2 *e2t = a.ex2 + (size_t)ii * a.n;
   const double2 *e3t = a.ex3 + (size_t)ii * a.n;
   const double *vl = a.vl + (size_t)ia * a.ngp;
   double f0 = 0.0, f1 = 0.0, f2 = 0.0;
   for (int p = p0 + (int)threadIdx.x; p < p1; p += 256) {
      /* structure factor of ion ii, as in k_structure */
      const double2 e1 = e1t[a.kx[p]], e2 = e2t[a.ky[p]], e3 = e3t[a.kz[p]];
      const double2 t = make_double2(e2.x * e3.x - e2.y * e3.y, e2.x * e3.y + e2.y * e3.x);
      const double2 e = make_double2(e1.x * t.x - e1.y * t.y, e1.y * t.x + e1.x * t.y);
      const double2 d = a.dng[p], sg = a.s[p];
      const double v = vl[p];
      const double vr = e.x * v, vi = e.y * v;            /* vtmp = S_I vl */
      const double xloc = d.y * vr - d.x * vi;             /* :252 */
      const double xew = (e.x * sg.y - e.y * sg.x) * a.vg[p];   /* :327 */
      const double tv = a.w[2 * p] * (xloc + zsc * xew);
      f0 += a.Gp[p] * tv;
      f1 += a.Gp[(size_t)a.ngp + p] * tv;
      f2 += a.Gp[2 * (size_t)a.ngp + p] * tv;
   }
   f0 = fsum<256>(f0, red);
   f1 = fsum<256>(f1, red);
   f2 = fsum<256>(f2, red);
   if (threadIdx.x == 0) {
      double *o = a.part + ((size_t)ii * a.nchunk + ch) * 3;
      o[0] = f0; o[1] = f1; o[2] = f2;
   }
}

/* block (projector p, direction d): sum of the nonlocal partials over the
 * K blocks of the tall kernel */
__global__ void __launch_bounds__(256) k_force_nlsum(const double *__restrict__ nlpart, int kblocks,
                                                     int nproj, double *__restrict__ nlsum)
{
   __shared__ double red[256];
   const int pd = blockIdx.x;   /* p*3 + d */
   double v = 0.0;
   for (int kb = threadIdx.x; kb < kblocks; kb += 256) v += nlpart[(size_t)kb * nproj * 3 + pd];
   v = fsum<256>(v, red);
   if (threadIdx.x == 0) nlsum[pd] = v;
}

/* one thread per ion: add the parts, the real-space Ewald pairs, then move
 * the ion */
__global__ void k_force_finish(ForceFinishArgs a)
{
   const int ii = blockIdx.x * blockDim.x + threadIdx.x;
   if (ii >= a.nion) return;
   double f[3] = {0.0, 0.0, 0.0};
   for (int ch = 0; ch < a.nchunk; ++ch)
      for (int d = 0; d < 3; ++d) f[d] += a.gpart[((size_t)ii * a.nchunk + ch) * 3 + d];
   const int ia = a.katm[ii];
   for (int l = 0; l < a.lmmax[ia]; ++l)
      for (int d = 0; d < 3; ++d) f[d] -= 4.0 * a.nlsum[(a.poff[ii] + l) * 3 + d];
   /* real space, :334-362, with the erfc of :11-17 */
   const double CERFC = 1.128379167;
   const double B1 = 0.0705230784, B2 = 0.0422820123, B3 = 0.0092705272;
   const double B4 = 0.0001520143, B5 = 0.0002765672, B6 = 0.0000430638;
   const double zi = a.zv[ia];
   for (int jj = 0; jj < a.nion; ++jj) {
      if (jj == ii) continue;
      const double dx = a.r1[3 * ii] - a.r1[3 * jj];
      const double dy = a.r1[3 * ii + 1] - a.r1[3 * jj + 1];
      const double dz = a.r1[3 * ii + 2] - a.r1[3 * jj + 2];
      const double zz = zi * a.zv[a.katm[jj]];
      double s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int l = 0; l < a.nsh; ++l) {
         const double x = a.rcell[l] + dx, y = a.rcell[l + a.nsh] + dy, z = a.rcell[l + 2 * a.nsh] + dz;
         const double r = sqrt(x * x + y * y + z * z);
         const double w = r / a.rcut;
         double aa = 1.0 + w * (B1 + w * (B2 + w * (B3 + w * (B4 + w * (B5 + w * B6)))));
         aa *= aa * aa * aa;
         const double erfc = 1.0 / (aa * aa * aa * aa);
         const double adiff = zz * (erfc + CERFC * w * exp(-w * w)) / (r * r * r);
         s1 += x * adiff;
         s2 += y * adiff;
         s3 += z * adiff;
      }
      f[0] += s1; f[1] += s2; f[2] += s3;
   }
   for (int d = 0; d < 3; ++d) {
      a.fion[3 * ii + d] = f[d];
      a.r2[3 * ii + d] = a.r1[3 * ii + d] + a.dti[ii] * f[d];
   }
}

int sd_force_gspace(const ForceArgs &a, hipStream_t s)
{
   hipLaunchKernelGGL(k_force_gspace, dim3(a.nchunk, a.nion), dim3(256), 0, s, a);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_force_nlsum(const double *nlpart, int kblocks, int nproj, double *nlsum, hipStream_t s)
{
   if (nproj <= 0) return 0;
   hipLaunchKernelGGL(k_force_nlsum, dim3(3 * nproj), dim3(256), 0, s, nlpart, kblocks, nproj, nlsum);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_force_finish(const ForceFinishArgs &a, hipStream_t s)
{
   hipLaunchKernelGGL(k_force_finish, dim3((a.nion + 63) / 64), dim3(64), 0, s, a);
   CPSD_CHECK(hipGetLastError());
   return 0;
}
