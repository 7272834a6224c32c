/*
 * kbpp_kernels.hip -- the Kleinman-Bylander transform of a radial
 * pseudopotential table onto the G grid (the .vpp generator,
 * ckbppv3/dkbppv3.c:335-444) on the device.
 *
 * For every G != 0 (|G| = q, unit vector g) the reference forms, with
 * composite Simpson sums over the radial grid rho_j (ckbppv3/dkbppv3.c:310-328):
 *   s:     vnl_0 = sqrt(4pi)   S[sin(q r)          wp_0 dv_0] / q
 *   p:     vnl_p = sqrt(12pi)  S[j1-like(q r)      wp_1 dv_1] / q  x g
 *   d:     vnl_d = sqrt(60pi)  S[j2-like(q r)      wp_2 dv_2] / q  x angular
 *   local: vl    = 4pi S[r sin(q r) dv_lmax] / q - 4pi zv cos(q r_last) / q^2
 * with dv_l = V_l - V_lmax for l < lmax.  Everything that does not depend on
 * G (the Simpson weight times wp_l dv_l, or r dv_lmax) is folded into one
 * table on the host, so one thread per G runs a single pass over rho with
 * one sincos per point; the table is staged through LDS in chunks that every
 * thread of the workgroup reads in the same order.  The p and d sums skip
 * j = 0 as the reference does (its f[0] = 0).  G = 0, the normalisation
 * integrals vnlnrm and the table are O(nrho) host work.
 */
#include "cpsd_internal.h"
#include <cmath>
#include <vector>

#define KB_THREADS 256
#define KB_CHUNK 512

struct KbppArgs {
   const double *G;    /* [3][nfft3d] */
   const double *tab;  /* [5][nrho]: rho, then the weighted s, p, d and local products */
   int nfft3d, nrho, lmax;
   double zv, rlast;
   double *vl, *vnl;   /* vl[nfft3d], vnl[lmax^2][nfft3d] */
};

__global__ void __launch_bounds__(KB_THREADS) k_kbpp(KbppArgs a)
{
   __shared__ double t[5][KB_CHUNK];
   const int k = blockIdx.x * KB_THREADS + threadIdx.x;
   const bool act = k >= 1 && k < a.nfft3d;
   double gx = 0.0, gy = 0.0, gz = 0.0, q = 1.0;
   if (act) {
      gx = a.G[k];
      gy = a.G[k + (size_t)a.nfft3d];
      gz = a.G[k + 2 * (size_t)a.nfft3d];
      q = sqrt(gx * gx + gy * gy + gz * gz);
   }
   double s0 = 0.0, s1 = 0.0, s2 = 0.0, sl = 0.0;
   for (int j0 = 0; j0 < a.nrho; j0 += KB_CHUNK) {
      __syncthreads();
      for (int e = threadIdx.x; e < 5 * KB_CHUNK; e += KB_THREADS) {
         const int c = e / KB_CHUNK, jj = e % KB_CHUNK;
         t[c][jj] = j0 + jj < a.nrho ? a.tab[(size_t)c * a.nrho + j0 + jj] : 0.0;
      }
      __syncthreads();
      if (act) {
         const int jn = min(KB_CHUNK, a.nrho - j0);
         for (int jj = 0; jj < jn; ++jj) {
            const double x = q * t[0][jj];
            double sn, cs;
            sincos(x, &sn, &cs);
            s0 += sn * t[1][jj];
            sl += sn * t[4][jj];
            if (j0 + jj > 0) {
               const double b1 = sn / x - cs;
               s1 += b1 * t[2][jj];
               s2 += (3.0 * b1 / x - sn) * t[3][jj];
            }
         }
      }
   }
   if (!act) return;
   const double fourpi = 4.0 * M_PI;
   const double p0 = sqrt(fourpi), p1 = sqrt(3.0 * fourpi), p2 = sqrt(15.0 * fourpi);
   const size_t n = a.nfft3d;
   gx /= q; gy /= q; gz /= q;
   if (a.lmax > 2) {
      const double d = p2 * s2 / q;
      a.vnl[k + 4 * n] = d * (3.0 * gz * gz - 1.0) / (2.0 * sqrt(3.0));
      a.vnl[k + 5 * n] = d * gx * gy;
      a.vnl[k + 6 * n] = d * gy * gz;
      a.vnl[k + 7 * n] = d * gz * gx;
      a.vnl[k + 8 * n] = d * (gx * gx - gy * gy) / 2.0;
   }
   if (a.lmax > 1) {
      const double p = p1 * s1 / q;
      a.vnl[k + 1 * n] = p * gx;
      a.vnl[k + 2 * n] = p * gy;
      a.vnl[k + 3 * n] = p * gz;
   }
   if (a.lmax > 0) a.vnl[k] = p0 * s0 / q;
   a.vl[k] = sl * fourpi / q - a.zv * fourpi * cos(q * a.rlast) / (q * q);
}

/* Simpson sum of ckbppv3/dkbppv3.c:310-328 */
static double simp(int n, const double *y, double h)
{
   double s2 = 0.0, s4 = 0.0;
   for (int i = 0; i < n; i += 2) s2 += y[i];
   for (int i = 1; i < n; i += 2) s4 += y[i];
   return (2.0 * s2 + 4.0 * s4 - y[0] - y[n - 1]) * (h / 3.0);
}

/* same arguments and outputs as the host transform cpsdh_kbpp (vp is not
 * modified; vnl channels at and above lmax^2 are left as they are) */
extern "C" int cpsd_kbpp_device(double zv, int lmax, int nrho, double drho, const double *rho,
                                const double *vp, const double *wp, int nfft3d, const double *G,
                                double *vl, double *vnl, double *vnlnrm)
{
   if (lmax < 0 || lmax > 3 || nrho < 1 || nfft3d < 1) {
      cpsd_set_error("kbpp: lmax=%d nrho=%d nfft3d=%d out of range", lmax, nrho, nfft3d);
      return -1;
   }
   const int lmmax = lmax * lmax;
   const double fourpi = 4.0 * M_PI;
   std::vector<double> dv((size_t)(lmax + 1) * nrho), f(nrho);
   for (int l = 0; l <= lmax; ++l)
      for (int i = 0; i < nrho; ++i)
         dv[i + (size_t)l * nrho] = l < lmax ? vp[i + (size_t)l * nrho] - vp[i + (size_t)lmax * nrho]
                                             : vp[i + (size_t)l * nrho];
   for (int l = 0; l < lmax; ++l) {
      for (int i = 0; i < nrho; ++i) f[i] = dv[i + (size_t)l * nrho] * wp[i + (size_t)l * nrho] * wp[i + (size_t)l * nrho];
      const double a = simp(nrho, f.data(), drho);
      for (int i = l * l; i < (l + 1) * (l + 1); ++i) vnlnrm[i] = a;
   }
   for (int i = lmmax; i < 9; ++i) vnlnrm[i] = 0.0;
   /* G-independent factors with the Simpson weights folded in */
   std::vector<double> tab((size_t)5 * nrho, 0.0);
   for (int j = 0; j < nrho; ++j) {
      const double ws = ((j % 2 == 0) ? 2.0 : 4.0) - (j == 0) - (j == nrho - 1);
      const double w = ws * (drho / 3.0);
      tab[j] = rho[j];
      if (lmax > 0) tab[nrho + j] = w * wp[j] * dv[j];
      if (lmax > 1) tab[2 * (size_t)nrho + j] = w * wp[j + nrho] * dv[j + nrho];
      if (lmax > 2) tab[3 * (size_t)nrho + j] = w * wp[j + 2 * (size_t)nrho] * dv[j + 2 * (size_t)nrho];
      tab[4 * (size_t)nrho + j] = w * rho[j] * dv[j + (size_t)lmax * nrho];
   }
   double *dG = nullptr, *dtab = nullptr, *dvl = nullptr, *dvnl = nullptr;
   const size_t n = nfft3d;
   hipStream_t st = nullptr;
   auto run = [&]() -> int {
      CPSD_CHECK(hipStreamCreate(&st));
      CPSD_CHECK(hipMalloc(&dG, 3 * n * sizeof(double)));
      CPSD_CHECK(hipMalloc(&dtab, tab.size() * sizeof(double)));
      CPSD_CHECK(hipMalloc(&dvl, n * sizeof(double)));
      CPSD_CHECK(hipMalloc(&dvnl, (lmmax > 0 ? lmmax : 1) * n * sizeof(double)));
      CPSD_CHECK(hipMemcpyAsync(dG, G, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
      CPSD_CHECK(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st));
      KbppArgs a;
      a.G = dG; a.tab = dtab; a.nfft3d = nfft3d; a.nrho = nrho; a.lmax = lmax;
      a.zv = zv; a.rlast = rho[nrho - 1]; a.vl = dvl; a.vnl = dvnl;
      hipLaunchKernelGGL(k_kbpp, dim3((nfft3d + KB_THREADS - 1) / KB_THREADS), dim3(KB_THREADS), 0, st, a);
      CPSD_CHECK(hipGetLastError());
      if (n > 1) {
         CPSD_CHECK(hipMemcpyAsync(vl + 1, dvl + 1, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, st));
         for (int l = 0; l < lmmax; ++l)
            CPSD_CHECK(hipMemcpyAsync(vnl + l * n + 1, dvnl + l * n + 1, (n - 1) * sizeof(double),
                                      hipMemcpyDeviceToHost, st));
      }
      CPSD_CHECK(hipStreamSynchronize(st));
      return 0;
   };
   const int rc = run();
   (void)hipFree(dG); (void)hipFree(dtab); (void)hipFree(dvl); (void)hipFree(dvnl);
   if (st) (void)hipStreamDestroy(st);
   if (rc) return rc;
   /* G = 0 (ckbppv3/dkbppv3.c:428-438) */
   for (int i = 0; i < nrho; ++i) f[i] = dv[i + (size_t)lmax * nrho] * (rho[i] * rho[i]);
   vl[0] = fourpi * simp(nrho, f.data(), drho) + 2.0 * M_PI * zv * rho[nrho - 1] * rho[nrho - 1];
   for (int i = 0; i < nrho; ++i) f[i] = rho[i] * wp[i] * dv[i];
   vnl[0] = sqrt(fourpi) * simp(nrho, f.data(), drho);
   for (int l = 1; l < lmmax; ++l) vnl[l * n] = 0.0;
   return 0;
}
