/*
 * sd_kernels.hip -- elementwise, reduction and small dense kernels of the SD
 * step (FP64): packing of the half-complex grid, structure factors and
 * projectors, density reduction + LDA, overlap matrices, the Lagrange
 * multiplier fixed point, and the energy terms.
 */
#include "cpsd_internal.h"
#include "sd_kernels.h"
#include "mfma64.h"

/* ------------------------------------------------- packing (boundary) */

__global__ void k_pack(const double2 *__restrict__ full, size_t fstride, const int *__restrict__ gidx,
                       int ng, int ngp, double2 *__restrict__ packed, size_t pstride, int ne)
{
   const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   const int i = blockIdx.y;
   if (p >= (size_t)ngp || i >= ne) return;
   packed[i * pstride + p] = (p < (size_t)ng) ? full[i * fstride + gidx[p]] : make_double2(0.0, 0.0);
}

__global__ void k_check_unmasked(const double2 *__restrict__ full, size_t fstride,
                                 const int *__restrict__ pidx, int nfft3d, int ne,
                                 int *__restrict__ flag)
{
   const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   const int i = blockIdx.y;
   if (g >= (size_t)nfft3d || i >= ne) return;
   if (pidx[g] < 0) {
      const double2 v = full[i * fstride + g];
      if (v.x != 0.0 || v.y != 0.0) atomicOr(flag, 1);
   }
}

__global__ void k_unpack(const double2 *__restrict__ packed, size_t pstride, const int *__restrict__ pidx,
                         int nfft3d, double2 *__restrict__ full, size_t fstride, int ne)
{
   const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   const int i = blockIdx.y;
   if (g >= (size_t)nfft3d || i >= ne) return;
   const int p = pidx[g];
   full[i * fstride + g] = (p >= 0) ? packed[i * pstride + p] : make_double2(0.0, 0.0);
}

/* dense N^3 real field -> padded (N+2) rows with zeroed pads */
__global__ void k_pad_real(const double *__restrict__ dense, int n, double *__restrict__ padded)
{
   const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   const size_t rows = (size_t)n * n;
   if (idx >= rows * (n + 2)) return;
   const size_t r = idx / (n + 2);
   const int x = (int)(idx % (n + 2));
   padded[idx] = (x < n) ? dense[r * n + x] : 0.0;
}

/* ------------------------------------- structure factors and projectors */

/* Per packed G: exi_I = ex1[kx] * (ex2[ky] * ex3[kz]) (ccpsdv3/strfac.c:93-123),
 * s = sum_I zv exi_I, vall = sum_I vl exi_I (ccpsdv3/cpsd.c:232-245), and the
 * projector rows beta_{I,l} = exi_I vnl_l (times i for l = 1..3,
 * ccpsdv3/cpsd.c:260-286), ions in input order. */
__global__ void k_structure(StructArgs a)
{
   const int p = blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= a.ngp) return;
   double2 s = make_double2(0.0, 0.0), vall = make_double2(0.0, 0.0);
   if (p >= a.ng) {
      a.s[p] = s; a.vall[p] = vall;
      for (int m = 0; m < a.nproj; ++m) a.B[(size_t)m * a.ngp + p] = s;
      return;
   }
   const int kx = a.kx[p], ky = a.ky[p], kz = a.kz[p];
   int m = 0;
   for (int ii = 0; ii < a.nion; ++ii) {
      const int ia = a.katm[ii];
      const double2 e1 = a.ex1[ii * (a.n + 2) / 2 + kx];
      const double2 e2 = a.ex2[ii * a.n + ky];
      const double2 e3 = a.ex3[ii * a.n + kz];
      const double2 t = make_double2(e2.x * e3.x - e2.y * e3.y, e2.x * e3.y + e2.y * e3.x);
      const double2 e = make_double2(e1.x * t.x - e1.y * t.y, e1.y * t.x + e1.x * t.y);
      const double z = a.zv[ia];
      s.x += z * e.x; s.y += z * e.y;
      const double vl = a.vl[(size_t)ia * a.ngp + p];
      vall.x += e.x * vl; vall.y += e.y * vl;
      for (int l = 0; l < a.lmmax[ia]; ++l) {
         const double v = a.vnl[((size_t)ia * 9 + l) * a.ngp + p];
         double2 b;
         if (l == 0 || l >= 4) b = make_double2(e.x * v, e.y * v);
         else b = make_double2(-e.y * v, e.x * v);
         a.B[(size_t)m * a.ngp + p] = b;
         ++m;
      }
   }
   a.s[p] = s;
   a.vall[p] = vall;
}

/* Phase factors by repeated complex multiplication (ccpsdv3/strfac.c:10-77):
 * ex1[k] = e^{-i(b1.R+pi)k}, k = 0..N/2; ex2/ex3 over k = 0..N-1 with
 * ex[N-k] = conj(ex[k]); the Nyquist entries are zero.  One thread per
 * (ion, direction); pi = 4 atan(1) as the reference passes it. */
__global__ void k_phafac(int n, const double *__restrict__ unitg, int nion, const double *__restrict__ r1,
                         double2 *__restrict__ ex1, double2 *__restrict__ ex2, double2 *__restrict__ ex3)
{
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= 3 * nion) return;
   const int ii = t / 3, d = t % 3;
   const double pi = 4.0 * atan(1.0);
   const double sw = unitg[3 * d] * r1[3 * ii] + unitg[3 * d + 1] * r1[3 * ii + 1]
                   + unitg[3 * d + 2] * r1[3 * ii + 2] + pi;
   const double cr = cos(sw), ci = -sin(sw);
   const int nh = n / 2;
   double2 *x = (d == 0) ? ex1 + (size_t)ii * (nh + 1) : ((d == 1) ? ex2 : ex3) + (size_t)ii * n;
   double pr = 1.0, pim = 0.0;
   x[0] = make_double2(1.0, 0.0);
   for (int k = 1; k <= nh; ++k) {
      const double nr = pr * cr - pim * ci;
      const double ni = pr * ci + pim * cr;
      pr = nr; pim = ni;
      x[k] = make_double2(nr, ni);
      if (d > 0) x[n - k] = make_double2(nr, -ni);
   }
   x[nh] = make_double2(0.0, 0.0);
}

int sd_phafac(int n, const double *unitg, int nion, const double *r1, double2 *ex1, double2 *ex2,
              double2 *ex3, hipStream_t s)
{
   hipLaunchKernelGGL(k_phafac, dim3((3 * nion + 63) / 64), dim3(64), 0, s, n, unitg, nion, r1, ex1, ex2, ex3);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ------------------------------------------------------ density + LDA */

__device__ __forceinline__ void lda_point(double d, double &xcp, double &xce)
{
   /* Slater exchange, ccpsdv3/dirac_exchange.c:84-91 */
   const double alpha = 0.6666666666666666667;
   const double onethird_x = 0.33333333333333333333;
   const double pi_x = 3.14159265358979311600;
   const double nn = d + d;
   const double n13 = pow(3.0 * nn / pi_x, onethird_x);
   const double ex = -(9.0 / 8.0) * alpha * n13;
   const double vx = -(3.0 / 2.0) * alpha * n13;
   /* VWN correlation, ccpsdv3/vosko.c:99-167 with nup = ndown (xi = 0) */
   const double bp = 3.727440e+00, cp = 1.293520e+01;
   const double cp1 = 3.109070e-02, cp2 = 9.690228e-04, cp3 = 1.049800e-01;
   const double cp4 = 3.878329e-02, cp5 = 3.075995e+00, cp6 = 1.863720e+00;
   const double dp1 = 6.218140e-02, dp2 = 1.938045e-03;
   const double dp4 = -3.205972e-02, dp5 = -1.192972e-01, dp6 = 1.863720e+00;
   const double dp7 = 9.461748e+00;
   const double onesixth = 0.166666666666667, onethird = 0.333333333333333;
   const double rs_scale = 0.620350490899400;
   const double nup = d + 1.0e-30;
   const double nt = nup + nup;
   const double rs = rs_scale / pow(nt, onethird);
   const double x = sqrt(rs);
   const double xxp = rs + bp * x + cp;
   const double dxxp = 2.0 * x + bp;
   const double ecp = cp1 * log(rs / xxp) + cp2 * log((x + cp3) * (x + cp3) / xxp)
                    + cp4 * atan(cp5 / (x + cp6));
   const double ucp = ecp - onesixth * x * (dp1 / x + dp2 / (x + cp3) + dp4 * dxxp / xxp
                    + dp5 / ((x + dp6) * (x + dp6) + dp7));
   xce = ex + ecp;
   xcp = vx + ucp;
}

__global__ void k_dn_reduce_xc(const double *__restrict__ dnpart, int nparts, size_t n3, double scal2,
                               double *__restrict__ dn, double *__restrict__ xcp, double *__restrict__ xce)
{
   const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (g >= n3) return;
   /* the parts added in order, loaded four at a time */
   double s = 0.0;
   int b = 0;
   for (; b + 4 <= nparts; b += 4) {
      const double v0 = dnpart[(size_t)b * n3 + g], v1 = dnpart[(size_t)(b + 1) * n3 + g];
      const double v2 = dnpart[(size_t)(b + 2) * n3 + g], v3 = dnpart[(size_t)(b + 3) * n3 + g];
      s += v0;
      s += v1;
      s += v2;
      s += v3;
   }
   for (; b < nparts; ++b) s += dnpart[(size_t)b * n3 + g];
   const double d = s * scal2;
   dn[g] = d;
   double p, e;
   lda_point(d, p, e);
   xcp[g] = p;
   xce[g] = e;
}

/* standalone LDA on a padded n2ft3d array (test hook of the XC kernel) */
__global__ void k_lda(const double *__restrict__ dn, size_t n, double *__restrict__ xcp, double *__restrict__ xce)
{
   const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (g >= n) return;
   double p, e;
   lda_point(dn[g], p, e);
   xcp[g] = p;
   xce[g] = e;
}

/* --------------------------------------------------------- reductions */

template<int NT>
__device__ __forceinline__ double block_sum(double v, double *red)
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

/* per block: sums over packed G of w|s|^2 vg, w|dng|^2 vc, w Re(vall* dng) */
__global__ void __launch_bounds__(256)
k_gsums(const double2 *__restrict__ s, const double2 *__restrict__ dng, const double2 *__restrict__ vall,
        const double *__restrict__ vg, const double *__restrict__ vc, const double *__restrict__ w,
        int ngp, double *__restrict__ part)
{
   __shared__ double red[256];
   double a = 0.0, b = 0.0, c = 0.0;
   for (int p = blockIdx.x * 256 + threadIdx.x; p < ngp; p += gridDim.x * 256) {
      const double wp = w[2 * p];
      const double2 sp = s[p], dp = dng[p], vp = vall[p];
      a += wp * ((sp.x * sp.x + sp.y * sp.y) * vg[p]);
      b += wp * ((dp.x * dp.x + dp.y * dp.y) * vc[p]);
      c += wp * (vp.x * dp.x + vp.y * dp.y);
   }
   a = block_sum<256>(a, red);
   b = block_sum<256>(b, red);
   c = block_sum<256>(c, red);
   if (threadIdx.x == 0) { part[3 * blockIdx.x] = a; part[3 * blockIdx.x + 1] = b; part[3 * blockIdx.x + 2] = c; }
}

/* per block: sum dn*xce, dn*xcp over the dense grid */
__global__ void __launch_bounds__(256)
k_xcsums(const double *__restrict__ dn, const double *__restrict__ xce, const double *__restrict__ xcp,
         size_t n3, double *__restrict__ part)
{
   __shared__ double red[256];
   double a = 0.0, b = 0.0;
   for (size_t g = (size_t)blockIdx.x * 256 + threadIdx.x; g < n3; g += (size_t)gridDim.x * 256) {
      a += dn[g] * xce[g];
      b += dn[g] * xcp[g];
   }
   a = block_sum<256>(a, red);
   b = block_sum<256>(b, red);
   if (threadIdx.x == 0) { part[2 * blockIdx.x] = a; part[2 * blockIdx.x + 1] = b; }
}

/* per orbital: eke_i = sum w tg |c1|^2 ; cp = c2 - c1 ; dc_i = sum w |cp|^2.
 * grid (blocks, ne): partials [ne][blocks][2] */
__global__ void __launch_bounds__(256)
k_orbsums(const double2 *__restrict__ c1, const double2 *__restrict__ c2, double2 *__restrict__ cp,
          const double *__restrict__ tg, const double *__restrict__ w, int ngp, double *__restrict__ part)
{
   __shared__ double red[256];
   const int i = blockIdx.y;
   const size_t off = (size_t)i * ngp;
   double a = 0.0, b = 0.0;
   for (int p = blockIdx.x * 256 + threadIdx.x; p < ngp; p += gridDim.x * 256) {
      const double2 x = c1[off + p], y = c2[off + p];
      const double2 d = make_double2(y.x - x.x, y.y - x.y);
      cp[off + p] = d;
      const double wp = w[2 * p];
      a += wp * (x.x * (x.x * tg[p]) + x.y * (x.y * tg[p]));
      b += wp * (d.x * d.x + d.y * d.y);
   }
   a = block_sum<256>(a, red);
   b = block_sum<256>(b, red);
   if (threadIdx.x == 0) {
      part[((size_t)i * gridDim.x + blockIdx.x) * 2] = a;
      part[((size_t)i * gridDim.x + blockIdx.x) * 2 + 1] = b;
   }
}

/* erfc of ccpsdv3/cpsd.c:11-17, used as (1/aa^4)^4 at :541-543 */
__device__ __forceinline__ double erfc_as(double w)
{
   const double B1 = 0.0705230784, B2 = 0.0422820123, B3 = 0.0092705272;
   const double B4 = 0.0001520143, B5 = 0.0002765672, B6 = 0.0000430638;
   const double aa = 1.0 + w * (B1 + w * (B2 + w * (B3 + w * (B4 + w * (B5 + w * B6)))));
   const double e = 1.0 / (aa * aa * aa * aa);
   return e * e * e * e;
}

/* real-space Ewald pairs (ccpsdv3/cpsd.c:527-545): one thread per pair
 * (ii, jj > ii) summing its nsh images in order; part[ii*nion + jj]
 * (zero for jj <= ii) */
__global__ void __launch_bounds__(256)
k_ewald_real(const double *__restrict__ r1, const int *__restrict__ katm, const double *__restrict__ zv,
             int nion, const double *__restrict__ rcell, int nsh, double rcut, double *__restrict__ part)
{
   const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= (size_t)nion * nion) return;
   const int ii = (int)(t / nion), jj = (int)(t % nion);
   double e = 0.0;
   if (jj > ii) {
      const double dx = r1[3 * ii] - r1[3 * jj];
      const double dy = r1[3 * ii + 1] - r1[3 * jj + 1];
      const double dz = r1[3 * ii + 2] - r1[3 * jj + 2];
      const double zz = zv[katm[ii]] * zv[katm[jj]];
      for (int l = 0; l < nsh; ++l) {
         const double x = rcell[l] + dx, y = rcell[l + nsh] + dy, z = rcell[l + 2 * nsh] + dz;
         const double r = sqrt(x * x + y * y + z * z);
         e += zz * erfc_as(r / rcut) / r;
      }
   }
   part[t] = e;
}

__global__ void k_erfc_as(const double *__restrict__ w, int n, double *__restrict__ out)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) out[i] = erfc_as(w[i]);
}

/* ------------------------------------------- overlaps -> s matrices */

/* Gram matrix Gm (2ne x 2ne, row-major) of the two orbital blocks (C1 first
 * unless c1_second) -> the four ne x ne
 * matrices of ccpsdv3/cpsd.c:415-437 (column-major).  s21 and s12 are each
 * filled from one triangle and mirrored, as the reference does. */
__global__ void k_smat(const double *__restrict__ Gm, const double *__restrict__ g11c, int ne, int c1_second,
                       double dte, double *__restrict__ s11, double *__restrict__ s12, double *__restrict__ s21,
                       double *__restrict__ s22, unsigned *__restrict__ bar)
{
   const int idx = blockIdx.x * blockDim.x + threadIdx.x;
   if (bar && idx == 0) *bar = 0u;
   if (idx >= ne * ne) return;
   const int r = idx % ne, c = idx / ne;   /* element (r, c) at r + c*ne */
   const int i = min(r, c), j = max(r, c);
   const int n2 = 2 * ne;
   const int o1 = c1_second ? ne : 0, o2 = c1_second ? 0 : ne;   /* row offsets of C1, C2 */
   const double g11 = g11c ? g11c[(size_t)i * ne + j] : Gm[(size_t)(o1 + i) * n2 + o1 + j];
   const double g22 = Gm[(size_t)(o2 + i) * n2 + o2 + j];
   const double g21 = Gm[(size_t)(o2 + i) * n2 + o1 + j];      /* <c2_i|c1_j> */
   const double g12 = Gm[(size_t)(o1 + i) * n2 + o2 + j];      /* <c1_i|c2_j> */
   if (i == j) {
      s22[idx] = (1.0 - g22) * 0.5 / dte;
      s21[idx] = (1.0 - g21) * 0.5;
      s12[idx] = s21[idx];
      s11[idx] = -g11 * 0.5 * dte;
   } else {
      s22[idx] = -g22 * 0.5 / dte;
      s21[idx] = -g21 * 0.5;
      s12[idx] = -g12 * 0.5;
      s11[idx] = -g11 * 0.5 * dte;
   }
}

/* The preconditioned step's overlaps.  K o R is not orthogonal to c1 where
 * K varies, so <c1|c2> moves away from the identity and the Lagrange fixed
 * point above stops converging; the step's component along c1 is taken out
 * first: c2' = c2 - c1 D with D = <c1|c2> - I, and the s matrices are formed
 * from
 *   A12' = A12 - A11 D,   A22' = A22 - Q - Q^T + D^T A11 D,   Q = D^T A12
 * (A_ab(i, j) = <a_i|b_j>), in full, with no mirroring.  The constraint then
 * adds dte c1 (X - D/dte) to the unprojected c2 (k_precond_lmd).  First
 * kernel: D, P = A11 D and Q, column-major, each sum over k in order.  A
 * wave holds 64 consecutive i at one j: A11(k, i) (symmetric) and
 * A12(k, i) are consecutive in memory, A12(k, j) is one address. */
__global__ void __launch_bounds__(256) k_precond_dpq(const double *__restrict__ Gm, const double *__restrict__ g11c,
                                                     int ne, int c1_second, double *__restrict__ D,
                                                     double *__restrict__ P, double *__restrict__ Q)
{
   const int idx = blockIdx.x * blockDim.x + threadIdx.x;
   if (idx >= ne * ne) return;
   const int i = idx % ne, j = idx / ne;
   const int n2 = 2 * ne, o1 = c1_second ? ne : 0, o2 = c1_second ? 0 : ne;
   const double *a11 = g11c ? g11c + i : Gm + (size_t)o1 * n2 + o1 + i;
   const size_t s11 = g11c ? (size_t)ne : (size_t)n2;
   const double *a12 = Gm + (size_t)o1 * n2 + o2;      /* A12(k, c) = a12[k n2 + c] */
   double p = 0.0, q = 0.0;
   for (int k = 0; k < ne; ++k) {
      const double dkj = a12[(size_t)k * n2 + j] - (k == j ? 1.0 : 0.0);
      const double dki = a12[(size_t)k * n2 + i] - (k == i ? 1.0 : 0.0);
      p += a11[k * s11] * dkj;
      q += dki * a12[(size_t)k * n2 + j];
   }
   D[idx] = a12[(size_t)i * n2 + j] - (i == j ? 1.0 : 0.0);
   P[idx] = p;
   Q[idx] = q;
}

/* second kernel: R = D^T P and the s matrices of the projected overlaps
 * (column-major, the layout of k_smat); also clears the Lagrange kernel's
 * barrier counter */
__global__ void __launch_bounds__(256) k_precond_smat(const double *__restrict__ Gm, const double *__restrict__ g11c,
                                                      int ne, int c1_second, double dte,
                                                      const double *__restrict__ P, const double *__restrict__ Q,
                                                      double *__restrict__ s11, double *__restrict__ s12,
                                                      double *__restrict__ s21, double *__restrict__ s22,
                                                      unsigned *__restrict__ bar)
{
   const int idx = blockIdx.x * blockDim.x + threadIdx.x;
   if (bar && idx == 0) *bar = 0u;
   if (idx >= ne * ne) return;
   const int i = idx % ne, j = idx / ne;
   const int n2 = 2 * ne, o1 = c1_second ? ne : 0, o2 = c1_second ? 0 : ne;
   const double *a12 = Gm + (size_t)o1 * n2 + o2;
   const double *pj = P + (size_t)j * ne;
   double r = 0.0;
   for (int k = 0; k < ne; ++k) r += (a12[(size_t)k * n2 + i] - (k == i ? 1.0 : 0.0)) * pj[k];
   const double a22 = Gm[(size_t)(o2 + i) * n2 + o2 + j] - Q[idx] - Q[(size_t)i * ne + j] + r;
   const double a12p = a12[(size_t)i * n2 + j] - P[idx];
   const double a21p = a12[(size_t)j * n2 + i] - P[(size_t)i * ne + j];
   const double g11 = g11c ? g11c[(size_t)i * ne + j] : Gm[(size_t)(o1 + i) * n2 + o1 + j];
   const double d = i == j ? 1.0 : 0.0;
   s22[idx] = (d - a22) * 0.5 / dte;
   s21[idx] = (d - a21p) * 0.5;
   s12[idx] = (d - a12p) * 0.5;
   s11[idx] = -g11 * 0.5 * dte;
}

/* lmd <- X - D/dte: the Lagrange matrix of the unprojected c2 */
__global__ void k_precond_lmd(double *__restrict__ lmd, const double *__restrict__ D, int nn, double dte)
{
   const int idx = blockIdx.x * blockDim.x + threadIdx.x;
   if (idx < nn) lmd[idx] -= D[idx] / dte;
}

/* ------------------------------------------ 16 x 16 tiles on the MFMA */

/* acc += P Q over k in [kb, kb + 4n) for a 16 x 16 tile on the FP64 4x4x4
 * MFMA (mfma64.h); Pr[i][k] = P(i, k) and Qc[j][k] = Q(k, j) in LDS.  The
 * result lands in the rotated layout that tile16_reduce undoes. */
template<int LP>
__device__ __forceinline__ void tile16_mfma(double (*Pr)[LP], double (*Qc)[LP], int kb, int ke, int lane,
                                            double (&acc)[4])
{
   for (int m0 = kb; m0 < ke; m0 += 4) {
      const int m = m0 + (lane >> 4);
      const double pa = Pr[lane & 15][m];
      double bq[4];
      rotations(Qc[lane & 15][m], bq);
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = mfma4(pa, bq[t], acc[t]);
   }
}

/* the four waves' partial tiles (each over a quarter of K) added in wave
 * order at element (rl, cl) of the tile; all 256 threads call it */
__device__ __forceinline__ double tile16_reduce(double (*part)[16][17], const double (&acc)[4], int wave, int lane,
                                                int s1, int rl, int cl)
{
   const int q = lane >> 4, ab = (lane >> 2) & 3, bb = lane & 3;
#pragma unroll
   for (int t = 0; t < 4; ++t) part[wave][4 * ab + q][4 * ((ab + s1 * t) & 3) + bb] = acc[t];
   __syncthreads();
   const double v = ((part[0][rl][cl] + part[1][rl][cl]) + part[2][rl][cl]) + part[3][rl][cl];
   __syncthreads();
   return v;
}

/* W = G11 lmd and T1 = G21 lmd for ne <= NM, one 16 x 16 tile of each per
 * workgroup (tile (ti, tj): rows i, columns b), the sums over j on the MFMA
 * with the four waves taking a quarter each.  Operands as in k_gram_next_w
 * below: G11(i, j) = g11c[j ne + i] (symmetric) or the Gram's C1 x C1 block,
 * G21(i, j) = <c1_j|c2_i> from the Gram's C1 x C2 block, lmd(j, b) =
 * lmd[j + b ne]. */
template<int NM>
__global__ void __launch_bounds__(256) k_gram_next_w_mf(const double *__restrict__ Gm,
                                                        const double *__restrict__ g11c, int ne, int c1_second,
                                                        const double *__restrict__ lmd, double *__restrict__ W,
                                                        double *__restrict__ T1)
{
   constexpr int TS = 16, P = NM + 2;
   __shared__ double g11r[TS][P], g21r[TS][P], lmc[TS][P];
   __shared__ double part[4][TS][TS + 1];
   const int nt = (ne + TS - 1) / TS;
   const int ti = blockIdx.x % nt, tj = blockIdx.x / nt;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int kp = (ne + 15) & ~15, kq = kp / 4, kb = wave * kq, ke = kb + kq;
   const int s1 = rot_step(lane);
   const int n2 = 2 * ne, o1 = c1_second ? ne : 0, o2 = c1_second ? 0 : ne;
   const double *gb = g11c ? g11c : Gm + (size_t)o1 * n2 + o1;
   const size_t gld = g11c ? (size_t)ne : (size_t)n2;
   const double *tb = Gm + (size_t)o1 * n2 + o2;
   constexpr int LQ = TS * NM / 256;
   static_assert(TS * NM % 256 == 0, "panel loads cover TS x NM in whole rounds");
   /* row panels: R[r][k] = base[k ld + ti*TS + r]; column panel: lmc[c][k] =
    * lmd(k, tj*TS + c); zero outside ne, up to the padded K */
   double vg[LQ], vt[LQ], vl[LQ];
#pragma unroll
   for (int u = 0; u < LQ; ++u) {
      const int x = tid + 256 * u;
      const int r = x % TS, k = x / TS, row = ti * TS + r;
      const bool in = row < ne && k < ne;
      vg[u] = in ? gb[(size_t)k * gld + row] : 0.0;
      vt[u] = in ? tb[(size_t)k * n2 + row] : 0.0;
      const int kk = x % NM, c = x / NM, col = tj * TS + c;
      vl[u] = (col < ne && kk < ne) ? lmd[kk + (size_t)col * ne] : 0.0;
   }
#pragma unroll
   for (int u = 0; u < LQ; ++u) {
      const int x = tid + 256 * u;
      if (x / TS < kp) {
         g11r[x % TS][x / TS] = vg[u];
         g21r[x % TS][x / TS] = vt[u];
      }
      if (x % NM < kp) lmc[x / NM][x % NM] = vl[u];
   }
   __syncthreads();
   double aw[4] = {0.0, 0.0, 0.0, 0.0}, at[4] = {0.0, 0.0, 0.0, 0.0};
   tile16_mfma<P>(g11r, lmc, kb, ke, lane, aw);
   tile16_mfma<P>(g21r, lmc, kb, ke, lane, at);
   const int rl = tid % TS, cl = tid / TS, i = ti * TS + rl, b = tj * TS + cl;
   const double w = tile16_reduce(part, aw, wave, lane, s1, rl, cl);
   const double t = tile16_reduce(part, at, wave, lane, s1, rl, cl);
   if (i < ne && b < ne) {
      W[(size_t)i * ne + b] = w;
      T1[(size_t)i * ne + b] = t;
   }
}

/* W = G11 lmd and T1 = G21 lmd (G21[i][j] = <c2_i|c1_j>), both stored as
 * [i][b].  A workgroup forms 64 outputs (i, b), i fastest; its four thread
 * groups sum over the four quarters of j and the quarters are added in a
 * fixed order.  G11[j][i] (symmetric) and the mirrored Gram entry
 * <c1_j|c2_i> make a wave read consecutive entries; lmd(j, b) = lmd[j + b ne]
 * is the same for the whole wave. */
__global__ void __launch_bounds__(256) k_gram_next_w(const double *__restrict__ Gm, const double *__restrict__ g11c,
                                                     int ne, int c1_second, const double *__restrict__ lmd,
                                                     double *__restrict__ W, double *__restrict__ T1)
{
   __shared__ double red[2][4][64];
   const int g = threadIdx.x >> 6, l = threadIdx.x & 63;
   const int idx = blockIdx.x * 64 + l;
   const bool valid = idx < ne * ne;
   const int i = idx % ne, b = idx / ne;
   const int n2 = 2 * ne, o1 = c1_second ? ne : 0, o2 = c1_second ? 0 : ne;
   double w = 0.0, t = 0.0;
   if (valid) {
      const int q = ne / 4, j0 = g * q;
      const double *gp = g11c ? g11c + i : Gm + (size_t)o1 * n2 + o1 + i;
      const size_t gs = g11c ? (size_t)ne : (size_t)n2;
      const double *tp = Gm + (size_t)o1 * n2 + o2 + i;
      const double *lp = lmd + (size_t)b * ne;
      /* unrolled so that a batch of independent loads is in flight at once */
#pragma unroll 8
      for (int j = j0; j < j0 + q; ++j) {
         const double lv = lp[j];
         w += gp[j * gs] * lv;
         t += tp[(size_t)j * n2] * lv;
      }
   }
   red[0][g][l] = w;
   red[1][g][l] = t;
   __syncthreads();
   if (g == 0 && valid) {
      W[(size_t)i * ne + b] = (red[0][0][l] + red[0][1][l]) + (red[0][2][l] + red[0][3][l]);
      T1[(size_t)i * ne + b] = (red[1][0][l] + red[1][1][l]) + (red[1][2][l] + red[1][3][l]);
   }
}

/* out[a][b] = G22[a][b] + dte (T1[a][b] + T1[b][a]) + dte^2 sum_i lmd(i,a) W[i][b]
 * (T1[b][a] = sum_i lmd(i,a) G21[b][i]); 64 outputs per workgroup, b
 * fastest, the sum over i in four fixed quarters */
__global__ void __launch_bounds__(256) k_gram_next(const double *__restrict__ Gm, int ne, int c1_second,
                                                   const double *__restrict__ lmd, double dte,
                                                   const double *__restrict__ W, const double *__restrict__ T1,
                                                   double *__restrict__ out)
{
   __shared__ double red[4][64];
   const int g = threadIdx.x >> 6, l = threadIdx.x & 63;
   const int idx = blockIdx.x * 64 + l;
   const bool valid = idx < ne * ne;
   const int b = idx % ne, a = idx / ne;
   double t3 = 0.0;
   if (valid) {
      const int q = ne / 4, i0 = g * q;
      const double *la = lmd + (size_t)a * ne;
#pragma unroll 8
      for (int i = i0; i < i0 + q; ++i) t3 += la[i] * W[(size_t)i * ne + b];
   }
   red[g][l] = t3;
   __syncthreads();
   if (g == 0 && valid) {
      const int n2 = 2 * ne, o2 = c1_second ? 0 : ne;
      const double s3 = (red[0][l] + red[1][l]) + (red[2][l] + red[3][l]);
      out[(size_t)a * ne + b] = Gm[(size_t)(o2 + a) * n2 + o2 + b] +
                                dte * (T1[(size_t)a * ne + b] + T1[(size_t)b * ne + a]) + dte * dte * s3;
   }
}

/* P (nproj x ne, row-major, P[m][n] = gcdotc(beta_m, c1_n)) -> the apply
 * coefficients X[m + n*nproj] = -P scal2 / vnlnrm_m (ccpsdv3/cpsd.c:289) */
__global__ void k_kbcoef(const double *__restrict__ P, int nproj, int ne, const double *__restrict__ pnrm,
                         double scal2, double *__restrict__ X)
{
   const int idx = blockIdx.x * blockDim.x + threadIdx.x;
   if (idx >= nproj * ne) return;
   const int m = idx % nproj, n = idx / nproj;
   X[idx] = -P[(size_t)m * ne + n] * scal2 / pnrm[m];
}

/* ---------------------------------------- Lagrange multiplier fixed point */

/* thread 0 publishes the workgroup's writes (agent-scope release), arrives,
 * and polls with relaxed agent-scope loads, which read past the XCD's L2
 * without invalidating it; one acquire fence after the poll makes the other
 * workgroups' writes visible to the whole workgroup */
__device__ __forceinline__ void grid_barrier(unsigned *bar, unsigned target, int *err)
{
   __syncthreads();
   if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      atomicAdd(bar, 1u);
      unsigned spins = 0;
      while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
         if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
         __builtin_amdgcn_s_sleep(4);
         if (++spins > (1u << 21)) { atomicExch(err, 1); break; }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
   }
   __syncthreads();
}

/* a grid barrier, or a workgroup barrier when the grid is one workgroup
 * (global writes of the workgroup are then visible after __syncthreads) */
__device__ __forceinline__ void lag_barrier(unsigned *bar, unsigned nbar, int nwg, int *err)
{
   if (nwg == 1) __syncthreads();
   else grid_barrier(bar, nbar * nwg, err);
}

/* the largest per-workgroup maximum: thread t reads entries t, t + 256, ...
 * (loads in flight together rather than one dependent load per workgroup),
 * then a tree over the 256 threads; every workgroup reduces the same values
 * and max does not depend on order, so all of them take the same decision */
__device__ __forceinline__ double grid_max(const double *wgmax, int nwg, double *red)
{
   double v = 0.0;
   for (int q = threadIdx.x; q < nwg; q += 256)
      v = fmax(v, __hip_atomic_load(wgmax + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
   red[threadIdx.x] = v;
   __syncthreads();
   for (int s2 = 128; s2 > 0; s2 >>= 1) {
      if ((int)threadIdx.x < s2) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s2]);
      __syncthreads();
   }
   const double g = red[0];
   __syncthreads();
   return g;
}

/* C_tile += A * B over k (column-major ne x ne), TS x TS tile at (ti, tj);
 * one or more outputs per thread accumulated in acc[]. */
template<int TS, int PER>
__device__ __forceinline__ void tile_mm(const double *A, const double *B, int ne, int ti, int tj,
                                        double (&acc)[PER], double (*sa)[33], double (*sb)[33])
{
   /* chunks of 32 along k: sa[r][kk] = A[ti*TS + r, k0+kk], sb[c][kk] = B[k0+kk, tj*TS + c] */
   /* 256 threads; each chunk's loads are issued together before the LDS stores */
   constexpr int CQ = TS * 32 / 256;
   for (int k0 = 0; k0 < ne; k0 += 32) {
      double va[CQ], vb[CQ];
#pragma unroll
      for (int u = 0; u < CQ; ++u) {
         const int e = threadIdx.x + 256 * u;
         const int r = e % TS, kk = e / TS;
         const int row = ti * TS + r, k = k0 + kk;
         va[u] = (row < ne && k < ne) ? A[row + (size_t)k * ne] : 0.0;
         const int kb = e % 32, c = e / 32;
         const int col = tj * TS + c, k2 = k0 + kb;
         vb[u] = (col < ne && k2 < ne) ? B[k2 + (size_t)col * ne] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < CQ; ++u) {
         const int e = threadIdx.x + 256 * u;
         sa[e % TS][e / TS] = va[u];
         sb[e / 32][e % 32] = vb[u];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PER; ++q) {
         const int o = threadIdx.x + q * blockDim.x;
         const int r = o % TS, c = o / TS;
         double s = 0.0;
#pragma unroll 8
         for (int kk = 0; kk < 32; ++kk) s += sa[r][kk] * sb[c][kk];
         acc[q] += s;
      }
      __syncthreads();
   }
}

/* X <- s22 + s21 X + X s12 + X (s11 X), from X0 = s22, until max|dX| < 1e-15
 * or a.itmax iterations (30 for the reference's step: ccpsdv3/cpsd.c:439-478,
 * ITERLMD/CONVGLMD at :19-20).
 * Persistent grid of (ne/TS)^2 workgroups (all co-resident) with a bounded
 * spinning grid barrier; info[0] = iterations, info[1] = last max|dX|. */
template<int TS>
__global__ void __launch_bounds__(256) k_lagrange(LagArgs a)
{
   constexpr int PER = TS * TS / 256;
   __shared__ double sa[TS][33];
   __shared__ double sb[TS][33];
   __shared__ double red[256];
   const int nt = (a.ne + TS - 1) / TS;
   const int ti = blockIdx.x % nt, tj = blockIdx.x / nt;
   const int nwg = gridDim.x;
   const size_t nn = (size_t)a.ne * a.ne;
   unsigned nbar = 0;
   int it = 0;
   double gmax = 0.0;
   /* X_0 = s22 */
   for (int q = 0; q < PER; ++q) {
      const int o = threadIdx.x + q * 256, r = ti * TS + o % TS, c = tj * TS + o / TS;
      if (r < a.ne && c < a.ne) a.X[r + (size_t)c * a.ne] = a.s22[r + (size_t)c * a.ne];
   }
   lag_barrier(a.bar, ++nbar, nwg, a.err);
   while (it < a.itmax) {
      const double *X = a.X + (size_t)(it & 1) * nn;
      double *Xn = a.X + (size_t)((it + 1) & 1) * nn;
      ++it;
      double acc[PER];
      for (int q = 0; q < PER; ++q) acc[q] = 0.0;
      tile_mm<TS, PER>(a.s11, X, a.ne, ti, tj, acc, sa, sb);
      for (int q = 0; q < PER; ++q) {
         const int o = threadIdx.x + q * 256, r = ti * TS + o % TS, c = tj * TS + o / TS;
         if (r < a.ne && c < a.ne) a.st1[r + (size_t)c * a.ne] = acc[q];
      }
      lag_barrier(a.bar, ++nbar, nwg, a.err);
      for (int q = 0; q < PER; ++q) acc[q] = 0.0;
      tile_mm<TS, PER>(a.s21, X, a.ne, ti, tj, acc, sa, sb);
      tile_mm<TS, PER>(X, a.s12, a.ne, ti, tj, acc, sa, sb);
      tile_mm<TS, PER>(X, a.st1, a.ne, ti, tj, acc, sa, sb);
      double lmax = 0.0;
      for (int q = 0; q < PER; ++q) {
         const int o = threadIdx.x + q * 256, r = ti * TS + o % TS, c = tj * TS + o / TS;
         if (r < a.ne && c < a.ne) {
            const size_t e = r + (size_t)c * a.ne;
            const double v = a.s22[e] + acc[q];
            Xn[e] = v;
            lmax = fmax(lmax, fabs(v - X[e]));
         }
      }
      red[threadIdx.x] = lmax;
      __syncthreads();
      for (int s = 128; s > 0; s >>= 1) {
         if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
         __syncthreads();
      }
      if (threadIdx.x == 0) a.wgmax[blockIdx.x] = red[0];
      lag_barrier(a.bar, ++nbar, nwg, a.err);
      gmax = grid_max(a.wgmax, nwg, red);
      if (gmax < 1e-15) break;
      if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
   }
   /* lmd = the last iterate (sa1 of the final iteration) */
   const double *Xf = a.X + (size_t)(it & 1) * nn;
   for (int q = 0; q < PER; ++q) {
      const int o = threadIdx.x + q * 256, r = ti * TS + o % TS, c = tj * TS + o / TS;
      if (r < a.ne && c < a.ne) a.lmd[r + (size_t)c * a.ne] = Xf[r + (size_t)c * a.ne];
   }
   if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.info[0] = (double)it;
      a.info[1] = gmax;
   }
}

/* ne <= NM: each workgroup owns one 16 x 16 tile of X and keeps its
 * constant panels (rows ti of s11 and s21, columns tj of s12) in LDS for all
 * iterations; an iteration loads the X panels and, after the first barrier,
 * the st1 panel, i.e. two round trips to L2 and two grid barriers per
 * iteration (the general kernel above makes one round trip per 32-wide K
 * chunk of each of its four products).  The tile products run on the FP64
 * 4x4x4 MFMA (mfma64.h): the four waves take a quarter of K each and their
 * 16 x 16 partial tiles are added in wave order, so results do not depend on
 * timing.  Matrices are column-major: M(r, c) = M[r + c*ne]. */
template<int NM>
__global__ void __launch_bounds__(256) k_lagrange_res(LagArgs a)
{
   constexpr int TS = 16;
   constexpr int P = NM + 2;   /* row pitch: a 64-bit read of 16 rows x 2 k hits distinct banks */
   extern __shared__ double lpan[];
   double (*s11r)[P] = reinterpret_cast<double (*)[P]>(lpan);
   double (*s21r)[P] = s11r + TS;
   double (*s12c)[P] = s21r + TS;
   double (*xc)[P] = s12c + TS;
   double (*xr)[P] = xc + TS;
   double (*stc)[P] = xr + TS;
   __shared__ double part[4][TS][TS + 1];
   __shared__ double red[256];
   const int ne = a.ne, nt = (ne + TS - 1) / TS;
   const int ti = blockIdx.x % nt, tj = blockIdx.x / nt;
   const int nwg = gridDim.x;
   const size_t nn = (size_t)ne * ne;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   /* K padded to whole 4-wide steps for each of the four waves */
   const int kp = (ne + 15) & ~15, kq = kp / 4, kb = wave * kq, ke = kb + kq;
   const int s1 = rot_step(lane);
   /* this thread's element of the tile */
   const int rl = tid % TS, cl = tid / TS, r = ti * TS + rl, c = tj * TS + cl;
   const bool mine = r < ne && c < ne;
   const size_t e = r + (size_t)c * ne;

   /* panel loads: all LQ loads of a thread are issued before any LDS store */
   constexpr int LQ = TS * NM / 256;
   static_assert(TS * NM % 256 == 0, "panel loads cover TS x NM in whole rounds");
   /* R[r][k] = M(ti*TS + r, k), k < ne */
   auto load_rows = [&](const double *M, double (*R)[P]) {
      double v[LQ];
#pragma unroll
      for (int u = 0; u < LQ; ++u) {
         const int x = tid + 256 * u;
         const int rr = x % TS, k = x / TS, row = ti * TS + rr;
         v[u] = (row < ne && k < ne) ? M[row + (size_t)k * ne] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < LQ; ++u) {
         const int x = tid + 256 * u;
         if (x / TS < ne) R[x % TS][x / TS] = v[u];
      }
   };
   /* Cp[c][k] = M(k, tj*TS + c), k < ne */
   auto load_cols = [&](const double *M, double (*Cp)[P]) {
      double v[LQ];
#pragma unroll
      for (int u = 0; u < LQ; ++u) {
         const int x = tid + 256 * u;
         const int k = x % NM, cc = x / NM, col = tj * TS + cc;
         v[u] = (col < ne && k < ne) ? M[k + (size_t)col * ne] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < LQ; ++u) {
         const int x = tid + 256 * u;
         if (x % NM < ne) Cp[x / NM][x % NM] = v[u];
      }
   };
   /* acc += P Q over this wave's K quarter; the four waves' tiles summed */
   auto tile = [&](double (*Pr)[P], double (*Qc)[P], double (&acc)[4]) {
      tile16_mfma<P>(Pr, Qc, kb, ke, lane, acc);
   };
   auto reduce = [&](const double (&acc)[4]) { return tile16_reduce(part, acc, wave, lane, s1, rl, cl); };

   /* K padding read by the MFMA steps stays zero: the loads write k < ne only */
   for (int x = tid; x < 6 * TS * (kp - ne); x += 256) {
      const int pnl = x / (TS * (kp - ne)), y = x % (TS * (kp - ne));
      (s11r + pnl * TS)[y / (kp - ne)][ne + y % (kp - ne)] = 0.0;
   }
   unsigned nbar = 0;
   int it = 0;
   double gmax = 0.0;
   load_rows(a.s11, s11r);
   load_rows(a.s21, s21r);
   load_cols(a.s12, s12c);
   const double s22v = mine ? a.s22[e] : 0.0;
   /* X_0 = s22 */
   if (mine) a.X[e] = s22v;
   lag_barrier(a.bar, ++nbar, nwg, a.err);
   while (it < a.itmax) {
      const double *X = a.X + (size_t)(it & 1) * nn;
      double *Xn = a.X + (size_t)((it + 1) & 1) * nn;
      ++it;
      load_cols(X, xc);
      load_rows(X, xr);
      __syncthreads();
      /* st1 = s11 X */
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      tile(s11r, xc, acc);
      const double st = reduce(acc);
      if (mine) a.st1[e] = st;
      lag_barrier(a.bar, ++nbar, nwg, a.err);
      load_cols(a.st1, stc);
      __syncthreads();
      /* X_new = s22 + s21 X + X s12 + X st1 */
      acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
      tile(s21r, xc, acc);
      tile(xr, s12c, acc);
      tile(xr, stc, acc);
      const double v = s22v + reduce(acc);
      double lmax = 0.0;
      if (mine) {
         Xn[e] = v;
         lmax = fabs(v - xc[cl][r]);   /* X(r, c) */
      }
      red[tid] = lmax;
      __syncthreads();
      for (int s2 = 128; s2 > 0; s2 >>= 1) {
         if (tid < s2) red[tid] = fmax(red[tid], red[tid + s2]);
         __syncthreads();
      }
      if (tid == 0) a.wgmax[blockIdx.x] = red[0];
      lag_barrier(a.bar, ++nbar, nwg, a.err);
      gmax = grid_max(a.wgmax, nwg, red);
      if (gmax < 1e-15) break;
      if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
   }
   const double *Xf = a.X + (size_t)(it & 1) * nn;
   if (mine) a.lmd[e] = Xf[e];
   if (blockIdx.x == 0 && tid == 0) {
      a.info[0] = (double)it;
      a.info[1] = gmax;
   }
}


/* 128 < ne <= NM: the kernel above, with the panels streamed through three
 * LDS slots instead of kept (six resident panels of 256 would need 198 KB).
 * An iteration loads s11 rows and X columns, forms st1 = s11 X, passes the
 * grid barrier, then forms s21 X (X columns still in their slot), loads the X
 * rows and s12 columns for X s12, and the st1 columns for X st1.  Panels are
 * rows of NM doubles with the k index XOR-swizzled by the row (element
 * (r, k) at column k ^ (r & 15)): a 16-lane group reads 16 rows at one k,
 * which then fall on 16 distinct bank pairs. */
template<int LP>
__device__ __forceinline__ void tile16_mfma_swz(double (*Pr)[LP], double (*Qc)[LP], int kb, int ke, int lane,
                                                double (&acc)[4])
{
   const int sw = lane & 15;
   for (int m0 = kb; m0 < ke; m0 += 4) {
      const int m = (m0 + (lane >> 4)) ^ sw;
      const double pa = Pr[sw][m];
      double bq[4];
      rotations(Qc[sw][m], bq);
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = mfma4(pa, bq[t], acc[t]);
   }
}

template<int NM>
__global__ void __launch_bounds__(256) k_lagrange_str(LagArgs a)
{
   constexpr int TS = 16;
   extern __shared__ double lpan[];
   double (*slA)[NM] = reinterpret_cast<double (*)[NM]>(lpan);
   double (*slB)[NM] = slA + TS;
   double (*slC)[NM] = slB + TS;
   __shared__ double part[4][TS][TS + 1];
   __shared__ double red[256];
   const int ne = a.ne, nt = (ne + TS - 1) / TS;
   const int ti = blockIdx.x % nt, tj = blockIdx.x / nt;
   const int nwg = gridDim.x;
   const size_t nn = (size_t)ne * ne;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int kp = (ne + 15) & ~15, kq = kp / 4, kb = wave * kq, ke = kb + kq;
   const int s1 = rot_step(lane);
   const int rl = tid % TS, cl = tid / TS, r = ti * TS + rl, c = tj * TS + cl;
   const bool mine = r < ne && c < ne;
   const size_t e = r + (size_t)c * ne;
   constexpr int LQ = TS * NM / 256;
   static_assert(TS * NM % 256 == 0, "panel loads cover TS x NM in whole rounds");
   /* R[rr][k ^ rr] = M(ti*TS + rr, k), k < ne */
   auto load_rows = [&](const double *M, double (*R)[NM]) {
      double v[LQ];
#pragma unroll
      for (int u = 0; u < LQ; ++u) {
         const int x = tid + 256 * u;
         const int rr = x % TS, k = x / TS, row = ti * TS + rr;
         v[u] = (row < ne && k < ne) ? M[row + (size_t)k * ne] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < LQ; ++u) {
         const int x = tid + 256 * u;
         const int rr = x % TS, k = x / TS;
         if (k < ne) R[rr][k ^ rr] = v[u];
      }
   };
   /* Cp[cc][k ^ cc] = M(k, tj*TS + cc), k < ne */
   auto load_cols = [&](const double *M, double (*Cp)[NM]) {
      double v[LQ];
#pragma unroll
      for (int u = 0; u < LQ; ++u) {
         const int x = tid + 256 * u;
         const int k = x % NM, cc = x / NM, col = tj * TS + cc;
         v[u] = (col < ne && k < ne) ? M[k + (size_t)col * ne] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < LQ; ++u) {
         const int x = tid + 256 * u;
         const int k = x % NM, cc = x / NM;
         if (k < ne) Cp[cc][k ^ cc] = v[u];
      }
   };
   auto tile = [&](double (*Pr)[NM], double (*Qc)[NM], double (&acc)[4]) {
      tile16_mfma_swz<NM>(Pr, Qc, kb, ke, lane, acc);
   };
   auto reduce = [&](const double (&acc)[4]) { return tile16_reduce(part, acc, wave, lane, s1, rl, cl); };

   /* the K padding k in [ne, kp) of every slot stays zero (the loads write
    * k < ne only; the swizzle keeps each 16-wide block of k in place) */
   for (int x = tid; x < 3 * TS * (kp - ne); x += 256) {
      const int sl = x / (TS * (kp - ne)), y = x % (TS * (kp - ne));
      const int rr = y / (kp - ne), k = ne + y % (kp - ne);
      (slA + sl * TS)[rr][k ^ rr] = 0.0;
   }
   unsigned nbar = 0;
   int it = 0;
   double gmax = 0.0;
   const double s22v = mine ? a.s22[e] : 0.0;
   if (mine) a.X[e] = s22v;
   lag_barrier(a.bar, ++nbar, nwg, a.err);
   while (it < a.itmax) {
      const double *X = a.X + (size_t)(it & 1) * nn;
      double *Xn = a.X + (size_t)((it + 1) & 1) * nn;
      ++it;
      /* st1 = s11 X */
      load_rows(a.s11, slA);
      load_cols(X, slB);
      __syncthreads();
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      tile(slA, slB, acc);
      const double st = reduce(acc);
      if (mine) a.st1[e] = st;
      lag_barrier(a.bar, ++nbar, nwg, a.err);
      /* X_new = s22 + s21 X + X s12 + X st1 */
      load_rows(a.s21, slA);
      __syncthreads();
      acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
      tile(slA, slB, acc);
      __syncthreads();
      load_rows(X, slA);
      load_cols(a.s12, slC);
      __syncthreads();
      tile(slA, slC, acc);
      __syncthreads();
      load_cols(a.st1, slC);
      __syncthreads();
      tile(slA, slC, acc);
      const double v = s22v + reduce(acc);
      double lmax = 0.0;
      if (mine) {
         Xn[e] = v;
         lmax = fabs(v - slB[cl][r ^ cl]);   /* X(r, c) */
      }
      red[tid] = lmax;
      __syncthreads();
      for (int s2 = 128; s2 > 0; s2 >>= 1) {
         if (tid < s2) red[tid] = fmax(red[tid], red[tid + s2]);
         __syncthreads();
      }
      if (tid == 0) a.wgmax[blockIdx.x] = red[0];
      lag_barrier(a.bar, ++nbar, nwg, a.err);
      gmax = grid_max(a.wgmax, nwg, red);
      if (gmax < 1e-15) break;
      if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
   }
   const double *Xf = a.X + (size_t)(it & 1) * nn;
   if (mine) a.lmd[e] = Xf[e];
   if (blockIdx.x == 0 && tid == 0) {
      a.info[0] = (double)it;
      a.info[1] = gmax;
   }
}

/* ------------------------------------------------------ energy assembly */

/* one workgroup of 256 threads: combine the partial sums into E[0..9],
 * deltae, deltac (ccpsdv3/cpsd.c:515-660); every sum is a fixed strided
 * loop followed by a fixed tree, so results do not depend on timing */
__global__ void __launch_bounds__(256) k_energy(EnergyArgs a)
{
   __shared__ double red[256];
   const int t = threadIdx.x;
   double v;
   v = 0.0;
   for (int i = t; i < a.ne; i += 256) v += a.hml[i + (size_t)i * a.ne];
   const double eorbit = 2.0 * block_sum<256>(v, red);
   double sg = 0.0, sh = 0.0, sl = 0.0;
   for (int b = t; b < a.ngsum_parts; b += 256) {
      sg += a.gsum_part[3 * b]; sh += a.gsum_part[3 * b + 1]; sl += a.gsum_part[3 * b + 2];
   }
   sg = block_sum<256>(sg, red);
   sh = block_sum<256>(sh, red);
   sl = block_sum<256>(sl, red);
   v = 0.0;
   for (size_t q = t; q < (size_t)a.nion * a.nion; q += 256) v += a.ewald_part[q];
   const double ereal = block_sum<256>(v, red);
   const double eion = sg * 0.5 * a.scal2 + a.cewald + ereal;
   const double ehartr = sh * a.omega * 0.5;
   double exc = 0.0, pxc = 0.0;
   for (int b = t; b < a.nxc_parts; b += 256) { exc += a.xc_part[2 * b]; pxc += a.xc_part[2 * b + 1]; }
   exc = block_sum<256>(exc, red);
   pxc = block_sum<256>(pxc, red);
   exc += exc; pxc += pxc;
   exc *= a.dv; pxc *= a.dv;
   double eke = 0.0, dcmax = 0.0;
   for (int i = t; i < a.ne; i += 256) {
      double ek = 0.0, dc = 0.0;
      for (int b = 0; b < a.norb_parts; ++b) {
         ek += a.orb_part[((size_t)i * a.norb_parts + b) * 2];
         dc += a.orb_part[((size_t)i * a.norb_parts + b) * 2 + 1];
      }
      eke += ek;
      dcmax = fmax(dcmax, dc);
   }
   eke = 2.0 * block_sum<256>(eke, red);
   red[t] = dcmax;
   __syncthreads();
   for (int s = 128; s > 0; s >>= 1) {
      if (t < s) red[t] = fmax(red[t], red[t + s]);
      __syncthreads();
   }
   dcmax = red[0];
   __syncthreads();
   const double elocal = sl;
   v = 0.0;
   for (size_t q = t; q < (size_t)a.nproj * a.ne; q += 256) {
      const double p = a.P[q];
      v += p * p * a.scal2 / a.pnrm[q / a.ne];
   }
   const double enl = 2.0 * block_sum<256>(v, red);
   if (t != 0) return;
   const double eold = a.E[0];
   double *E = a.E;
   E[0] = eorbit + eion + exc - ehartr - pxc;
   E[1] = eorbit;
   E[2] = ehartr;
   E[3] = exc;
   E[4] = eion;
   E[5] = eke;
   E[6] = elocal;
   E[7] = enl;
   E[8] = 2.0 * ehartr;
   E[9] = pxc;
   E[10] = (E[0] - eold) / (a.dt * a.inner);   /* deltae */
   E[11] = dcmax / a.dte;                        /* deltac */
}

/* hml from M[i][j] = gcdotc(c1_i, cp_j) (row-major): upper triangle
 * -M[i][j], mirrored (ccpsdv3/cpsd.c:504-513) */
__global__ void k_hml(const double *__restrict__ M, int ne, double *__restrict__ hml)
{
   const int idx = blockIdx.x * blockDim.x + threadIdx.x;
   if (idx >= ne * ne) return;
   const int r = idx % ne, c = idx / ne;
   const int i = min(r, c), j = max(r, c);
   hml[idx] = -M[(size_t)i * ne + j];
}

/* ------------------------------------------------------------ launchers */

static inline dim3 g1(size_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

int sd_pack(const double2 *full, size_t fstride, const int *gidx, int ng, int ngp, double2 *packed,
            size_t pstride, int ne, hipStream_t s)
{
   hipLaunchKernelGGL(k_pack, dim3((ngp + 255) / 256, ne), dim3(256), 0, s, full, fstride, gidx, ng, ngp,
                      packed, pstride, ne);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_check_unmasked(const double2 *full, size_t fstride, const int *pidx, int nfft3d, int ne,
                      int *flag, hipStream_t s)
{
   hipLaunchKernelGGL(k_check_unmasked, dim3((nfft3d + 255) / 256, ne), dim3(256), 0, s, full, fstride,
                      pidx, nfft3d, ne, flag);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_unpack(const double2 *packed, size_t pstride, const int *pidx, int nfft3d, double2 *full,
              size_t fstride, int ne, hipStream_t s)
{
   hipLaunchKernelGGL(k_unpack, dim3((nfft3d + 255) / 256, ne), dim3(256), 0, s, packed, pstride, pidx,
                      nfft3d, full, fstride, ne);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_pad_real(const double *dense, int n, double *padded, hipStream_t s)
{
   const size_t tot = (size_t)n * n * (n + 2);
   hipLaunchKernelGGL(k_pad_real, g1(tot), dim3(256), 0, s, dense, n, padded);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_structure(const StructArgs &a, hipStream_t s)
{
   hipLaunchKernelGGL(k_structure, g1(a.ngp), dim3(256), 0, s, a);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_dn_reduce_xc(const double *dnpart, int nparts, size_t n3, double scal2, double *dn, double *xcp,
                    double *xce, hipStream_t s)
{
   hipLaunchKernelGGL(k_dn_reduce_xc, g1(n3), dim3(256), 0, s, dnpart, nparts, n3, scal2, dn, xcp, xce);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_lda(const double *dn, size_t n, double *xcp, double *xce, hipStream_t s)
{
   hipLaunchKernelGGL(k_lda, g1(n), dim3(256), 0, s, dn, n, xcp, xce);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_gsums(const double2 *sg, const double2 *dng, const double2 *vall, const double *vg, const double *vc,
             const double *w, int ngp, double *part, int nparts, hipStream_t s)
{
   hipLaunchKernelGGL(k_gsums, dim3(nparts), dim3(256), 0, s, sg, dng, vall, vg, vc, w, ngp, part);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_xcsums(const double *dn, const double *xce, const double *xcp, size_t n3, double *part, int nparts,
              hipStream_t s)
{
   hipLaunchKernelGGL(k_xcsums, dim3(nparts), dim3(256), 0, s, dn, xce, xcp, n3, part);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_orbsums(const double2 *c1, const double2 *c2, double2 *cp, const double *tg, const double *w, int ngp,
               int ne, double *part, int nparts, hipStream_t s)
{
   hipLaunchKernelGGL(k_orbsums, dim3(nparts, ne), dim3(256), 0, s, c1, c2, cp, tg, w, ngp, part);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_ewald_real(const double *r1, const int *katm, const double *zv, int nion, const double *rcell,
                  int nsh, double rcut, double *part, hipStream_t s)
{
   hipLaunchKernelGGL(k_ewald_real, g1((size_t)nion * nion), dim3(256), 0, s, r1, katm, zv, nion, rcell, nsh, rcut, part);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_erfc_as(const double *w, int n, double *out, hipStream_t s)
{
   hipLaunchKernelGGL(k_erfc_as, g1(n), dim3(256), 0, s, w, n, out);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_smat(const double *Gm, const double *g11, int ne, int c1_second, double dte, double *s11, double *s12,
            double *s21, double *s22, unsigned *bar, hipStream_t s)
{
   hipLaunchKernelGGL(k_smat, g1((size_t)ne * ne), dim3(256), 0, s, Gm, g11, ne, c1_second, dte, s11, s12, s21,
                      s22, bar);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_smat_precond(const double *Gm, const double *g11, int ne, int c1_second, double dte, double *DP, double *s11,
                    double *s12, double *s21, double *s22, unsigned *bar, hipStream_t s)
{
   const size_t ee = (size_t)ne * ne;
   hipLaunchKernelGGL(k_precond_dpq, g1(ee), dim3(256), 0, s, Gm, g11, ne, c1_second, DP, DP + ee, DP + 2 * ee);
   CPSD_CHECK(hipGetLastError());
   hipLaunchKernelGGL(k_precond_smat, g1(ee), dim3(256), 0, s, Gm, g11, ne, c1_second, dte, DP + ee, DP + 2 * ee,
                      s11, s12, s21, s22, bar);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_precond_lmd(double *lmd, const double *DP, int ne, double dte, hipStream_t s)
{
   const size_t ee = (size_t)ne * ne;
   hipLaunchKernelGGL(k_precond_lmd, g1(ee), dim3(256), 0, s, lmd, DP, (int)ee, dte);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_gram_next(const double *Gm, const double *g11, int ne, int c1_second, const double *lmd, double dte,
                 double *work, double *out, hipStream_t s)
{
   double *W = work, *T1 = work + (size_t)ne * ne;
   if (ne % 4) { cpsd_set_error("gram_next: ne=%d is not a multiple of 4", ne); return -1; }
   const dim3 grid((unsigned)(((size_t)ne * ne + 63) / 64));
   if (ne <= 128) {
      const int nt = (ne + 15) / 16;
      hipLaunchKernelGGL(k_gram_next_w_mf<128>, dim3(nt * nt), dim3(256), 0, s, Gm, g11, ne, c1_second, lmd, W,
                         T1);
   } else if (ne <= 256) {
      const int nt = (ne + 15) / 16;
      hipLaunchKernelGGL(k_gram_next_w_mf<256>, dim3(nt * nt), dim3(256), 0, s, Gm, g11, ne, c1_second, lmd, W,
                         T1);
   } else {
      hipLaunchKernelGGL(k_gram_next_w, grid, dim3(256), 0, s, Gm, g11, ne, c1_second, lmd, W, T1);
   }
   CPSD_CHECK(hipGetLastError());
   hipLaunchKernelGGL(k_gram_next, grid, dim3(256), 0, s, Gm, ne, c1_second, lmd, dte, W, T1, out);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_kbcoef(const double *P, int nproj, int ne, const double *pnrm, double scal2, double *X, hipStream_t s)
{
   hipLaunchKernelGGL(k_kbcoef, g1((size_t)nproj * ne), dim3(256), 0, s, P, nproj, ne, pnrm, scal2, X);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* The persistent Lagrange kernels meet at grid barriers, so every workgroup
 * of the grid must be resident at once: the grid may not exceed what the
 * device holds of the kernel (occupancy per CU times CUs).  Contexts that
 * run concurrently share the CUs, which is why the barrier wait is bounded
 * (err) as well. */
static int lagrange_grid_fits(const void *k, int grid, int lds)
{
   int dev = 0, cus = 0, per = 0;
   CPSD_CHECK(hipGetDevice(&dev));
   CPSD_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
   CPSD_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, 256, lds));
   if ((long)per * cus < grid) {
      cpsd_set_error("lagrange: %d workgroups must be resident at once but the device holds %d x %d", grid, per,
                     cus);
      return -1;
   }
   return 0;
}

int sd_lagrange(const LagArgs &a, hipStream_t s)
{
   static std::atomic<int> checked[3] = {0, 0, 0};   /* largest grid shown to fit, per kernel */
   if (a.ne <= 128) {
      const int nt = (a.ne + 15) / 16;
      const int lds = 6 * 16 * (128 + 2) * (int)sizeof(double);
      static std::atomic<unsigned long long> attr{0};
      if (dyn_lds_once((const void *)k_lagrange_res<128>, lds, attr) != 0) return -1;
      if (nt * nt > checked[0].load()) {
         if (lagrange_grid_fits((const void *)k_lagrange_res<128>, nt * nt, lds) != 0) return -1;
         checked[0].store(nt * nt);
      }
      hipLaunchKernelGGL((k_lagrange_res<128>), dim3(nt * nt), dim3(256), lds, s, a);
   } else if (a.ne <= 256) {
      const int nt = (a.ne + 15) / 16;
      const int lds = 3 * 16 * 256 * (int)sizeof(double);
      static std::atomic<unsigned long long> attr{0};
      if (dyn_lds_once((const void *)k_lagrange_str<256>, lds, attr) != 0) return -1;
      if (nt * nt > checked[1].load()) {
         if (lagrange_grid_fits((const void *)k_lagrange_str<256>, nt * nt, lds) != 0) return -1;
         checked[1].store(nt * nt);
      }
      hipLaunchKernelGGL((k_lagrange_str<256>), dim3(nt * nt), dim3(256), lds, s, a);
   } else if (a.ne <= 512) {
      const int nt = (a.ne + 31) / 32;
      if (nt * nt > checked[2].load()) {
         if (lagrange_grid_fits((const void *)k_lagrange<32>, nt * nt, 0) != 0) return -1;
         checked[2].store(nt * nt);
      }
      hipLaunchKernelGGL(k_lagrange<32>, dim3(nt * nt), dim3(256), 0, s, a);
   } else {
      cpsd_set_error("lagrange: ne=%d above 512 is not supported", a.ne);
      return -1;
   }
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_energy(const EnergyArgs &a, hipStream_t s)
{
   hipLaunchKernelGGL(k_energy, dim3(1), dim3(256), 0, s, a);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int sd_hml(const double *M, int ne, double *hml, hipStream_t s)
{
   hipLaunchKernelGGL(k_hml, g1((size_t)ne * ne), dim3(256), 0, s, M, ne, hml);
   CPSD_CHECK(hipGetLastError());
   return 0;
}
