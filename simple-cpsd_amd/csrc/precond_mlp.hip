/*
 * precond_mlp.hip -- the learned per-G preconditioner: a small MLP that maps
 * features of each plane-wave coefficient G to the factor K(G) of the
 * preconditioned SD step (EPI_PRECOND), with its training kernels.
 *
 * Network (FP32 weights, FP32 arithmetic): x[NF] -> tanh(x W1 + b1)[NH] ->
 * tanh(h1 W2 + b2)[NH] -> y = h2 . w3 + b3, K = exp(y).
 * Parameter vector (MLP_NP floats): W1[NF][NH], b1[NH], W2[NH][NH] (input
 * row, output column), b2[NH], w3[NH], b3.
 *
 * Training target.  For states c1 along SD runs of a cell, the residual is
 * R = cpsi + c1 hml (= -(H c1 - c1 <c1|H|c1>)) and the remaining distance to
 * the converged orbitals is D = c_inf - c1.  Near the minimum D ~ J^-1 R for
 * the Hessian J, so the best diagonal step factor at G is the least-squares
 * fit of K R ~ D over the states and orbitals, which only needs the per-G
 * sums A(G) = sum |dte R|^2 and B(G) = sum Re(conj(dte R) D) (factors in
 * the units of the step c1 + K dte R).  The loss of a parameter set is
 *     L = sum_G A(G) K(G)^2 - 2 B(G) K(G),
 * minimised per G by K = B/A; the MLP ties the factors of all G together
 * through their features, so it generalises across cells.
 *
 * Every reduction is in a fixed order (block partials, then one ordered sum
 * per parameter), so training is bit-reproducible on one device.
 */
#include "cpsd_internal.h"

#include <math.h>

#define MLP_NF 4
#define MLP_NH 32
#define MLP_NP (MLP_NF * MLP_NH + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NH + 1)
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + MLP_NF * MLP_NH)
#define OFF_W2 (OFF_B1 + MLP_NH)
#define OFF_B2 (OFF_W2 + MLP_NH * MLP_NH)
#define OFF_W3 (OFF_B2 + MLP_NH)
#define OFF_B3 (OFF_W3 + MLP_NH)
static_assert(MLP_NP == CPSD_MLP_NPARAM, "parameter count");

#define TB 128          /* samples per chunk = threads per workgroup */
#define YMAX 12.0f      /* exp(y) is evaluated at min(y, YMAX) */

__device__ __forceinline__ void mlp_forward(const float *P, const float *x, float *h1, float *h2, float &y)
{
#pragma unroll
   for (int j = 0; j < MLP_NH; ++j) {
      float a = P[OFF_B1 + j];
#pragma unroll
      for (int f = 0; f < MLP_NF; ++f) a = fmaf(x[f], P[OFF_W1 + f * MLP_NH + j], a);
      h1[j] = tanhf(a);
   }
#pragma unroll
   for (int j = 0; j < MLP_NH; ++j) {
      float a = P[OFF_B2 + j];
#pragma unroll
      for (int i = 0; i < MLP_NH; ++i) a = fmaf(h1[i], P[OFF_W2 + i * MLP_NH + j], a);
      h2[j] = tanhf(a);
   }
   float o = P[OFF_B3];
#pragma unroll
   for (int j = 0; j < MLP_NH; ++j) o = fmaf(h2[j], P[OFF_W3 + j], o);
   y = o;
}

/* One pass over the samples: the loss and its gradient.  Workgroup w takes
 * chunks w, w + grid, ... of TB samples; thread t owns the gradient entries
 * q = t, t + TB, ... and sums them over the workgroup's chunks in order;
 * gpart[w][q] and lpart[w] are the workgroup's partial sums. */
__global__ void __launch_bounds__(TB) k_mlp_loss_grad(const float *X, const float *A, const float *B, long n,
                                                      const float *Pg, float *gpart, float *lpart)
{
   __shared__ float P[MLP_NP];
   __shared__ float sx[TB][MLP_NF];
   __shared__ float sh1[TB][MLP_NH + 1], sh2[TB][MLP_NH + 1];
   __shared__ float sz1[TB][MLP_NH + 1], sz2[TB][MLP_NH + 1];
   __shared__ float sdy[TB];
   __shared__ float sl[TB];
   const int t = threadIdx.x;
   for (int q = t; q < MLP_NP; q += TB) P[q] = Pg[q];
   constexpr int QPT = (MLP_NP + TB - 1) / TB;
   float g[QPT];
#pragma unroll
   for (int u = 0; u < QPT; ++u) g[u] = 0.f;
   float lsum = 0.f;
   __syncthreads();
   const long nchunk = (n + TB - 1) / TB;
   for (long ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
      const long i = ch * TB + t;
      float x[MLP_NF], h1[MLP_NH], h2[MLP_NH], y = 0.f;
      float a = 0.f, b = 0.f;
      if (i < n) {
#pragma unroll
         for (int f = 0; f < MLP_NF; ++f) x[f] = X[i * MLP_NF + f];
         a = A[i];
         b = B[i];
      } else {
#pragma unroll
         for (int f = 0; f < MLP_NF; ++f) x[f] = 0.f;
      }
      mlp_forward(P, x, h1, h2, y);
      const float K = expf(fminf(y, YMAX));
      /* dL/dy = 2 K (A K - B); zero above YMAX, where K is held constant */
      float dy = (y < YMAX) ? 2.f * K * (a * K - b) : 0.f;
      float l = a * K * K - 2.f * b * K;
      if (i >= n) { dy = 0.f; l = 0.f; }
      lsum += l;
      float z2[MLP_NH];
#pragma unroll
      for (int j = 0; j < MLP_NH; ++j) z2[j] = dy * P[OFF_W3 + j] * (1.f - h2[j] * h2[j]);
#pragma unroll
      for (int i1 = 0; i1 < MLP_NH; ++i1) {
         float s = 0.f;
#pragma unroll
         for (int j = 0; j < MLP_NH; ++j) s = fmaf(z2[j], P[OFF_W2 + i1 * MLP_NH + j], s);
         sz1[t][i1] = s * (1.f - h1[i1] * h1[i1]);
         sh1[t][i1] = h1[i1];
         sh2[t][i1] = h2[i1];
         sz2[t][i1] = z2[i1];
      }
#pragma unroll
      for (int f = 0; f < MLP_NF; ++f) sx[t][f] = x[f];
      sdy[t] = dy;
      __syncthreads();
      /* this chunk's contribution to the entries the thread owns */
#pragma unroll
      for (int u = 0; u < QPT; ++u) {
         const int q = t + TB * u;
         if (q >= MLP_NP) break;
         float s = 0.f;
         if (q < OFF_B1) {
            const int f = q / MLP_NH, j = q % MLP_NH;
            for (int r = 0; r < TB; ++r) s = fmaf(sx[r][f], sz1[r][j], s);
         } else if (q < OFF_W2) {
            const int j = q - OFF_B1;
            for (int r = 0; r < TB; ++r) s += sz1[r][j];
         } else if (q < OFF_B2) {
            const int i1 = (q - OFF_W2) / MLP_NH, j = (q - OFF_W2) % MLP_NH;
            for (int r = 0; r < TB; ++r) s = fmaf(sh1[r][i1], sz2[r][j], s);
         } else if (q < OFF_W3) {
            const int j = q - OFF_B2;
            for (int r = 0; r < TB; ++r) s += sz2[r][j];
         } else if (q < OFF_B3) {
            const int j = q - OFF_W3;
            for (int r = 0; r < TB; ++r) s = fmaf(sh2[r][j], sdy[r], s);
         } else {
            for (int r = 0; r < TB; ++r) s += sdy[r];
         }
         g[u] += s;
      }
      __syncthreads();
   }
#pragma unroll
   for (int u = 0; u < QPT; ++u) {
      const int q = t + TB * u;
      if (q < MLP_NP) gpart[(size_t)blockIdx.x * MLP_NP + q] = g[u];
   }
   /* workgroup loss: fixed tree over the threads */
   sl[t] = lsum;
   __syncthreads();
   for (int s = TB / 2; s > 0; s >>= 1) {
      if (t < s) sl[t] += sl[t + s];
      __syncthreads();
   }
   if (t == 0) lpart[blockIdx.x] = sl[0];
}

/* grad[q] = sum over workgroups w (in order) of gpart[w][q]; loss likewise */
__global__ void k_mlp_reduce(const float *gpart, const float *lpart, int nblk, float *grad, double *loss)
{
   const int q = blockIdx.x * blockDim.x + threadIdx.x;
   if (q < MLP_NP) {
      float s = 0.f;
      for (int w = 0; w < nblk; ++w) s += gpart[(size_t)w * MLP_NP + q];
      grad[q] = s;
   }
   if (q == MLP_NP) {
      double s = 0.0;
      for (int w = 0; w < nblk; ++w) s += (double)lpart[w];
      *loss = s;
   }
}

/* Adam (Kingma & Ba 2015) on the whole parameter vector, step t >= 1 */
__global__ void k_mlp_adam(float *P, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                           int t)
{
   const int q = blockIdx.x * blockDim.x + threadIdx.x;
   if (q >= MLP_NP) return;
   const float gq = g[q];
   const float mq = b1 * m[q] + (1.f - b1) * gq;
   const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
   m[q] = mq;
   v[q] = vq;
   const float mh = mq / (1.f - powf(b1, (float)t));
   const float vh = vq / (1.f - powf(b2, (float)t));
   P[q] -= lr * mh / (sqrtf(vh) + eps);
}

/* y[i] of every sample; with Kout, also Kout[i] = clamp(alpha exp(y), kmin,
 * kmax) in FP64 (the factors the preconditioned step reads) */
__global__ void __launch_bounds__(TB) k_mlp_predict(const float *X, long n, const float *Pg, float *yout,
                                                    double *Kout, double alpha, double kmin, double kmax)
{
   __shared__ float P[MLP_NP];
   for (int q = threadIdx.x; q < MLP_NP; q += TB) P[q] = Pg[q];
   __syncthreads();
   const long i = (long)blockIdx.x * TB + threadIdx.x;
   if (i >= n) return;
   float x[MLP_NF], h1[MLP_NH], h2[MLP_NH], y;
#pragma unroll
   for (int f = 0; f < MLP_NF; ++f) x[f] = X[i * MLP_NF + f];
   mlp_forward(P, x, h1, h2, y);
   if (yout) yout[i] = y;
   if (Kout) {
      double k = alpha * exp((double)fminf(y, YMAX));
      k = !(k >= kmin) ? kmin : (k > kmax ? kmax : k);   /* a NaN output takes kmin */
      Kout[i] = k;
   }
}

/* features of packed coefficient p < ng (zeros in the padded tail):
 *   x0 = log(1 + T) / 5,       T = |G|^2 / 2 (Ha), the kinetic diagonal
 *   x1 = T / (1 + T)
 *   x2 = asinh(Dnl),           Dnl = -(2/omega) sum_m |beta_m(G)|^2 / vnlnrm_m,
 *                              the diagonal of the KB operator at G (Ha)
 *   x3 = asinh(10 |vall(G)| / omega), the local pseudopotential at G (Ha) */
__global__ void k_precond_features(const double *tg, const double2 *B, const double *pnrm, int nproj,
                                   const double2 *vall, double omega, int ng, int ngp, float *X)
{
   const int p = blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= ngp) return;
   float *x = X + (size_t)p * MLP_NF;
   if (p >= ng) {
      x[0] = x[1] = x[2] = x[3] = 0.f;
      return;
   }
   const double T = tg[p];
   double d = 0.0;
   for (int m = 0; m < nproj; ++m) {
      const double2 b = B[(size_t)m * ngp + p];
      d += (b.x * b.x + b.y * b.y) / pnrm[m];
   }
   d *= -2.0 / omega;
   const double2 v = vall[p];
   x[0] = (float)(log1p(T) / 5.0);
   x[1] = (float)(T / (1.0 + T));
   x[2] = (float)asinh(d);
   x[3] = (float)asinh(10.0 * sqrt(v.x * v.x + v.y * v.y) / omega);
}

/* training statistics of the current state, per packed coefficient p:
 * A[p] += sum_i |dte R_i(p)|^2, B[p] += sum_i Re(conj(dte R_i(p)) D_i(p)),
 * R = CP + Rh (Rh = C1 hml), D = Cinf - C1 */
__global__ void k_precond_stats(const double2 *CP, const double2 *Rh, const double2 *C1, const double2 *Cinf,
                                int ne, int ngp, double dte, double *A, double *B)
{
   const int p = blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= ngp) return;
   double a = 0.0, b = 0.0;
   for (int i = 0; i < ne; ++i) {
      const size_t k = (size_t)i * ngp + p;
      const double2 cp = CP[k], rh = Rh[k], c1 = C1[k], ci = Cinf[k];
      const double rx = dte * (cp.x + rh.x), ry = dte * (cp.y + rh.y);
      a += rx * rx + ry * ry;
      b += rx * (ci.x - c1.x) + ry * (ci.y - c1.y);
   }
   A[p] += a;
   B[p] += b;
}

/* ------------------------------------------------------------ launchers */

int mlp_nparam(void) { return MLP_NP; }

int launch_mlp_loss_grad(const float *X, const float *A, const float *B, long n, const float *P, float *gpart,
                         float *lpart, int nblk, float *grad, double *loss, hipStream_t s)
{
   hipLaunchKernelGGL(k_mlp_loss_grad, dim3(nblk), dim3(TB), 0, s, X, A, B, n, P, gpart, lpart);
   CPSD_CHECK(hipGetLastError());
   hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NP + 1 + 255) / 256), dim3(256), 0, s, gpart, lpart, nblk, grad,
                      loss);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int launch_mlp_adam(float *P, const float *g, float *m, float *v, float lr, int t, hipStream_t s)
{
   hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, P, g, m, v, lr, 0.9f, 0.999f,
                      1e-8f, t);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int launch_mlp_predict(const float *X, long n, const float *P, float *y, double *K, double alpha, double kmin,
                       double kmax, hipStream_t s)
{
   if (n <= 0) return 0;
   hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)((n + TB - 1) / TB)), dim3(TB), 0, s, X, n, P, y, K, alpha,
                      kmin, kmax);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int launch_precond_features(const double *tg, const double2 *B, const double *pnrm, int nproj,
                            const double2 *vall, double omega, int ng, int ngp, float *X, hipStream_t s)
{
   hipLaunchKernelGGL(k_precond_features, dim3((ngp + 255) / 256), dim3(256), 0, s, tg, B, pnrm, nproj, vall,
                      omega, ng, ngp, X);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

int launch_precond_stats(const double2 *CP, const double2 *Rh, const double2 *C1, const double2 *Cinf, int ne,
                         int ngp, double dte, double *A, double *B, hipStream_t s)
{
   hipLaunchKernelGGL(k_precond_stats, dim3((ngp + 255) / 256), dim3(256), 0, s, CP, Rh, C1, Cinf, ne, ngp,
                      dte, A, B);
   CPSD_CHECK(hipGetLastError());
   return 0;
}

/* ------------------------------------------------------- C-ABI (training) */

#include <vector>

namespace {
struct DevBuf {
   void *p = nullptr;
   ~DevBuf() { if (p) (void)hipFree(p); }
};
}

template<typename T>
static int upload(DevBuf &b, const T *h, size_t n, hipStream_t s)
{
   CPSD_CHECK(hipMalloc(&b.p, (n ? n : 1) * sizeof(T)));
   if (n) CPSD_CHECK(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
   return 0;
}

static int alloc(DevBuf &b, size_t bytes)
{
   CPSD_CHECK(hipMalloc(&b.p, bytes ? bytes : 1));
   return 0;
}

/* the training set on the device: X as given, A and B rounded to FP32 */
static int upload_set(const float *X, const double *A, const double *B, long n, DevBuf &dX, DevBuf &dA,
                      DevBuf &dB, hipStream_t s)
{
   std::vector<float> a(n), b(n);
   for (long i = 0; i < n; ++i) { a[i] = (float)A[i]; b[i] = (float)B[i]; }
   if (upload(dX, X, (size_t)n * MLP_NF, s) || upload(dA, a.data(), (size_t)n, s) ||
       upload(dB, b.data(), (size_t)n, s))
      return -1;
   CPSD_CHECK(hipStreamSynchronize(s));
   return 0;
}

static int grad_blocks(long n)
{
   const long chunks = (n + TB - 1) / TB;
   return (int)(chunks < 1024 ? (chunks > 0 ? chunks : 1) : 1024);
}

extern "C" int cpsd_mlp_nparam(void) { return MLP_NP; }

extern "C" int cpsd_mlp_loss_grad(const float *X, const double *A, const double *B, long n, const float *params,
                                  float *grad, double *loss)
{
   if (n < 1 || !X || !A || !B || !params || !grad || !loss) {
      cpsd_set_error("cpsd_mlp_loss_grad: needs n >= 1 samples and every array");
      return -1;
   }
   hipStream_t s = nullptr;
   DevBuf dX, dA, dB, dP, gp, lp, dg, dl;
   const int nblk = grad_blocks(n);
   if (upload_set(X, A, B, n, dX, dA, dB, s) || upload(dP, params, MLP_NP, s) ||
       alloc(gp, (size_t)nblk * MLP_NP * sizeof(float)) || alloc(lp, nblk * sizeof(float)) ||
       alloc(dg, MLP_NP * sizeof(float)) || alloc(dl, sizeof(double)))
      return -1;
   if (launch_mlp_loss_grad((float *)dX.p, (float *)dA.p, (float *)dB.p, n, (float *)dP.p, (float *)gp.p,
                            (float *)lp.p, nblk, (float *)dg.p, (double *)dl.p, s) != 0)
      return -1;
   CPSD_CHECK(hipMemcpy(grad, dg.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
   CPSD_CHECK(hipMemcpy(loss, dl.p, sizeof(double), hipMemcpyDeviceToHost));
   return 0;
}

extern "C" int cpsd_mlp_train(const float *X, const double *A, const double *B, long n, float *params, int iters,
                              float lr, double *loss_hist)
{
   if (n < 1 || iters < 1 || !X || !A || !B || !params || !(lr > 0.f)) {
      cpsd_set_error("cpsd_mlp_train: needs n >= 1 samples, iters >= 1, lr > 0 and every array");
      return -1;
   }
   hipStream_t s = nullptr;
   DevBuf dX, dA, dB, dP, gp, lp, dg, dm, dv, dl;
   const int nblk = grad_blocks(n);
   if (upload_set(X, A, B, n, dX, dA, dB, s) || upload(dP, params, MLP_NP, s) ||
       alloc(gp, (size_t)nblk * MLP_NP * sizeof(float)) || alloc(lp, nblk * sizeof(float)) ||
       alloc(dg, MLP_NP * sizeof(float)) || alloc(dm, MLP_NP * sizeof(float)) ||
       alloc(dv, MLP_NP * sizeof(float)) || alloc(dl, (size_t)iters * sizeof(double)))
      return -1;
   CPSD_CHECK(hipMemsetAsync(dm.p, 0, MLP_NP * sizeof(float), s));
   CPSD_CHECK(hipMemsetAsync(dv.p, 0, MLP_NP * sizeof(float), s));
   for (int it = 0; it < iters; ++it) {
      if (launch_mlp_loss_grad((float *)dX.p, (float *)dA.p, (float *)dB.p, n, (float *)dP.p, (float *)gp.p,
                               (float *)lp.p, nblk, (float *)dg.p, (double *)dl.p + it, s) != 0 ||
          launch_mlp_adam((float *)dP.p, (float *)dg.p, (float *)dm.p, (float *)dv.p, lr, it + 1, s) != 0)
         return -1;
   }
   CPSD_CHECK(hipMemcpy(params, dP.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
   if (loss_hist) CPSD_CHECK(hipMemcpy(loss_hist, dl.p, (size_t)iters * sizeof(double), hipMemcpyDeviceToHost));
   return 0;
}

extern "C" int cpsd_mlp_predict(const float *X, long n, const float *params, float *y)
{
   if (n < 1 || !X || !params || !y) { cpsd_set_error("cpsd_mlp_predict: needs n >= 1 and every array"); return -1; }
   hipStream_t s = nullptr;
   DevBuf dX, dP, dy;
   if (upload(dX, X, (size_t)n * MLP_NF, s) || upload(dP, params, MLP_NP, s) || alloc(dy, (size_t)n * sizeof(float)))
      return -1;
   if (launch_mlp_predict((float *)dX.p, n, (float *)dP.p, (float *)dy.p, nullptr, 0, 0, 0, s) != 0) return -1;
   CPSD_CHECK(hipMemcpy(y, dy.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
   return 0;
}
