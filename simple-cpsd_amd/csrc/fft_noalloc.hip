/*
 * fft_noalloc.hip -- the batched 3D FFT of cpsd_fft3d_batch() on caller-owned
 * device memory, with no allocation of any kind.
 *
 * cpsd_fft3d_batch() (and the Fortran-ABI cr_fft3b_/rc_fft3f_ built on it)
 * stages the caller's host arrays through device buffers that the library
 * allocates.  Callers that may not allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc) call
 * cpsd_fft3d_batch_noalloc() instead: the grid stays in their own device
 * buffer and is transformed in place, and the two things the transform needs
 * besides it -- the z-transformed intermediate T and the row/twiddle tables
 * of the kernels -- live in `scratch`, a file-scope device array of
 * CPSD_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends.
 *
 * The kernels are the ones cpsd_fft3d_batch() launches (z pass and fused
 * (y,x) plane pass); only where their buffers come from differs.
 */
#include "cpsd_internal.h"
#include "../../include/cpsd_abi.h"

#include <math.h>
#include <string.h>
#include <mutex>

static __device__ __align__(256) unsigned char scratch[CPSD_NOALLOC_SCRATCH_BYTES];

/* device address of `scratch` on each device ordinal, looked up once per
 * device with that device current (the array's name is a host-side
 * stand-in and must never be passed to a kernel) */
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_addr[64];

static int scratch_on_current_device(unsigned char **out)
{
   int dev = 0;
   CPSD_CHECK(hipGetDevice(&dev));
   if (dev < 0 || dev >= 64) {
      cpsd_set_error("cpsd_fft3d_batch_noalloc: device ordinal %d out of range", dev);
      return -1;
   }
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   if (!g_scratch_addr[dev]) {
      void *p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
         cpsd_set_error("cpsd_fft3d_batch_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                        hipGetErrorString(e));
         return -1;
      }
      g_scratch_addr[dev] = static_cast<unsigned char *>(p);
   }
   *out = g_scratch_addr[dev];
   return 0;
}

/* byte offsets of the pieces of the scratch array for grid n */
struct NoallocLayout {
   size_t frs, frl, fro, tw, T, end;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static NoallocLayout noalloc_layout(int n, int nbatch)
{
   const size_t rows = (size_t)n * n * sizeof(int);
   NoallocLayout L;
   L.frs = 0;
   L.frl = align256(L.frs + rows);
   L.fro = align256(L.frl + rows);
   L.tw = align256(L.fro + rows);
   L.T = align256(L.tw + (size_t)n * sizeof(double2));
   L.end = L.T + (size_t)nbatch * n * n * (n / 2 + 1) * sizeof(double2);
   return L;
}

/* the full-grid row table (every (ky,kz) row holds kx = 0..N/2) and the
 * twiddles exp(+2 pi i m/N), written on the device on every call, so a call
 * never depends on what an earlier call on another grid left there */
__global__ void k_noalloc_tables(int n, int *frs, int *frl, int *fro, double2 *tw)
{
   const int nh = n / 2 + 1;
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r < n * n) {
      frs[r] = 0;
      frl[r] = nh;
      fro[r] = r * nh;
   }
   if (r < n) {
      double s, c;
      sincospi(2.0 * (double)r / (double)n, &s, &c);
      tw[r] = make_double2(c, s);
   }
}

extern "C" long cpsd_fft3d_batch_noalloc_max(int nfft)
{
   if (!fft_supported(nfft)) return 0;
   const NoallocLayout one = noalloc_layout(nfft, 1), none = noalloc_layout(nfft, 0);
   const size_t per = one.end - none.end;
   if (none.end >= (size_t)CPSD_NOALLOC_SCRATCH_BYTES) return 0;
   return (long)(((size_t)CPSD_NOALLOC_SCRATCH_BYTES - none.end) / per);
}

extern "C" int cpsd_fft3d_batch_noalloc(int dir, int nfft, int nbatch, double *A, void *stream)
{
   if (dir != 1 && dir != -1) {
      cpsd_set_error("cpsd_fft3d_batch_noalloc: dir must be +1 (cr) or -1 (rc), got %d", dir);
      return -1;
   }
   if (!fft_supported(nfft)) {
      cpsd_set_error("cpsd_fft3d_batch_noalloc: FFT grid %d is not supported", nfft);
      return -1;
   }
   if (nbatch < 1 || !A) {
      cpsd_set_error("cpsd_fft3d_batch_noalloc: needs nbatch >= 1 and a device array");
      return -1;
   }
   const long maxb = cpsd_fft3d_batch_noalloc_max(nfft);
   if (nbatch > maxb) {
      cpsd_set_error("cpsd_fft3d_batch_noalloc: %d transforms of %d^3 need %zu bytes of scratch, "
                     "the library holds %u (at most %ld transforms of this grid per call)",
                     nbatch, nfft, noalloc_layout(nfft, nbatch).end, (unsigned)CPSD_NOALLOC_SCRATCH_BYTES, maxb);
      return -1;
   }
   unsigned char *base = nullptr;
   if (scratch_on_current_device(&base) != 0) return -1;
   const NoallocLayout L = noalloc_layout(nfft, nbatch);
   int *frs = reinterpret_cast<int *>(base + L.frs);
   int *frl = reinterpret_cast<int *>(base + L.frl);
   int *fro = reinterpret_cast<int *>(base + L.fro);
   double2 *tw = reinterpret_cast<double2 *>(base + L.tw);
   double2 *T = reinterpret_cast<double2 *>(base + L.T);
   hipStream_t st = static_cast<hipStream_t>(stream);

   const int n = nfft, nh = n / 2 + 1;
   hipLaunchKernelGGL(k_noalloc_tables, dim3((n * n + 255) / 256), dim3(256), 0, st, n, frs, frl, fro, tw);
   CPSD_CHECK(hipGetLastError());

   const size_t per = (size_t)n * n * nh;  /* complex values per transform */
   double2 *Ac = reinterpret_cast<double2 *>(A);
   ZArgs z;
   memset(&z, 0, sizeof z);
   z.rows.rstart = frs; z.rows.rlen = frl; z.rows.roff = fro; z.rows.nhc = nh;
   z.tw = tw;
   PlaneArgs pa;
   memset(&pa, 0, sizeof pa);
   pa.T = T; pa.nhload = nh; pa.tw = tw; pa.nbatch = nbatch;
   if (dir > 0) {
      z.src = Ac; z.sstride = per; z.T = T;
      if (launch_zpass_g2r(n, SRC_PACKED, z, nbatch, st) != 0) return -1;
      pa.rout = A;
      if (launch_plane(n, PM_C2R_OUT, pa, nbatch, st) != 0) return -1;
   } else {
      pa.rin = A;
      if (launch_plane(n, PM_R2C_IN, pa, nbatch, st) != 0) return -1;
      z.T = T; z.out = Ac; z.ostride = per;
      if (launch_zpass_r2g(n, DST_FULL, z, nbatch, st) != 0) return -1;
   }
   return 0;
}
