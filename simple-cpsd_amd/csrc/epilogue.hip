/*
 * epilogue.hip -- host side of the run epilogue (ccpsdv3/main.c:466-484):
 * the eigenpairs of the ne x ne matrix hml, which the reference takes from
 * LAPACK dsyev('V', 'U') (ascending eigenvalues) and then reorders with
 * eigsrt (ccpsdv3/eigsrt.c:11-39).  The rotation c1 = c2 V that follows runs
 * on the device (cpsd_api.hip, tall MFMA kernel with EPI_STORE).
 *
 * Eigensolver: the cyclic Jacobi method (Golub and Van Loan, Matrix
 * Computations, 4th ed., section 8.5: the symmetric 2 x 2 Schur step
 * "sym.schur2" and cyclic-by-row sweeps).  Each rotation J(p, q) zeroes the
 * (p, q) pair of A <- J^T A J and is accumulated into V <- V J; sweeps stop
 * when the off-diagonal Frobenius norm falls below 1e-16 of the whole.  The
 * method is accurate to a few ulps of the largest eigenvalue; ne <= 512 takes
 * at most a few seconds once per run.
 */
#include "cpsd_internal.h"
#include <cmath>
#include <vector>
#include <algorithm>

/* a: n x n row-major symmetric, overwritten; v: eigenvectors in columns */
static void jacobi_sweeps(int n, std::vector<double> &a, std::vector<double> &v)
{
   v.assign((size_t)n * n, 0.0);
   for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
   for (int sweep = 0; sweep < 100; ++sweep) {
      double off = 0.0, tot = 0.0;
      for (int i = 0; i < n; ++i)
         for (int j = 0; j < n; ++j) {
            const double x = a[(size_t)i * n + j] * a[(size_t)i * n + j];
            tot += x;
            if (i != j) off += x;
         }
      if (off == 0.0 || off <= 1e-32 * tot) break;
      for (int p = 0; p < n - 1; ++p)
         for (int q = p + 1; q < n; ++q) {
            const double apq = a[(size_t)p * n + q];
            if (apq == 0.0) continue;
            const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
            const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
            for (int k = 0; k < n; ++k) {   /* A <- A J (columns p, q) */
               const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
               a[(size_t)k * n + p] = c * akp - s * akq;
               a[(size_t)k * n + q] = s * akp + c * akq;
            }
            for (int k = 0; k < n; ++k) {   /* A <- J^T A (rows p, q) */
               const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
               a[(size_t)p * n + k] = c * apk - s * aqk;
               a[(size_t)q * n + k] = s * apk + c * aqk;
            }
            a[(size_t)p * n + q] = 0.0;
            a[(size_t)q * n + p] = 0.0;
            for (int k = 0; k < n; ++k) {   /* V <- V J */
               const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
               v[(size_t)k * n + p] = c * vkp - s * vkq;
               v[(size_t)k * n + q] = s * vkp + c * vkq;
            }
         }
   }
}

extern "C" int cpsd_syev_sorted(int n, double *a, double *eig)
{
   if (n < 1) { cpsd_set_error("syev: n=%d", n); return -1; }
   /* the upper triangle of the column-major input, as dsyev('U') reads it */
   std::vector<double> m((size_t)n * n), v;
   for (int j = 0; j < n; ++j)
      for (int i = 0; i <= j; ++i) {
         const double x = a[i + (size_t)j * n];
         if (!std::isfinite(x)) { cpsd_set_error("syev: non-finite matrix element (%d, %d)", i, j); return -1; }
         m[(size_t)i * n + j] = x;
         m[(size_t)j * n + i] = x;
      }
   jacobi_sweeps(n, m, v);
   /* dsyev's order first (ascending eigenvalues) ... */
   std::vector<int> idx(n);
   for (int i = 0; i < n; ++i) idx[i] = i;
   std::stable_sort(idx.begin(), idx.end(),
                    [&](int x, int y) { return m[(size_t)x * n + x] < m[(size_t)y * n + y]; });
   for (int i = 0; i < n; ++i) {
      eig[i] = m[(size_t)idx[i] * n + idx[i]];
      for (int j = 0; j < n; ++j) a[j + (size_t)i * n] = v[(size_t)j * n + idx[i]];
   }
   /* ... then eigsrt's selection sort (ccpsdv3/eigsrt.c:18-38): descending,
    * the last of equal maxima taken, eigenvector columns swapped along */
   for (int i = 0; i < n - 1; ++i) {
      int k = i;
      double p = eig[i];
      for (int j = i + 1; j < n; ++j)
         if (eig[j] >= p) { k = j; p = eig[j]; }
      if (k != i) {
         eig[k] = eig[i];
         eig[i] = p;
         for (int j = 0; j < n; ++j) std::swap(a[j + (size_t)i * n], a[j + (size_t)k * n]);
      }
   }
   return 0;
}
