/* CPU oracle / baseline: dense single-layer collocation matvec in FP64.
 *
 * TEST INFRASTRUCTURE ONLY (timed as the CPU baseline, used as a checker).
 * Restates _bem_matvec_core_torch (core/bem_kernel.py:157-230):
 *   V_i = sum_{j != i} K_E sigma_j A_j / sqrt(max(r_ij^2, 1e-24))
 *         + self_integrals[i] sigma_i A_i      (0 when self_integrals NULL)
 * for target rows [row0, row1), OpenMP over rows.  K_E is the reference
 * double 1/(4 pi 8.854187817e-12) (electrodrive/utils/config.py:34-35).
 */
#include <math.h>
#include <omp.h>

static const double kKE = 0x1.0bd99415ffbb9p+33;

int oracle_dense_rows(const double* C, const double* A, const double* sigma,
                      const double* selfi, double* out, long n, long row0, long row1,
                      int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (long i = row0; i < row1; ++i) {
    const double xi = C[3 * i], yi = C[3 * i + 1], zi = C[3 * i + 2];
    double acc = 0.0;
    for (long j = 0; j < n; ++j) {
      double k;
      if (j == i) {
        k = selfi ? selfi[i] : 0.0;
      } else {
        const double dx = xi - C[3 * j], dy = yi - C[3 * j + 1], dz = zi - C[3 * j + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < 1e-24) r2 = 1e-24;
        k = kKE / sqrt(r2);
      }
      acc += k * (sigma[j] * A[j]);
    }
    out[i - row0] = isfinite(acc) ? acc : 0.0;
  }
  return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
