// Near-pair search shared by the quadrature set-up and the sparse approximate
// inverse: a host-built uniform cell grid and device count / fill passes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace ed {

// R = sqrt(clip(A, 0, inf) / pi): a NaN area stays NaN (no partner at all)
__device__ __forceinline__ double near_radius(double a) {
#pragma clang fp contract(off)
  a = a < 0.0 ? 0.0 : a;
  return sqrt(a / 3.141592653589793);
}

// The near-pair acceptance test of core/bem.py:475-555 for panels i and j:
// 0 < |c_i - c_j| <= df (R_i + R_j), with FP contraction off so that it
// takes the same decisions as a brute-force scan in FP64.
__device__ __forceinline__ bool near_accept(double xi, double yi, double zi, double ri,
                                            double xj, double yj, double zj, double aj,
                                            double df) {
#pragma clang fp contract(off)
  const double ex = xi - xj, ey = yi - yj, ez = zi - zj;
  const double d = sqrt(ex * ex + ey * ey + ez * ez);
  const double th = df * (ri + near_radius(aj));
  return isfinite(d) && isfinite(th) && d > 0.0 && d <= th;
}

struct NearGrid {
  double lo[3];
  double inv_h;
  int32_t dim[3];
};

// Pairs (i, j), i != j, 0 < |c_i - c_j| <= df (R_i + R_j), R = sqrt(A / pi),
// rows i-major with j ascending.  C (n, 3) and A (n) are device FP64.
struct NearSearch {
  const double* C = nullptr;
  const double* A = nullptr;
  int64_t n = 0;
  double df = 0.0;
  NearGrid g{};
  int32_t* d_start = nullptr;
  int32_t* d_pts = nullptr;
  bool any = false;  // false: no pair can exist
  ~NearSearch();
  int build(const double* C, const double* A, int64_t n, double df, hipStream_t st);
  // row offsets (n + 1, host, exclusive prefix of the row counts); a
  // factor df_use <= the build's df searches the same grid (0: the build's)
  int count(hipStream_t st, std::vector<int64_t>& ptr, double df_use = 0.0);
  // pairs (int64 (i, j)) and/or int32 columns j + col_off at the row offsets
  int fill(hipStream_t st, const std::vector<int64_t>& ptr, int64_t* pairs, int32_t* cols32,
           int64_t col_off, double df_use = 0.0);
};

}  // namespace ed
