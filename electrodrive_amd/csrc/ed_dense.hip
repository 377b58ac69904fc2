// Dense single-layer collocation matvec (core/bem_kernel.py:157-230, :715-810):
//   V_i = sum_{j != i} K_E sigma_j A_j rsqrt(max(r_ij^2, 1e-24)) + K_ii sigma_i A_i
// with K_ii = self_integrals[i] (per-area) or 0.  One thread per target row,
// source panels staged through LDS in tiles of 256; up to 4 right-hand sides
// per pass.  The FP64 kernel refines a hardware FP32 reciprocal square root
// with two FP64 Newton steps (relative error at FP64 rounding level).
#include <hip/hip_runtime.h>

#include "ed_internal.h"

namespace ed {

constexpr int kDenseBlock = 256;
constexpr int kDenseRhs = 4;

template <typename T>
__device__ __forceinline__ T inv_sqrt(T r2);

template <>
__device__ __forceinline__ double inv_sqrt<double>(double r2) {
  double y = (double)__builtin_amdgcn_rsqf((float)r2);
  y = y * (1.5 - 0.5 * r2 * y * y);
  y = y * (1.5 - 0.5 * r2 * y * y);
  return y;
}

template <>
__device__ __forceinline__ float inv_sqrt<float>(float r2) {
  return __builtin_amdgcn_rsqf(r2);
}

template <typename T>
__global__ __launch_bounds__(kDenseBlock) void k_dense(const T* __restrict__ C,
                                                       const T* __restrict__ A,
                                                       const T* __restrict__ sigma,
                                                       const T* __restrict__ selfi,
                                                       T* __restrict__ out, int64_t n,
                                                       int r0, int nr, T ke) {
  __shared__ T sx[kDenseBlock], sy[kDenseBlock], sz[kDenseBlock];
  __shared__ T sq[kDenseRhs][kDenseBlock];
  const int64_t i = (int64_t)blockIdx.x * kDenseBlock + threadIdx.x;
  const bool active = i < n;
  T xi = 0, yi = 0, zi = 0;
  if (active) { xi = C[3 * i]; yi = C[3 * i + 1]; zi = C[3 * i + 2]; }
  T acc[kDenseRhs];
#pragma unroll
  for (int r = 0; r < kDenseRhs; ++r) acc[r] = 0;
  for (int64_t j0 = 0; j0 < n; j0 += kDenseBlock) {
    const int64_t j = j0 + threadIdx.x;
    __syncthreads();
    if (j < n) {
      sx[threadIdx.x] = C[3 * j];
      sy[threadIdx.x] = C[3 * j + 1];
      sz[threadIdx.x] = C[3 * j + 2];
      const T a = A[j];
#pragma unroll
      for (int r = 0; r < kDenseRhs; ++r)
        sq[r][threadIdx.x] = (r < nr) ? sigma[(int64_t)(r0 + r) * n + j] * a : (T)0;
    } else {
      sx[threadIdx.x] = sy[threadIdx.x] = sz[threadIdx.x] = 0;
#pragma unroll
      for (int r = 0; r < kDenseRhs; ++r) sq[r][threadIdx.x] = 0;
    }
    __syncthreads();
    const int nt = (int)((n - j0) < kDenseBlock ? (n - j0) : kDenseBlock);
    const bool diag_tile = active && (i >= j0) && (i < j0 + nt);
    for (int t = 0; t < nt; ++t) {
      const T dx = xi - sx[t], dy = yi - sy[t], dz = zi - sz[t];
      T r2 = dx * dx + dy * dy + dz * dz;
      r2 = r2 > (T)1e-24 ? r2 : (T)1e-24;
      T k = ke * inv_sqrt<T>(r2);
      if (diag_tile && (j0 + t == i)) k = selfi ? selfi[i] : (T)0;
#pragma unroll
      for (int r = 0; r < kDenseRhs; ++r) acc[r] += k * sq[r][t];
    }
  }
  if (active) {
#pragma unroll
    for (int r = 0; r < kDenseRhs; ++r)
      if (r < nr) {
        T v = acc[r];
        if (!isfinite(v)) v = 0;  // core/bem_kernel.py:375-381
        out[(int64_t)(r0 + r) * n + i] = v;
      }
  }
}

template <typename T>
static int dense_launch(const T* C, const T* A, const T* sigma, const T* selfi, T* out,
                        int64_t n, int64_t nrhs, hipStream_t st) {
  if (n < 0 || nrhs < 1) return fail(ED_ERR_ARG, "dense matvec: bad sizes");
  if (n == 0) return ED_OK;
  if (!C || !A || !sigma || !out) return fail(ED_ERR_ARG, "dense matvec: NULL pointer");
  const unsigned grid = (unsigned)((n + kDenseBlock - 1) / kDenseBlock);
  for (int64_t r0 = 0; r0 < nrhs; r0 += kDenseRhs) {
    const int nr = (int)((nrhs - r0) < kDenseRhs ? (nrhs - r0) : kDenseRhs);
    hipLaunchKernelGGL(k_dense<T>, dim3(grid), dim3(kDenseBlock), 0, st, C, A, sigma, selfi,
                       out, n, (int)r0, nr, (T)kKE);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("dense matvec: ") + hipGetErrorString(e));
  }
  return ED_OK;
}

int dense_launch_f64(const double* C, const double* A, const double* sigma, const double* selfi,
                     double* out, int64_t n, int64_t nrhs, hipStream_t st) {
  return dense_launch<double>(C, A, sigma, selfi, out, n, nrhs, st);
}

}  // namespace ed

extern "C" int ed_dense_matvec_f64(const double* C, const double* A, const double* sigma,
                                   const double* selfi, double* out, int64_t n, int64_t nrhs,
                                   ed_stream stream) {
  return ed::dense_launch<double>(C, A, sigma, selfi, out, n, nrhs, (hipStream_t)stream);
}

extern "C" int ed_dense_matvec_f32(const float* C, const float* A, const float* sigma,
                                   const float* selfi, float* out, int64_t n, int64_t nrhs,
                                   ed_stream stream) {
  return ed::dense_launch<float>(C, A, sigma, selfi, out, n, nrhs, (hipStream_t)stream);
}
