// Dense single-layer collocation matvec (core/bem_kernel.py:157-230, :715-810):
//   V_i = sum_{j != i} K_E sigma_j A_j rsqrt(max(r_ij^2, 1e-24)) + K_ii sigma_i A_i
// with K_ii = self_integrals[i] (per-area) or 0.  Two target rows per thread,
// source panels staged through LDS in tiles of 256; up to 4 right-hand sides
// per pass.  The FP64 kernel refines a hardware FP32 reciprocal square root
// with one FP64 Newton step (relative error ~1e-14).
#include <hip/hip_runtime.h>

#include "ed_internal.h"

namespace ed {

constexpr int kDenseBlock = 256;
constexpr int kDenseRhs = 4;

template <typename T>
__device__ __forceinline__ T inv_sqrt(T r2);

// FP32 hardware reciprocal square root (relative error ~1e-7) refined by one
// FP64 Newton step: relative error ~1e-14, below what a 1e-8 residual needs.
template <>
__device__ __forceinline__ double inv_sqrt<double>(double r2) {
  double y = (double)__builtin_amdgcn_rsqf((float)r2);
  return y * (1.5 - 0.5 * r2 * y * y);
}

template <>
__device__ __forceinline__ float inv_sqrt<float>(float r2) {
  return __builtin_amdgcn_rsqf(r2);
}

// Two target rows per thread (i and i + kDenseBlock); sources staged through
// LDS; the diagonal is excluded with an index test only in the tile that
// holds it, and added back with the self term after the loop.
constexpr int kDenseRows = 2;

template <typename T, int NR>
__device__ __forceinline__ void dense_tile(const T (&xi)[kDenseRows], const T (&yi)[kDenseRows],
                                           const T (&zi)[kDenseRows], const T* sx, const T* sy,
                                           const T* sz, const T (*sq)[kDenseBlock], int nt,
                                           int skip0, int skip1, T (&acc)[kDenseRows][kDenseRhs]) {
  for (int t = 0; t < nt; ++t) {
    const T px = sx[t], py = sy[t], pz = sz[t];
#pragma unroll
    for (int u = 0; u < kDenseRows; ++u) {
      const T dx = xi[u] - px, dy = yi[u] - py, dz = zi[u] - pz;
      T r2 = dx * dx + dy * dy + dz * dz;
      r2 = r2 > (T)1e-24 ? r2 : (T)1e-24;
      T k = inv_sqrt<T>(r2);
      if (u == 0 && t == skip0) k = 0;
      if (u == 1 && t == skip1) k = 0;
#pragma unroll
      for (int r = 0; r < NR; ++r) acc[u][r] += k * sq[r][t];
    }
  }
}

template <typename T, int NR>
__global__ __launch_bounds__(kDenseBlock) void k_dense(const T* __restrict__ C,
                                                       const T* __restrict__ A,
                                                       const T* __restrict__ sigma,
                                                       const T* __restrict__ selfi,
                                                       T* __restrict__ out, int64_t n,
                                                       int r0, T ke) {
  __shared__ T sx[kDenseBlock], sy[kDenseBlock], sz[kDenseBlock];
  __shared__ T sq[kDenseRhs][kDenseBlock];
  const int64_t base = (int64_t)blockIdx.x * (kDenseBlock * kDenseRows);
  int64_t ii[kDenseRows];
  T xi[kDenseRows], yi[kDenseRows], zi[kDenseRows];
#pragma unroll
  for (int u = 0; u < kDenseRows; ++u) {
    ii[u] = base + u * kDenseBlock + threadIdx.x;
    const int64_t c = ii[u] < n ? ii[u] : n - 1;
    xi[u] = C[3 * c]; yi[u] = C[3 * c + 1]; zi[u] = C[3 * c + 2];
  }
  T acc[kDenseRows][kDenseRhs];
#pragma unroll
  for (int u = 0; u < kDenseRows; ++u)
#pragma unroll
    for (int r = 0; r < kDenseRhs; ++r) acc[u][r] = 0;
  for (int64_t j0 = 0; j0 < n; j0 += kDenseBlock) {
    const int64_t j = j0 + threadIdx.x;
    __syncthreads();
    if (j < n) {
      sx[threadIdx.x] = C[3 * j];
      sy[threadIdx.x] = C[3 * j + 1];
      sz[threadIdx.x] = C[3 * j + 2];
      const T a = A[j];
#pragma unroll
      for (int r = 0; r < NR; ++r) sq[r][threadIdx.x] = sigma[(int64_t)(r0 + r) * n + j] * a;
    } else {
      // far-away dummy source with zero charge
      sx[threadIdx.x] = sy[threadIdx.x] = sz[threadIdx.x] = (T)1e30;
#pragma unroll
      for (int r = 0; r < NR; ++r) sq[r][threadIdx.x] = 0;
    }
    __syncthreads();
    const int nt = (int)((n - j0) < kDenseBlock ? (n - j0) : kDenseBlock);
    // tiles that hold a diagonal entry of this block are uniform per block
    const bool d0 = (j0 == base), d1 = (j0 == base + kDenseBlock);
    if (d0 || d1) {
      const int s0 = d0 ? (int)threadIdx.x : -1, s1 = d1 ? (int)threadIdx.x : -1;
      dense_tile<T, NR>(xi, yi, zi, sx, sy, sz, sq, nt, s0, s1, acc);
    } else {
      dense_tile<T, NR>(xi, yi, zi, sx, sy, sz, sq, nt, -1, -1, acc);
    }
  }
#pragma unroll
  for (int u = 0; u < kDenseRows; ++u) {
    const int64_t i = ii[u];
    if (i >= n) continue;
    const T diag = selfi ? selfi[i] * A[i] : (T)0;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      T v = ke * acc[u][r] + diag * sigma[(int64_t)(r0 + r) * n + i];
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
  const int64_t per_block = (int64_t)kDenseBlock * kDenseRows;
  const unsigned grid = (unsigned)((n + per_block - 1) / per_block);
  for (int64_t r0 = 0; r0 < nrhs; r0 += kDenseRhs) {
    const int nr = (int)((nrhs - r0) < kDenseRhs ? (nrhs - r0) : kDenseRhs);
    switch (nr) {
      case 1: hipLaunchKernelGGL((k_dense<T, 1>), dim3(grid), dim3(kDenseBlock), 0, st, C, A, sigma, selfi, out, n, (int)r0, (T)kKE); break;
      case 2: hipLaunchKernelGGL((k_dense<T, 2>), dim3(grid), dim3(kDenseBlock), 0, st, C, A, sigma, selfi, out, n, (int)r0, (T)kKE); break;
      case 3: hipLaunchKernelGGL((k_dense<T, 3>), dim3(grid), dim3(kDenseBlock), 0, st, C, A, sigma, selfi, out, n, (int)r0, (T)kKE); break;
      default: hipLaunchKernelGGL((k_dense<T, 4>), dim3(grid), dim3(kDenseBlock), 0, st, C, A, sigma, selfi, out, n, (int)r0, (T)kKE); break;
    }
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
