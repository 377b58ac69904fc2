// Dense single-layer collocation matvec (core/bem_kernel.py:157-230, :715-810):
//   V_i = sum_{j != i} K_E sigma_j A_j rsqrt(max(r_ij^2, 1e-24)) + K_ii sigma_i A_i
// with K_ii = self_integrals[i] (per-area) or 0.  Two target rows per thread,
// source panels staged through LDS in tiles of 256; up to 4 right-hand sides
// per pass.  The FP64 kernel refines a hardware FP32 reciprocal square root
// with one FP64 Newton step (relative error ~1e-14).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ed_internal.h"
#include "ed_plan.h"

namespace ed {

constexpr int kDenseBlock = 256;
constexpr int kDenseRhs = 4;
constexpr int kDenseMaxSplit = 16;

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

// ---------------------------------------------------------------------------
// Split form: a work item is a block of up to 2 * kDenseBlock target rows and
// one contiguous range of source panels (of the rows' own geometry).  Items
// write partial sums [split][rhs][n]; k_dense_finish adds the splits of each
// row in a fixed order, the self term and K_E.  Splitting the sources keeps
// the GPU full (the whole-row form runs only n / 512 workgroups) and one
// launch covers every geometry of a batched plan.
// ---------------------------------------------------------------------------
template <typename T, int NR>
__global__ __launch_bounds__(kDenseBlock) void k_dense_part(const DenseItem* __restrict__ items,
                                                            const T* __restrict__ C,
                                                            const T* __restrict__ A,
                                                            const T* __restrict__ sigma,
                                                            int64_t n, int64_t sig_ld, int r0,
                                                            T* __restrict__ partial) {
  __shared__ T sx[kDenseBlock], sy[kDenseBlock], sz[kDenseBlock];
  __shared__ T sq[kDenseRhs][kDenseBlock];
  const DenseItem it = items[blockIdx.x];
  const int64_t rend = it.row0 + it.nrows;
  int64_t ii[kDenseRows];
  T xi[kDenseRows], yi[kDenseRows], zi[kDenseRows];
#pragma unroll
  for (int u = 0; u < kDenseRows; ++u) {
    ii[u] = it.row0 + u * kDenseBlock + threadIdx.x;
    const int64_t c = ii[u] < rend ? ii[u] : rend - 1;
    xi[u] = C[3 * c]; yi[u] = C[3 * c + 1]; zi[u] = C[3 * c + 2];
  }
  T acc[kDenseRows][kDenseRhs];
#pragma unroll
  for (int u = 0; u < kDenseRows; ++u)
#pragma unroll
    for (int r = 0; r < kDenseRhs; ++r) acc[u][r] = 0;
  for (int64_t j0 = it.src0; j0 < it.src1; j0 += kDenseBlock) {
    const int64_t j = j0 + threadIdx.x;
    __syncthreads();
    if (j < it.src1) {
      sx[threadIdx.x] = C[3 * j];
      sy[threadIdx.x] = C[3 * j + 1];
      sz[threadIdx.x] = C[3 * j + 2];
      const T a = A[j];
#pragma unroll
      for (int r = 0; r < NR; ++r) sq[r][threadIdx.x] = sigma[(int64_t)(r0 + r) * sig_ld + j] * a;
    } else {
      sx[threadIdx.x] = sy[threadIdx.x] = sz[threadIdx.x] = (T)1e30;
#pragma unroll
      for (int r = 0; r < NR; ++r) sq[r][threadIdx.x] = 0;
    }
    __syncthreads();
    const int nt = (int)((it.src1 - j0) < kDenseBlock ? (it.src1 - j0) : kDenseBlock);
    // the tile holds a diagonal entry of this block only if the ranges meet
    // (uniform per block)
    if (j0 < rend && j0 + nt > it.row0) {
      const int64_t d0 = ii[0] - j0, d1 = ii[1] - j0;
      const int s0 = (d0 >= 0 && d0 < nt) ? (int)d0 : -1;
      const int s1 = (d1 >= 0 && d1 < nt) ? (int)d1 : -1;
      dense_tile<T, NR>(xi, yi, zi, sx, sy, sz, sq, nt, s0, s1, acc);
    } else {
      dense_tile<T, NR>(xi, yi, zi, sx, sy, sz, sq, nt, -1, -1, acc);
    }
  }
#pragma unroll
  for (int u = 0; u < kDenseRows; ++u) {
    if (ii[u] >= rend) continue;
#pragma unroll
    for (int r = 0; r < NR; ++r) partial[((int64_t)it.split * NR + r) * n + ii[u]] = acc[u][r];
  }
}

template <typename T, int NR>
__global__ void k_dense_finish(int64_t n, const int32_t* __restrict__ row_nsplit,
                               const T* __restrict__ partial, const T* __restrict__ sigma,
                               int64_t sig_ld, const T* __restrict__ selfi,
                               const T* __restrict__ A, int r0, T ke, T* __restrict__ out,
                               int64_t out_ld) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int S = row_nsplit[i];
  const T diag = selfi ? selfi[i] * A[i] : (T)0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    T acc = 0;
    for (int sp = 0; sp < S; ++sp) acc += partial[((int64_t)sp * NR + r) * n + i];
    T v = ke * acc + diag * sigma[(int64_t)(r0 + r) * sig_ld + i];
    if (!isfinite(v)) v = 0;  // core/bem_kernel.py:375-381
    out[(int64_t)(r0 + r) * out_ld + i] = v;
  }
}

// Work items of the split form for geometry segments [off[g], off[g+1]).
void dense_schedule(const std::vector<int64_t>& off, std::vector<DenseItem>& items,
                    std::vector<int32_t>& row_nsplit, int* max_split) {
  const int64_t n = off.back();
  row_nsplit.assign((size_t)n, 1);
  items.clear();
  *max_split = 1;
  const int64_t per_block = (int64_t)kDenseBlock * kDenseRows;
  for (size_t g = 0; g + 1 < off.size(); ++g) {
    const int64_t b = off[g], e = off[g + 1], ng = e - b;
    if (ng <= 0) continue;
    const int S = (int)std::max<int64_t>(1, std::min<int64_t>(kDenseMaxSplit, (ng + 2047) / 2048));
    *max_split = std::max(*max_split, S);
    const int64_t chunk = (ng + S - 1) / S;
    for (int64_t i = b; i < e; ++i) row_nsplit[(size_t)i] = S;
    for (int64_t r = b; r < e; r += per_block)
      for (int sp = 0; sp < S; ++sp) {
        DenseItem it;
        it.row0 = r;
        it.nrows = (int32_t)std::min<int64_t>(per_block, e - r);
        it.split = sp;
        // every (split, row) partial is written, an empty range writes zeros
        it.src0 = std::min<int64_t>(e, b + sp * chunk);
        it.src1 = std::min<int64_t>(e, b + (sp + 1) * chunk);
        items.push_back(it);
      }
  }
}

template <typename T>
int dense_run(const DenseItem* items, int64_t n_items, const int32_t* row_nsplit, int max_split,
              const T* C, const T* A, const T* sigma, const T* selfi, T* out, int64_t n,
              int64_t nrhs, T* partial, hipStream_t st) {
  const unsigned gn = (unsigned)((n + 255) / 256);
  for (int64_t r0 = 0; r0 < nrhs; r0 += kDenseRhs) {
    const int nr = (int)((nrhs - r0) < kDenseRhs ? (nrhs - r0) : kDenseRhs);
#define ED_DENSE_NR(NRV)                                                                       \
  case NRV:                                                                                     \
    hipLaunchKernelGGL((k_dense_part<T, NRV>), dim3((unsigned)n_items), dim3(kDenseBlock), 0,  \
                       st, items, C, A, sigma, n, n, (int)r0, partial);                        \
    hipLaunchKernelGGL((k_dense_finish<T, NRV>), dim3(gn), dim3(256), 0, st, n, row_nsplit,    \
                       (const T*)partial, sigma, n, selfi, A, (int)r0, (T)kKE, out, n);        \
    break;
    switch (nr) {
      ED_DENSE_NR(1)
      ED_DENSE_NR(2)
      ED_DENSE_NR(3)
      default: ED_DENSE_NR(4)
    }
#undef ED_DENSE_NR
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("dense matvec: ") + hipGetErrorString(e));
  }
  (void)max_split;
  return ED_OK;
}

template int dense_run<double>(const DenseItem*, int64_t, const int32_t*, int, const double*,
                               const double*, const double*, const double*, double*, int64_t,
                               int64_t, double*, hipStream_t);

// Stand-alone call: schedule and partial buffer per call.
template <typename T>
static int dense_launch(const T* C, const T* A, const T* sigma, const T* selfi, T* out,
                        int64_t n, int64_t nrhs, hipStream_t st) {
  if (n < 0 || nrhs < 1) return fail(ED_ERR_ARG, "dense matvec: bad sizes");
  if (n == 0) return ED_OK;
  if (!C || !A || !sigma || !out) return fail(ED_ERR_ARG, "dense matvec: NULL pointer");
  std::vector<DenseItem> items;
  std::vector<int32_t> rns;
  int smax = 1;
  dense_schedule({0, n}, items, rns, &smax);
  DenseItem* ditems = nullptr;
  int32_t* drns = nullptr;
  T* part = nullptr;
  const int nrb = (int)std::min<int64_t>(nrhs, kDenseRhs);
  int rc = ED_OK;
  if (hipMalloc((void**)&ditems, sizeof(DenseItem) * items.size()) != hipSuccess ||
      hipMalloc((void**)&drns, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc((void**)&part, sizeof(T) * (size_t)smax * nrb * n) != hipSuccess)
    rc = fail(ED_ERR_ALLOC, "dense matvec: alloc");
  hipError_t e = hipSuccess;
  if (rc == ED_OK) {
    e = hipMemcpyAsync(ditems, items.data(), sizeof(DenseItem) * items.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(drns, rns.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = fail(ED_ERR_DEVICE, std::string("dense matvec: ") + hipGetErrorString(e));
  }
  if (rc == ED_OK)
    rc = dense_run<T>(ditems, (int64_t)items.size(), drns, smax, C, A, sigma, selfi, out, n, nrhs,
                      part, st);
  if (rc == ED_OK && (e = hipStreamSynchronize(st)) != hipSuccess)
    rc = fail(ED_ERR_DEVICE, std::string("dense matvec: ") + hipGetErrorString(e));
  if (ditems) (void)hipFree(ditems);
  if (drns) (void)hipFree(drns);
  if (part) (void)hipFree(part);
  return rc;
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
