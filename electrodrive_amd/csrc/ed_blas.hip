// FP64 level-1 vector kernels used by the generic (callable-operator) GMRES
// path: deterministic dot product (fixed-order two-stage reduction), axpy,
// scale and Jacobi apply.  They replace the torch.dot / axpy calls of
// core/bem_solver.py:536-541 when the operator is an arbitrary callable.
#include <hip/hip_runtime.h>

#include "ed_internal.h"

namespace ed {

constexpr int kDotBlocks = 240;

__global__ __launch_bounds__(256) void k_dot_partial(const double* __restrict__ x,
                                                     const double* __restrict__ y, int64_t n,
                                                     double* __restrict__ part) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void k_dot_final(const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < kDotBlocks; ++b) s += part[b];
    out[0] = s;
  }
}

__global__ void k_axpy64(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

__global__ void k_scale64(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

__global__ void k_mul64(const double* __restrict__ d, const double* __restrict__ x,
                        double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = d[i] * x[i];
}

}  // namespace ed

using namespace ed;

extern "C" {

// *result (host) = x . y ; synchronises the stream.  scratch: >= 241 doubles
// of device memory owned by the caller.
int ed_vec_dot_f64(const double* x, const double* y, int64_t n, double* scratch,
                   double* result_host, ed_stream stream) {
  if (!result_host || n < 0 || (n > 0 && (!x || !y || !scratch)))
    return fail(ED_ERR_ARG, "ed_vec_dot_f64: bad argument");
  if (n == 0) { *result_host = 0.0; return ED_OK; }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(256), 0, st, x, y, n, scratch);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(64), 0, st, scratch, scratch + kDotBlocks);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(result_host, scratch + kDotBlocks, sizeof(double),
                                          hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(ED_ERR_DEVICE, std::string("ed_vec_dot_f64: ") + hipGetErrorString(e));
  return ED_OK;
}

int ed_vec_axpy_f64(double a, const double* x, double* y, int64_t n, ed_stream stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return fail(ED_ERR_ARG, "ed_vec_axpy_f64: bad argument");
  if (n == 0) return ED_OK;
  hipLaunchKernelGGL(k_axpy64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, a, x, y, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? ED_OK : fail(ED_ERR_LAUNCH, "ed_vec_axpy_f64 launch");
}

int ed_vec_scale_f64(double a, const double* x, double* y, int64_t n, ed_stream stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return fail(ED_ERR_ARG, "ed_vec_scale_f64: bad argument");
  if (n == 0) return ED_OK;
  hipLaunchKernelGGL(k_scale64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, a, x, y, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? ED_OK : fail(ED_ERR_LAUNCH, "ed_vec_scale_f64 launch");
}

int ed_vec_mul_f64(const double* d, const double* x, double* y, int64_t n, ed_stream stream) {
  if (n < 0 || (n > 0 && (!d || !x || !y))) return fail(ED_ERR_ARG, "ed_vec_mul_f64: bad argument");
  if (n == 0) return ED_OK;
  hipLaunchKernelGGL(k_mul64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, d, x, y, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? ED_OK : fail(ED_ERR_LAUNCH, "ed_vec_mul_f64 launch");
}

}  // extern "C"
