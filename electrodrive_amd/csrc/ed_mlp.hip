// Kernels of the learned preconditioner's network (training and inference).
//
//  ed_gemm_f32        C = act(op(A) op(B) + bias), FP32 MFMA (v_mfma_f32_32x32x2f32),
//                     64x64 block tile, 4 waves, K staged through LDS in chunks
//                     of 32; op() = optional transpose; act in {none, SiLU}
//  ed_act_backward    dX = dY * act'(pre)
//  ed_colsum_f32      deterministic column sums (bias gradients)
//  ed_adam_f32        Adam update (bias-corrected)
//  ed_wmse_f32        weighted squared error: loss partials and gradient
//  ed_panel_features  per-panel features from geometry, RHS and the plan's leaves
//
// No counterpart in the reference (its learned models are out of scope); the
// network exists only to produce GMRES initial guesses and a coarse space.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ed_plan.h"

namespace ed {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kGT = 64;   // block tile M and N
constexpr int kGK = 32;   // K chunk

__device__ __forceinline__ float silu(float x) { return x / (1.0f + __expf(-x)); }
__device__ __forceinline__ float silu_grad(float x) {
  const float s = 1.0f / (1.0f + __expf(-x));
  return s * (1.0f + x * (1.0f - s));
}

__global__ __launch_bounds__(256) void k_gemm(int ta, int tb, int64_t M, int64_t N, int64_t K,
                                              const float* __restrict__ A, int64_t lda,
                                              const float* __restrict__ B, int64_t ldb,
                                              float* __restrict__ C, int64_t ldc,
                                              const float* __restrict__ bias, int act,
                                              float* __restrict__ pre, float beta) {
  __shared__ float As[kGT][kGK + 1];
  __shared__ float Bs[kGK][kGT + 1];
  const int64_t m0 = (int64_t)blockIdx.x * kGT, n0 = (int64_t)blockIdx.y * kGT;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 31, lk = lane >> 5;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int64_t k0 = 0; k0 < K; k0 += kGK) {
    __syncthreads();
    for (int t = threadIdx.x; t < kGT * kGK; t += 256) {
      const int i = ta ? t % kGT : t / kGK;
      const int k = ta ? t / kGT : t % kGK;
      const int64_t gi = m0 + i, gk = k0 + k;
      float v = 0.0f;
      if (gi < M && gk < K) v = ta ? A[gk * lda + gi] : A[gi * lda + gk];
      As[i][k] = v;
    }
    for (int t = threadIdx.x; t < kGK * kGT; t += 256) {
      const int j = tb ? t / kGK : t % kGT;
      const int k = tb ? t % kGK : t / kGT;
      const int64_t gj = n0 + j, gk = k0 + k;
      float v = 0.0f;
      if (gj < N && gk < K) v = tb ? B[gj * ldb + gk] : B[gk * ldb + gj];
      Bs[k][j] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kGK; kk += 2) {
      const float a = As[wm * 32 + li][kk + lk];
      const float b = Bs[kk + lk][wn * 32 + li];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
  }
  const int64_t j = n0 + wn * 32 + li;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int64_t i = m0 + wm * 32 + (r / 4) * 8 + lk * 4 + (r % 4);
    if (i < M && j < N) {
      float v = acc[r] + (bias ? bias[j] : 0.0f);
      if (beta != 0.0f) v += beta * C[i * ldc + j];
      if (pre) pre[i * ldc + j] = v;
      C[i * ldc + j] = act == 1 ? silu(v) : v;
    }
  }
}

__global__ void k_act_bwd(const float* __restrict__ pre, const float* __restrict__ dy,
                          float* __restrict__ dx, int64_t n, int act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dx[i] = act == 1 ? dy[i] * silu_grad(pre[i]) : dy[i];
}

// out[j] = sum_i X[i][j] over rows in fixed order: block b sums a row chunk,
// then a second pass adds the chunks in order.
constexpr int kColChunks = 64;
__global__ void k_colsum_part(const float* __restrict__ X, int64_t rows, int64_t cols,
                              float* __restrict__ part) {
  const int64_t j = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  const int64_t chunk = (rows + kColChunks - 1) / kColChunks;
  const int64_t r0 = (int64_t)blockIdx.x * chunk, r1 = min(rows, r0 + chunk);
  float s = 0.0f;
  for (int64_t i = r0; i < r1; ++i) s += X[i * cols + j];
  part[(int64_t)blockIdx.x * cols + j] = s;
}
__global__ void k_colsum_fin(const float* __restrict__ part, int64_t cols, float* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.0f;
  for (int b = 0; b < kColChunks; ++b) s += part[(int64_t)b * cols + j];
  out[j] = s;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// loss = sum w_i (y_i - t_i)^2 ; grad_i = 2 w_i (y_i - t_i)
__global__ __launch_bounds__(256) void k_wmse(const float* __restrict__ y, const float* __restrict__ t,
                                              const float* __restrict__ w, int64_t n,
                                              float* __restrict__ grad, float* __restrict__ part) {
  __shared__ float sh[4];
  float s = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float d = y[i] - t[i];
    const float wi = w ? w[i] : 1.0f;
    grad[i] = 2.0f * wi * d;
    s += wi * d * d;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Per-panel features (tree-local leaf summaries).  F = 12 columns:
//  0-2 centroid relative to its geometry centre, over the geometry half size
//  3-5 unit normal
//  6   log(A / mean A of the geometry)
//  7   b / max|b| of the geometry            (RHS potential)
//  8   (b / max|b|)^3
//  9   (b/diag) / max|b/diag| of the geometry  (Jacobi guess)
//  10  leaf mean of column 7
//  11  leaf mean of column 9
// Geometry scalars (centre, half size, mean area, max|b|, max|b/diag|) come in
// `gscal` (5+3 per geometry: cx cy cz half meanA maxb maxj pad).
constexpr int kFeat = 12;
__global__ void k_features(int64_t n, const double* __restrict__ C, const double* __restrict__ Nrm,
                           const double* __restrict__ A, const double* __restrict__ b,
                           const double* __restrict__ diag, const int32_t* __restrict__ geom,
                           const double* __restrict__ gscal, float* __restrict__ F) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* g = gscal + 8 * geom[i];
  float* f = F + i * kFeat;
  for (int a = 0; a < 3; ++a) f[a] = (float)((C[3 * i + a] - g[a]) / g[3]);
  for (int a = 0; a < 3; ++a) f[3 + a] = (float)Nrm[3 * i + a];
  f[6] = (float)log(A[i] / g[4]);
  const double bn = b[i] / g[5];
  f[7] = (float)bn;
  f[8] = (float)(bn * bn * bn);
  f[9] = (float)((b[i] / diag[i]) / g[6]);
  f[10] = 0.0f;
  f[11] = 0.0f;
}

// leaf means of columns 7 and 9 (original order via perm), one wave per leaf
__global__ void k_features_leaf(const int32_t* __restrict__ leaves, const int2* __restrict__ rng,
                                const int32_t* __restrict__ perm, float* __restrict__ F) {
  const int node = leaves[blockIdx.x];
  const int2 r = rng[node];
  float s7 = 0.0f, s9 = 0.0f;
  for (int t = r.x + threadIdx.x; t < r.y; t += 64) {
    const int o = perm[t];
    s7 += F[(int64_t)o * kFeat + 7];
    s9 += F[(int64_t)o * kFeat + 9];
  }
  for (int o = 32; o > 0; o >>= 1) {
    s7 += __shfl_xor(s7, o, 64);
    s9 += __shfl_xor(s9, o, 64);
  }
  const float inv = 1.0f / (float)(r.y - r.x);
  for (int t = r.x + threadIdx.x; t < r.y; t += 64) {
    const int o = perm[t];
    F[(int64_t)o * kFeat + 10] = s7 * inv;
    F[(int64_t)o * kFeat + 11] = s9 * inv;
  }
}

// Projection loss of a learned basis.  Solve s owns rows s*m .. s*m+m-1;
// its basis is Z = [Y | X] (ky network columns, kx fixed columns, k <= 32).
// c = argmin |t - Z c|^2 + lam |c|^2 with lam = ridge * trace(Z^T Z) / k,
// r = t - Z c, loss_s = |r|^2 / |t|^2 and dY = -2 r c_Y^T / |t|^2 (lam held
// fixed).  One block per solve; the k x k system is solved in FP64 by
// Cholesky in LDS.
constexpr int kProjMaxK = 32;
__global__ __launch_bounds__(256) void k_proj_loss(const float* __restrict__ Y, int ky,
                                                   const float* __restrict__ X, int kx,
                                                   const float* __restrict__ T, int m, float ridge,
                                                   float* __restrict__ dY,
                                                   float* __restrict__ loss) {
  __shared__ double G[kProjMaxK][kProjMaxK + 1];
  __shared__ double h[kProjMaxK], c[kProjMaxK];
  __shared__ double red[256];
  const int k = ky + kx;
  const int64_t r0 = (int64_t)blockIdx.x * m;
  auto z = [&](int64_t i, int q) -> double {
    return q < ky ? (double)Y[i * ky + q] : (double)X[i * kx + (q - ky)];
  };
  for (int e = threadIdx.x; e < k * k + k; e += blockDim.x) {
    double acc = 0.0;
    if (e < k * k) {
      const int a = e / k, b = e % k;
      if (b >= a)
        for (int i = 0; i < m; ++i) acc += z(r0 + i, a) * z(r0 + i, b);
      if (b >= a) { G[a][b] = acc; G[b][a] = acc; }
    } else {
      const int a = e - k * k;
      for (int i = 0; i < m; ++i) acc += z(r0 + i, a) * (double)T[r0 + i];
      h[a] = acc;
    }
  }
  double tt = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) tt += (double)T[r0 + i] * T[r0 + i];
  red[threadIdx.x] = tt;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  tt = red[0];
  if (threadIdx.x == 0) {
    double tr = 0.0;
    for (int a = 0; a < k; ++a) tr += G[a][a];
    const double lam = (double)ridge * (tr > 0.0 ? tr : 1.0) / k;
    for (int a = 0; a < k; ++a) G[a][a] += lam;
    // Cholesky G = L L^T in place (lower triangle)
    for (int j = 0; j < k; ++j) {
      double d = G[j][j];
      for (int q = 0; q < j; ++q) d -= G[j][q] * G[j][q];
      d = d > 1e-300 ? sqrt(d) : 1e-150;
      G[j][j] = d;
      for (int i = j + 1; i < k; ++i) {
        double v = G[i][j];
        for (int q = 0; q < j; ++q) v -= G[i][q] * G[j][q];
        G[i][j] = v / d;
      }
    }
    for (int i = 0; i < k; ++i) {  // L y = h
      double v = h[i];
      for (int q = 0; q < i; ++q) v -= G[i][q] * c[q];
      c[i] = v / G[i][i];
    }
    for (int i = k - 1; i >= 0; --i) {  // L^T c = y
      double v = c[i];
      for (int q = i + 1; q < k; ++q) v -= G[q][i] * c[q];
      c[i] = v / G[i][i];
    }
  }
  __syncthreads();
  const double inv_tt = tt > 0.0 ? 1.0 / tt : 0.0;
  double rr = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const int64_t row = r0 + i;
    double v = (double)T[row];
    for (int q = 0; q < k; ++q) v -= z(row, q) * c[q];
    rr += v * v;
    for (int q = 0; q < ky; ++q) dY[row * ky + q] = (float)(-2.0 * v * c[q] * inv_tt);
  }
  __syncthreads();
  red[threadIdx.x] = rr;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[blockIdx.x] = (float)(red[0] * inv_tt);
}

#define ED_MLP_CHECK(msg)                                                          \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess)                                                          \
      return fail(ED_ERR_LAUNCH, std::string(msg) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace ed

using namespace ed;

extern "C" {

int ed_gemm_f32(int32_t trans_a, int32_t trans_b, int64_t M, int64_t N, int64_t K,
                const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                int64_t ldc, const float* bias, int32_t act, float* pre_act, float beta,
                ed_stream stream) {
  if (M < 0 || N < 0 || K < 0 || act < 0 || act > 1) return fail(ED_ERR_ARG, "ed_gemm_f32: bad sizes");
  if (M == 0 || N == 0) return ED_OK;
  if (!A || !B || !C) return fail(ED_ERR_ARG, "ed_gemm_f32: NULL pointer");
  dim3 grid((unsigned)((M + kGT - 1) / kGT), (unsigned)((N + kGT - 1) / kGT));
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, (hipStream_t)stream, trans_a, trans_b, M, N, K,
                     A, lda, B, ldb, C, ldc, bias, act, pre_act, beta);
  ED_MLP_CHECK("ed_gemm_f32");
  return ED_OK;
}

int ed_act_backward_f32(const float* pre, const float* dy, float* dx, int64_t n, int32_t act,
                        ed_stream stream) {
  if (n < 0 || (n > 0 && (!pre || !dy || !dx))) return fail(ED_ERR_ARG, "ed_act_backward_f32");
  if (n == 0) return ED_OK;
  hipLaunchKernelGGL(k_act_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, pre, dy, dx, n, act);
  ED_MLP_CHECK("ed_act_backward_f32");
  return ED_OK;
}

// scratch: >= 64 * cols floats of device memory
int ed_colsum_f32(const float* X, int64_t rows, int64_t cols, float* scratch, float* out,
                  ed_stream stream) {
  if (rows < 0 || cols <= 0 || !X || !scratch || !out) return fail(ED_ERR_ARG, "ed_colsum_f32");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_colsum_part, dim3(kColChunks, (unsigned)((cols + 255) / 256)), dim3(256), 0,
                     st, X, rows, cols, scratch);
  hipLaunchKernelGGL(k_colsum_fin, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, st, scratch,
                     cols, out);
  ED_MLP_CHECK("ed_colsum_f32");
  return ED_OK;
}

int ed_adam_f32(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                float b2, float eps, int32_t step, float weight_decay, ed_stream stream) {
  if (n <= 0 || !p || !g || !m || !v || step < 1) return fail(ED_ERR_ARG, "ed_adam_f32");
  const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     p, g, m, v, n, lr, b1, b2, eps, c1, c2, weight_decay);
  ED_MLP_CHECK("ed_adam_f32");
  return ED_OK;
}

// part: >= 240 floats; loss (host) = sum of partials (the call synchronises)
int ed_wmse_f32(const float* y, const float* t, const float* w, int64_t n, float* grad,
                float* part, float* loss_host, ed_stream stream) {
  if (n <= 0 || !y || !t || !grad || !part) return fail(ED_ERR_ARG, "ed_wmse_f32");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_wmse, dim3(240), dim3(256), 0, st, y, t, w, n, grad, part);
  ED_MLP_CHECK("ed_wmse_f32");
  if (loss_host) {
    float hp[240];
    hipError_t e = hipMemcpyAsync(hp, part, sizeof(hp), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(ED_ERR_DEVICE, "ed_wmse_f32: copy");
    float s = 0.0f;
    for (int k = 0; k < 240; ++k) s += hp[k];
    *loss_host = s;
  }
  return ED_OK;
}

// Features of the plan's panels (original order) into F (n x 12, device FP32).
// centroids/normals (n,3), areas/b/diag (n) device FP64 in original order.
int ed_panel_features(ed_plan* P, const double* C, const double* Nrm, const double* A,
                      const double* b, const double* diag, float* F, ed_stream stream) {
  if (!P || !C || !Nrm || !A || !b || !diag || !F) return fail(ED_ERR_ARG, "ed_panel_features: NULL");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = P->n;
  const auto& go = P->host.geom_pt_off;
  const int64_t G = (int64_t)go.size() - 1;
  // geometry scalars on the host (set-up cost, O(n) host work on copies)
  std::vector<double> hC(3 * n), hA(n), hb(n), hd(n);
  hipError_t e = hipMemcpyAsync(hC.data(), C, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(hA.data(), A, sizeof(double) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(hb.data(), b, sizeof(double) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(hd.data(), diag, sizeof(double) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(ED_ERR_DEVICE, "ed_panel_features: copy");
  std::vector<double> gs(8 * G, 0.0);
  std::vector<int32_t> geom(n);
  for (int64_t g = 0; g < G; ++g) {
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300}, sa = 0, mb = 0, mj = 0;
    for (int64_t t = go[g]; t < go[g + 1]; ++t) {
      const int64_t o = P->host.perm[t];
      geom[o] = (int32_t)g;
      for (int a = 0; a < 3; ++a) { mn[a] = std::min(mn[a], hC[3 * o + a]); mx[a] = std::max(mx[a], hC[3 * o + a]); }
      sa += hA[o];
      mb = std::max(mb, std::fabs(hb[o]));
      mj = std::max(mj, std::fabs(hb[o] / hd[o]));
    }
    double half = 0;
    for (int a = 0; a < 3; ++a) { gs[8 * g + a] = 0.5 * (mn[a] + mx[a]); half = std::max(half, 0.5 * (mx[a] - mn[a])); }
    gs[8 * g + 3] = half > 0 ? half : 1.0;
    gs[8 * g + 4] = sa / (double)(go[g + 1] - go[g]);
    gs[8 * g + 5] = mb > 0 ? mb : 1.0;
    gs[8 * g + 6] = mj > 0 ? mj : 1.0;
  }
  double* dgs = nullptr;
  int32_t* dgeom = nullptr;
  if (hipMalloc((void**)&dgs, sizeof(double) * gs.size()) != hipSuccess ||
      hipMalloc((void**)&dgeom, sizeof(int32_t) * n) != hipSuccess)
    return fail(ED_ERR_ALLOC, "ed_panel_features: alloc");
  (void)hipMemcpyAsync(dgs, gs.data(), sizeof(double) * gs.size(), hipMemcpyHostToDevice, st);
  (void)hipMemcpyAsync(dgeom, geom.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st);
  hipLaunchKernelGGL(k_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, C, Nrm, A,
                     b, diag, dgeom, dgs, F);
  hipLaunchKernelGGL(k_features_leaf, dim3((unsigned)P->n_leaves), dim3(64), 0, st, P->leaves,
                     P->node_rng, P->perm, F);
  e = hipGetLastError();
  (void)hipStreamSynchronize(st);
  (void)hipFree(dgs);
  (void)hipFree(dgeom);
  if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("ed_panel_features: ") + hipGetErrorString(e));
  return ED_OK;
}

// Projection loss of a learned coarse basis (no reference counterpart), see
// k_proj_loss.  Y (n_sol*m x ky), X (n_sol*m x kx), T (n_sol*m), dY like Y,
// loss (n_sol): device FP32.
int ed_proj_loss_f32(const float* Y, int32_t ky, const float* X, int32_t kx, const float* T,
                     int64_t n_sol, int32_t m, float ridge, float* dY, float* loss,
                     ed_stream stream) {
  if (ky < 1 || kx < 0 || ky + kx > kProjMaxK || n_sol < 0 || m < 1 || !(ridge >= 0.0f))
    return fail(ED_ERR_ARG, "ed_proj_loss_f32: need 1 <= ky, ky + kx <= 32, m >= 1, ridge >= 0");
  if (n_sol == 0) return ED_OK;
  if (!Y || (kx > 0 && !X) || !T || !dY || !loss) return fail(ED_ERR_ARG, "ed_proj_loss_f32: NULL");
  hipLaunchKernelGGL(k_proj_loss, dim3((unsigned)n_sol), dim3(256), 0, (hipStream_t)stream, Y, ky,
                     X, kx, T, m, ridge, dY, loss);
  ED_MLP_CHECK("ed_proj_loss_f32");
  return ED_OK;
}

int32_t ed_panel_feature_count(void) { return kFeat; }

}  // extern "C"
