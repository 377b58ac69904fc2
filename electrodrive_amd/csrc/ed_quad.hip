// Panel quadrature, near-field correction and Morton keys (FP64 set-up path).
//
//  k_self_integrals  core/bem.py:1777-1816 with core/bem_quadrature.py:217-308
//  k_near_pairs      core/bem.py:475-555 (count + fill, row i-major, j ascending)
//  k_near_weights    core/bem.py:558-667 (uniform 4^depth refinement of panel j)
//  k_near_apply      core/bem_near_cuda_kernel.cu:110-232 semantics without
//                    atomics: one thread per output row over i-major pairs
//  k_morton          fmm3d/geometry.py:671-799
// Distance tests of the near-pair search are evaluated with FP contraction
// off, so the selected index set matches the reference bit for bit.
#include <hip/hip_runtime.h>

#include <vector>

#include "ed_internal.h"

namespace ed {

#define ED_CHECK_LAUNCH(msg)                                                      \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess)                                                         \
      return fail(ED_ERR_LAUNCH, std::string(msg) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Tri {
  double v[9];
};

__device__ __forceinline__ double tri_area(const double* v) {
#pragma clang fp contract(off)
  const double e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
  const double e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
  const double cx = e1y * e2z - e1z * e2y;
  const double cy = e1z * e2x - e1x * e2z;
  const double cz = e1x * e2y - e1y * e2x;
  return 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

// Standard rule (core/bem_quadrature.py:79-150) on one triangle, accumulated
// as sum K_E w / max(|t - y|, 1e-12).
__device__ __forceinline__ double rule_accum(const double* v, int order, double tx, double ty,
                                             double tz) {
#pragma clang fp contract(off)
  const double area = tri_area(v);
  if (area <= 0.0) return 0.0;
  double acc = 0.0;
  if (order <= 1) {
    const double b = 1.0 / 3.0;
    const double px = b * v[0] + b * v[3] + b * v[6];
    const double py = b * v[1] + b * v[4] + b * v[7];
    const double pz = b * v[2] + b * v[5] + b * v[8];
    const double dx = tx - px, dy = ty - py, dz = tz - pz;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    r = r > 1e-12 ? r : 1e-12;
    acc += (kKE / r) * (1.0 * area);
  } else {
    const double a = 2.0 / 3.0, b = 1.0 / 6.0;
    const double w = (1.0 / 3.0) * area;
    for (int q = 0; q < 3; ++q) {
      const double b0 = q == 0 ? a : b, b1 = q == 1 ? a : b, b2 = q == 2 ? a : b;
      const double px = b0 * v[0] + b1 * v[3] + b2 * v[6];
      const double py = b0 * v[1] + b1 * v[4] + b2 * v[7];
      const double pz = b0 * v[2] + b1 * v[5] + b2 * v[8];
      const double dx = tx - px, dy = ty - py, dz = tz - pz;
      double r = sqrt(dx * dx + dy * dy + dz * dz);
      r = r > 1e-12 ? r : 1e-12;
      acc += (kKE / r) * w;
    }
  }
  return acc;
}

__device__ __forceinline__ void split4(const double* t, int c, double* o) {
#pragma clang fp contract(off)
  double m01[3], m12[3], m20[3];
  for (int k = 0; k < 3; ++k) {
    m01[k] = 0.5 * (t[k] + t[3 + k]);
    m12[k] = 0.5 * (t[3 + k] + t[6 + k]);
    m20[k] = 0.5 * (t[6 + k] + t[k]);
  }
  for (int k = 0; k < 3; ++k) {
    if (c == 0) { o[k] = t[k];    o[3 + k] = m01[k];    o[6 + k] = m20[k]; }
    if (c == 1) { o[k] = m01[k];  o[3 + k] = t[3 + k];  o[6 + k] = m12[k]; }
    if (c == 2) { o[k] = m20[k];  o[3 + k] = m12[k];    o[6 + k] = t[6 + k]; }
    if (c == 3) { o[k] = m01[k];  o[3 + k] = m12[k];    o[6 + k] = m20[k]; }
  }
}

constexpr int kMaxSelfDepth = 6;

__global__ void k_self_integrals(const double* __restrict__ PV, const double* __restrict__ C,
                                 int64_t n, int depth, int order, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Tri tris[1 + 3 * kMaxSelfDepth];
  int cnt = 1;
  for (int k = 0; k < 9; ++k) tris[0].v[k] = PV[9 * i + k];
  const double tx = C[3 * i], ty = C[3 * i + 1], tz = C[3 * i + 2];
  for (int d = 0; d < depth; ++d) {
    int best = 0;
    double bd = 0.0;
    for (int t = 0; t < cnt; ++t) {
      const double* v = tris[t].v;
      const double cx = (v[0] + v[3] + v[6]) / 3.0;
      const double cy = (v[1] + v[4] + v[7]) / 3.0;
      const double cz = (v[2] + v[5] + v[8]) / 3.0;
      const double dx = cx - tx, dy = cy - ty, dz = cz - tz;
      const double dd = sqrt(dx * dx + dy * dy + dz * dz);
      if (t == 0 || dd < bd) { bd = dd; best = t; }
    }
    Tri sel = tris[best];
    for (int t = best; t + 1 < cnt; ++t) tris[t] = tris[t + 1];
    --cnt;
    for (int c = 0; c < 4; ++c) split4(sel.v, c, tris[cnt + c].v);
    cnt += 4;
  }
  double acc = 0.0;
  for (int t = 0; t < cnt; ++t) acc += rule_accum(tris[t].v, order, tx, ty, tz);
  out[i] = acc;
}

// Near pairs: phase 0 counts per row, phase 1 fills at row offsets.
constexpr int kNearBlock = 256;
__global__ __launch_bounds__(kNearBlock) void k_near_pairs(const double* __restrict__ C,
                                                           const double* __restrict__ A,
                                                           int64_t n, double df,
                                                           int64_t* __restrict__ counts,
                                                           const int64_t* __restrict__ offs,
                                                           int64_t* __restrict__ pairs) {
#pragma clang fp contract(off)
  __shared__ double sx[kNearBlock], sy[kNearBlock], sz[kNearBlock], sr[kNearBlock];
  const int64_t i = (int64_t)blockIdx.x * kNearBlock + threadIdx.x;
  const bool active = i < n;
  double xi = 0, yi = 0, zi = 0, ri = 0;
  const double pi = 3.141592653589793;
  if (active) {
    xi = C[3 * i]; yi = C[3 * i + 1]; zi = C[3 * i + 2];
    double a = A[i];
    a = a > 0.0 ? a : 0.0;
    ri = sqrt(a / pi);
  }
  int64_t cnt = 0;
  int64_t w = (offs && active) ? offs[i] : 0;
  for (int64_t j0 = 0; j0 < n; j0 += kNearBlock) {
    const int64_t j = j0 + threadIdx.x;
    __syncthreads();
    if (j < n) {
      sx[threadIdx.x] = C[3 * j]; sy[threadIdx.x] = C[3 * j + 1]; sz[threadIdx.x] = C[3 * j + 2];
      double a = A[j];
      a = a > 0.0 ? a : 0.0;
      sr[threadIdx.x] = sqrt(a / pi);
    }
    __syncthreads();
    if (!active) continue;
    const int nt = (int)((n - j0) < kNearBlock ? (n - j0) : kNearBlock);
    for (int t = 0; t < nt; ++t) {
      const double dx = xi - sx[t], dy = yi - sy[t], dz = zi - sz[t];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const double th = df * (ri + sr[t]);
      if (isfinite(d) && isfinite(th) && d > 0.0 && d <= th) {
        if (pairs) { pairs[2 * w] = i; pairs[2 * w + 1] = j0 + t; ++w; }
        ++cnt;
      }
    }
  }
  if (active && counts) counts[i] = cnt;
}

__global__ void k_near_weights(const double* __restrict__ C, const double* __restrict__ A,
                               const double* __restrict__ PV, const int64_t* __restrict__ pairs,
                               int64_t np, int order, int depth, double* __restrict__ wout) {
#pragma clang fp contract(off)
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int64_t i = pairs[2 * q], j = pairs[2 * q + 1];
  const double Aj = A[j];
  if (Aj <= 0.0) { wout[q] = 0.0; return; }  // core/bem.py:642-644 leaves 0
  const double tx = C[3 * i], ty = C[3 * i + 1], tz = C[3 * i + 2];
  double base[9];
  for (int k = 0; k < 9; ++k) base[k] = PV[9 * j + k];
  double acc = 0.0;
  const int64_t nsub = 1LL << (2 * depth);
  for (int64_t t = 0; t < nsub; ++t) {
    double cur[9], nxt[9];
    for (int k = 0; k < 9; ++k) cur[k] = base[k];
    for (int d = depth - 1; d >= 0; --d) {
      const int c = (int)((t >> (2 * d)) & 3);
      split4(cur, c, nxt);
      for (int k = 0; k < 9; ++k) cur[k] = nxt[k];
    }
    acc += rule_accum(cur, order, tx, ty, tz);
  }
  const double dx = tx - C[3 * j], dy = ty - C[3 * j + 1], dz = tz - C[3 * j + 2];
  double rf = sqrt(dx * dx + dy * dy + dz * dz);
  rf = rf > 1e-12 ? rf : 1e-12;
  wout[q] = acc - (kKE * Aj) / rf;
}

__device__ __forceinline__ int64_t lower_row(const int64_t* pairs, int64_t np, int64_t row) {
  int64_t lo = 0, hi = np;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (pairs[2 * mid] < row) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// On-the-fly delta (CUDA kernel semantics, core/bem_near_cuda_kernel.cu:140-231):
// I_near = sum_k K_E (w_k A_j)/max(|c_i - y_k|, 1e-12) with the reference-triangle
// rule (weights normalised to 1), I_far = K_E A_j / max(|c_i - c_j|, 1e-12).
__global__ void k_near_apply(const double* __restrict__ vfar, const double* __restrict__ sigma,
                             const double* __restrict__ C, const double* __restrict__ A,
                             const double* __restrict__ PV, const int64_t* __restrict__ pairs,
                             const double* __restrict__ W, int64_t np, int64_t n, int order,
                             double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b = lower_row(pairs, np, i), e = lower_row(pairs, np, i + 1);
  double acc = 0.0;
  for (int64_t q = b; q < e; ++q) {
    const int64_t j = pairs[2 * q + 1];
    if (j < 0 || j >= n) continue;
    const double s = sigma[j];
    double d;
    if (W) {
      d = W[q];
    } else {
      const double Aj = A[j];
      if (Aj <= 0.0 || s == 0.0) continue;
      const double* v = PV + 9 * j;
      const double tx = C[3 * i], ty = C[3 * i + 1], tz = C[3 * i + 2];
      double In = 0.0;
      const int Q = order <= 1 ? 1 : 3;
      for (int k = 0; k < Q; ++k) {
        double u, w, wk;
        if (Q == 1) { u = 1.0 / 3.0; w = 1.0 / 3.0; wk = 1.0; }
        else {
          u = k == 0 ? 1.0 / 6.0 : (k == 1 ? 2.0 / 3.0 : 1.0 / 6.0);
          w = k == 0 ? 1.0 / 6.0 : (k == 1 ? 1.0 / 6.0 : 2.0 / 3.0);
          wk = 1.0 / 3.0;
        }
        const double px = v[0] + u * (v[3] - v[0]) + w * (v[6] - v[0]);
        const double py = v[1] + u * (v[4] - v[1]) + w * (v[7] - v[1]);
        const double pz = v[2] + u * (v[5] - v[2]) + w * (v[8] - v[2]);
        const double dx = tx - px, dy = ty - py, dz = tz - pz;
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        r = r > 1e-12 ? r : 1e-12;
        In += kKE * (wk * Aj) / r;
      }
      const double dx = tx - C[3 * j], dy = ty - C[3 * j + 1], dz = tz - C[3 * j + 2];
      double rf = sqrt(dx * dx + dy * dy + dz * dz);
      rf = rf > 1e-12 ? rf : 1e-12;
      d = In - kKE * Aj / rf;
    }
    acc += s * d;
  }
  out[i] = vfar[i] + acc;
}

template <typename T>
__global__ void k_bbox(const T* __restrict__ P, int64_t n, T* __restrict__ mm) {
  __shared__ T smin[3][256], smax[3][256];
  T lo[3], hi[3];
  for (int a = 0; a < 3; ++a) { lo[a] = P[a]; hi[a] = P[a]; }
  for (int64_t i = threadIdx.x; i < n; i += 256)
    for (int a = 0; a < 3; ++a) {
      const T v = P[3 * i + a];
      lo[a] = v < lo[a] ? v : lo[a];
      hi[a] = v > hi[a] ? v : hi[a];
    }
  for (int a = 0; a < 3; ++a) { smin[a][threadIdx.x] = lo[a]; smax[a][threadIdx.x] = hi[a]; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < 256; ++t)
      for (int a = 0; a < 3; ++a) {
        lo[a] = smin[a][t] < lo[a] ? smin[a][t] : lo[a];
        hi[a] = smax[a][t] > hi[a] ? smax[a][t] : hi[a];
      }
    // geometry.py:678-693 degenerate-axis padding
    T e[3], me = 0;
    for (int a = 0; a < 3; ++a) { e[a] = hi[a] - lo[a]; const T ae = e[a] < 0 ? -e[a] : e[a]; me = ae > me ? ae : me; }
    const T padded = (T)1e-12 * ((T)1 + me);
    for (int a = 0; a < 3; ++a) { mm[a] = lo[a]; mm[3 + a] = lo[a] + (e[a] > 0 ? e[a] : padded); }
  }
}

template <typename T>
__global__ void k_morton(const T* __restrict__ P, int64_t n, int level, const T* __restrict__ mm,
                         T hi, T eps, int64_t* __restrict__ keys) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T scale = (T)(double)(1LL << level);
  const int64_t maxi = (1LL << level) - 1;
  uint64_t ijk[3];
  for (int a = 0; a < 3; ++a) {
    T ext = mm[3 + a] - mm[a];
    ext = ext > eps ? ext : eps;
    T v = (P[3 * i + a] - mm[a]) / ext;
    v = v < (T)0 ? (T)0 : v;
    v = v > hi ? hi : v;
    int64_t q = (int64_t)floor(v * scale);
    q = q < 0 ? 0 : (q > maxi ? maxi : q);
    ijk[a] = (uint64_t)q;
  }
  uint64_t k = 0;
  for (int b = 0; b < level; ++b) {
    k |= ((ijk[0] >> b) & 1ULL) << (3 * b);
    k |= ((ijk[1] >> b) & 1ULL) << (3 * b + 1);
    k |= ((ijk[2] >> b) & 1ULL) << (3 * b + 2);
  }
  keys[i] = (int64_t)k;
}

template <typename T>
static int morton_launch(const T* P, int64_t n, int level, const T* bc, const T* bh,
                         int64_t* keys, hipStream_t st, T hi, T eps) {
  if (level < 1 || level > 21) return fail(ED_ERR_ARG, "level must be in [1, 21]");
  if (n < 0 || (n > 0 && (!P || !keys))) return fail(ED_ERR_ARG, "morton keys: NULL pointer");
  if ((bc == nullptr) != (bh == nullptr)) return fail(ED_ERR_ARG, "bbox center/half mismatch");
  if (n == 0) return ED_OK;
  T* mm = nullptr;
  if (hipMalloc((void**)&mm, 6 * sizeof(T)) != hipSuccess) return fail(ED_ERR_ALLOC, "morton alloc");
  if (bc) {
    // geometry.py:261-268: mins/maxs = center -/+ half in the points' dtype
    T h[6];
    for (int a = 0; a < 3; ++a) { h[a] = bc[a] - bh[a]; h[3 + a] = bc[a] + bh[a]; }
    (void)hipMemcpyAsync(mm, h, sizeof(h), hipMemcpyHostToDevice, st);
  } else {
    hipLaunchKernelGGL(k_bbox<T>, dim3(1), dim3(256), 0, st, P, n, mm);
  }
  hipLaunchKernelGGL(k_morton<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P, n,
                     level, (const T*)mm, hi, eps, keys);
  hipError_t e = hipGetLastError();
  (void)hipStreamSynchronize(st);
  (void)hipFree(mm);
  if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("morton: ") + hipGetErrorString(e));
  return ED_OK;
}

}  // namespace ed

using namespace ed;

extern "C" {

int ed_self_integrals_f64(const double* PV, const double* C, int64_t n, int32_t depth,
                          int32_t order, double* out, ed_stream stream) {
  if (n < 0 || depth < 0 || depth > kMaxSelfDepth) return fail(ED_ERR_ARG, "self integrals: bad depth/n");
  if (n == 0) return ED_OK;
  if (!PV || !C || !out) return fail(ED_ERR_ARG, "self integrals: NULL pointer");
  hipLaunchKernelGGL(k_self_integrals, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                     (hipStream_t)stream, PV, C, n, depth, order, out);
  ED_CHECK_LAUNCH("self integrals");
  return ED_OK;
}

int ed_near_pairs_f64(const double* C, const double* A, int64_t n, double df, int64_t* pairs,
                      int64_t capacity, int64_t* n_pairs_out, ed_stream stream) {
  if (!n_pairs_out) return fail(ED_ERR_ARG, "near pairs: NULL count");
  *n_pairs_out = 0;
  if (n <= 1) return ED_OK;
  if (!C || !A) return fail(ED_ERR_ARG, "near pairs: NULL geometry");
  hipStream_t st = (hipStream_t)stream;
  int64_t *counts = nullptr, *offs = nullptr;
  if (hipMalloc((void**)&counts, sizeof(int64_t) * n) != hipSuccess ||
      hipMalloc((void**)&offs, sizeof(int64_t) * n) != hipSuccess) {
    if (counts) (void)hipFree(counts);
    return fail(ED_ERR_ALLOC, "near pairs: alloc");
  }
  const unsigned grid = (unsigned)((n + kNearBlock - 1) / kNearBlock);
  hipLaunchKernelGGL(k_near_pairs, dim3(grid), dim3(kNearBlock), 0, st, C, A, n, df, counts,
                     (const int64_t*)nullptr, (int64_t*)nullptr);
  std::vector<int64_t> hc(n);
  hipError_t e = hipMemcpyAsync(hc.data(), counts, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  int rc = ED_OK;
  if (e != hipSuccess) rc = fail(ED_ERR_DEVICE, std::string("near pairs: ") + hipGetErrorString(e));
  int64_t total = 0;
  if (rc == ED_OK) {
    for (int64_t i = 0; i < n; ++i) { const int64_t c = hc[i]; hc[i] = total; total += c; }
    *n_pairs_out = total;
    if (pairs) {
      if (capacity < total) rc = fail(ED_ERR_ARG, "near pairs: capacity too small");
      else {
        (void)hipMemcpyAsync(offs, hc.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st);
        hipLaunchKernelGGL(k_near_pairs, dim3(grid), dim3(kNearBlock), 0, st, C, A, n, df,
                           (int64_t*)nullptr, (const int64_t*)offs, pairs);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(ED_ERR_LAUNCH, std::string("near pairs: ") + hipGetErrorString(e));
      }
    }
  }
  (void)hipFree(counts);
  (void)hipFree(offs);
  return rc;
}

int ed_near_weights_f64(const double* C, const double* A, const double* PV, const int64_t* pairs,
                        int64_t np, int32_t order, int32_t depth, double* wout, ed_stream stream) {
  if (np < 0 || depth < 0 || depth > 6) return fail(ED_ERR_ARG, "near weights: bad sizes");
  if (np == 0) return ED_OK;
  if (!C || !A || !PV || !pairs || !wout) return fail(ED_ERR_ARG, "near weights: NULL pointer");
  hipLaunchKernelGGL(k_near_weights, dim3((unsigned)((np + 127) / 128)), dim3(128), 0,
                     (hipStream_t)stream, C, A, PV, pairs, np, order, depth, wout);
  ED_CHECK_LAUNCH("near weights");
  return ED_OK;
}

int ed_near_apply_f64(const double* vfar, const double* sigma, const double* C, const double* A,
                      const double* PV, const int64_t* pairs, const double* W, int64_t np,
                      int64_t n, int32_t order, double* out, ed_stream stream) {
  if (n < 0 || np < 0) return fail(ED_ERR_ARG, "near apply: bad sizes");
  if (n == 0) return ED_OK;
  if (!vfar || !sigma || !out || (np > 0 && !pairs)) return fail(ED_ERR_ARG, "near apply: NULL pointer");
  if (np > 0 && !W && (!C || !A || !PV)) return fail(ED_ERR_ARG, "near apply: geometry required without weights");
  hipLaunchKernelGGL(k_near_apply, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                     (hipStream_t)stream, vfar, sigma, C, A, PV, pairs, W, np, n, order, out);
  ED_CHECK_LAUNCH("near apply");
  return ED_OK;
}

int ed_morton_keys_f64(const double* P, int64_t n, int32_t level, const double* bc,
                       const double* bh, int64_t* keys, ed_stream stream) {
  return morton_launch<double>(P, n, level, bc, bh, keys, (hipStream_t)stream, 1.0 - 1e-9,
                               2.220446049250313e-16);
}

int ed_morton_keys_f32(const float* P, int64_t n, int32_t level, const float* bc,
                       const float* bh, int64_t* keys, ed_stream stream) {
  return morton_launch<float>(P, n, level, bc, bh, keys, (hipStream_t)stream, (float)(1.0 - 1e-9),
                              1.1920928955078125e-07f);
}

}  // extern "C"
