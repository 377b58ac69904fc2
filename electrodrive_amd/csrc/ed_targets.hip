// Induced potential and field at arbitrary target points from solved panel
// charges (post-solve evaluation):
//
//   bem_potential_targets  core/bem_kernel.py:233-277, :818-918
//     V(x_i) = sum_j K_E sigma_j A_j rsqrt(max(|x_i - c_j|^2, 1e-24))
//   bem_E_field_targets    core/bem_kernel.py:280-335, :921-1008
//     E(x_i) = sum_j K_E sigma_j A_j (x_i - c_j) rsqrt(max(r^2, 1e-24))^3
//   _apply_near_quadrature_potentials  core/bem.py:940-1047
//     for panels with |c_j - x_i| <= f R_j (R = sqrt(A/pi)), A_j > 0, sigma_j != 0:
//     V_i += sigma_j (I_near - K_E A_j / max(|c_j - x_i|, 1e-12)), I_near the
//     panel integral refined `depth` times toward x_i (the rule of the self
//     integrals, core/bem_quadrature.py:217-308).
//
// Direct sums as in the reference, on the GPU: target blocks x source ranges
// (split so the GPU is full for any M, N), partial sums combined in a fixed
// order.  The near correction finds its panels with the near-pair cell grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ed_internal.h"
#include "ed_near.h"

namespace ed {

constexpr int kTgtBlock = 256;
constexpr int kTgtMaxSplit = 32;

struct TargetItem {
  int64_t t0;
  int32_t nt;
  int32_t split;
  int64_t s0, s1;
};

template <typename T>
__device__ __forceinline__ T tgt_rsqrt(T r2);
template <>
__device__ __forceinline__ double tgt_rsqrt<double>(double r2) {
  double y = (double)__builtin_amdgcn_rsqf((float)r2);
  y = y * (1.5 - 0.5 * r2 * y * y);
  return y * (1.5 - 0.5 * r2 * y * y);
}
template <>
__device__ __forceinline__ float tgt_rsqrt<float>(float r2) {
  return __builtin_amdgcn_rsqf(r2);
}

// partial layout: [split][component][m], component 0 = V, 1..3 = E
template <typename T, bool WV, bool WE>
__global__ __launch_bounds__(kTgtBlock) void k_targets_part(const TargetItem* __restrict__ items,
                                                            const T* __restrict__ X,
                                                            const T* __restrict__ C,
                                                            const T* __restrict__ A,
                                                            const T* __restrict__ sigma,
                                                            int64_t m, T* __restrict__ part) {
  __shared__ T sx[kTgtBlock], sy[kTgtBlock], sz[kTgtBlock], sq[kTgtBlock];
  const TargetItem it = items[blockIdx.x];
  const int64_t i = it.t0 + threadIdx.x;
  const bool act = threadIdx.x < it.nt;
  const int64_t ic = act ? i : it.t0;
  const T xi = X[3 * ic], yi = X[3 * ic + 1], zi = X[3 * ic + 2];
  T v = 0, ex = 0, ey = 0, ez = 0;
  for (int64_t j0 = it.s0; j0 < it.s1; j0 += kTgtBlock) {
    const int64_t j = j0 + threadIdx.x;
    __syncthreads();
    if (j < it.s1) {
      sx[threadIdx.x] = C[3 * j];
      sy[threadIdx.x] = C[3 * j + 1];
      sz[threadIdx.x] = C[3 * j + 2];
      sq[threadIdx.x] = sigma[j] * A[j];
    } else {
      sx[threadIdx.x] = sy[threadIdx.x] = sz[threadIdx.x] = (T)1e30;
      sq[threadIdx.x] = 0;
    }
    __syncthreads();
    const int nt = (int)std::min<int64_t>(kTgtBlock, it.s1 - j0);
    for (int t = 0; t < nt; ++t) {
      const T dx = xi - sx[t], dy = yi - sy[t], dz = zi - sz[t];
      T r2 = dx * dx + dy * dy + dz * dz;
      r2 = r2 > (T)1e-24 ? r2 : (T)1e-24;
      const T ir = tgt_rsqrt<T>(r2);
      if (WV) v += sq[t] * ir;
      if (WE) {
        const T w = sq[t] * ir * ir * ir;
        ex += w * dx;
        ey += w * dy;
        ez += w * dz;
      }
    }
  }
  if (!act) return;
  T* p = part + (size_t)it.split * 4 * m;
  if (WV) p[i] = v;
  if (WE) {
    p[m + i] = ex;
    p[2 * m + i] = ey;
    p[3 * m + i] = ez;
  }
}

template <typename T, bool WV, bool WE>
__global__ void k_targets_finish(int64_t m, int nsplit, const T* __restrict__ part, T ke,
                                 T* __restrict__ V, T* __restrict__ E) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  T v = 0, ex = 0, ey = 0, ez = 0;
  for (int s = 0; s < nsplit; ++s) {
    const T* p = part + (size_t)s * 4 * m;
    if (WV) v += p[i];
    if (WE) { ex += p[m + i]; ey += p[2 * m + i]; ez += p[3 * m + i]; }
  }
  // non-finite -> 0 (core/bem_kernel.py:897-904, :987-994)
  if (WV) { v *= ke; V[i] = isfinite(v) ? v : (T)0; }
  if (WE) {
    ex *= ke; ey *= ke; ez *= ke;
    E[3 * i] = isfinite(ex) ? ex : (T)0;
    E[3 * i + 1] = isfinite(ey) ? ey : (T)0;
    E[3 * i + 2] = isfinite(ez) ? ez : (T)0;
  }
}

// ---------------------------------------------------------------------------
// near-singular target correction
// ---------------------------------------------------------------------------
__device__ __forceinline__ double tgt_tri_area(const double* v) {
#pragma clang fp contract(off)
  const double e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
  const double e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
  const double cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
  return 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

// order-1/2 rule on one triangle: sum K_E w / max(|x - y|, 1e-12)
__device__ __forceinline__ double tgt_rule(const double* v, int order, double tx, double ty,
                                           double tz) {
#pragma clang fp contract(off)
  const double area = tgt_tri_area(v);
  if (area <= 0.0) return 0.0;
  double acc = 0.0;
  const int nq = order <= 1 ? 1 : 3;
  for (int q = 0; q < nq; ++q) {
    double b0, b1, b2, w;
    if (order <= 1) { b0 = b1 = b2 = 1.0 / 3.0; w = 1.0 * area; }
    else {
      b0 = q == 0 ? 2.0 / 3.0 : 1.0 / 6.0;
      b1 = q == 1 ? 2.0 / 3.0 : 1.0 / 6.0;
      b2 = q == 2 ? 2.0 / 3.0 : 1.0 / 6.0;
      w = (1.0 / 3.0) * area;
    }
    const double px = b0 * v[0] + b1 * v[3] + b2 * v[6];
    const double py = b0 * v[1] + b1 * v[4] + b2 * v[7];
    const double pz = b0 * v[2] + b1 * v[5] + b2 * v[8];
    const double dx = tx - px, dy = ty - py, dz = tz - pz;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    r = r > 1e-12 ? r : 1e-12;
    acc += (kKE / r) * w;
  }
  return acc;
}

constexpr int kTgtMaxDepth = 6;

// panel integral refined `depth` times toward the target: the sub-triangle
// whose centroid is nearest (first on ties) is split 4-way and appended
__device__ double tgt_refined(const double* pv, int depth, int order, double tx, double ty,
                              double tz) {
#pragma clang fp contract(off)
  double tris[1 + 3 * kTgtMaxDepth][9];
  int cnt = 1;
  for (int k = 0; k < 9; ++k) tris[0][k] = pv[k];
  for (int d = 0; d < depth; ++d) {
    int best = 0;
    double bd = 0.0;
    for (int t = 0; t < cnt; ++t) {
      const double* v = tris[t];
      const double cx = (v[0] + v[3] + v[6]) / 3.0, cy = (v[1] + v[4] + v[7]) / 3.0,
                   cz = (v[2] + v[5] + v[8]) / 3.0;
      const double dx = cx - tx, dy = cy - ty, dz = cz - tz;
      const double dd = sqrt(dx * dx + dy * dy + dz * dz);
      if (t == 0 || dd < bd) { bd = dd; best = t; }
    }
    double sel[9];
    for (int k = 0; k < 9; ++k) sel[k] = tris[best][k];
    for (int t = best; t + 1 < cnt; ++t)
      for (int k = 0; k < 9; ++k) tris[t][k] = tris[t + 1][k];
    --cnt;
    double m01[3], m12[3], m20[3];
    for (int k = 0; k < 3; ++k) {
      m01[k] = 0.5 * (sel[k] + sel[3 + k]);
      m12[k] = 0.5 * (sel[3 + k] + sel[6 + k]);
      m20[k] = 0.5 * (sel[6 + k] + sel[k]);
    }
    for (int k = 0; k < 3; ++k) {
      tris[cnt][k] = sel[k];        tris[cnt][3 + k] = m01[k];       tris[cnt][6 + k] = m20[k];
      tris[cnt + 1][k] = m01[k];    tris[cnt + 1][3 + k] = sel[3 + k]; tris[cnt + 1][6 + k] = m12[k];
      tris[cnt + 2][k] = m20[k];    tris[cnt + 2][3 + k] = m12[k];   tris[cnt + 2][6 + k] = sel[6 + k];
      tris[cnt + 3][k] = m01[k];    tris[cnt + 3][3 + k] = m12[k];   tris[cnt + 3][6 + k] = m20[k];
    }
    cnt += 4;
  }
  double acc = 0.0;
  for (int t = 0; t < cnt; ++t) acc += tgt_rule(tris[t], order, tx, ty, tz);
  return acc;
}

__device__ __forceinline__ int tgt_cell(double x, double lo, double inv_h, int dim) {
  int c = (int)floor((x - lo) * inv_h);
  return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

// One thread per target; panels from the 27 cells around the target's cell
// of a grid whose edge is at least f R_max.
__global__ void k_targets_near(int64_t m, const double* __restrict__ X, const double* __restrict__ C,
                               const double* __restrict__ A, const double* __restrict__ PV,
                               const double* __restrict__ sigma, double df, int depth, int order,
                               NearGrid g, const int32_t* __restrict__ cell_start,
                               const int32_t* __restrict__ cell_pts, double* __restrict__ V) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double tx = X[3 * i], ty = X[3 * i + 1], tz = X[3 * i + 2];
  if (!isfinite(tx) || !isfinite(ty) || !isfinite(tz)) return;
  const int cx = tgt_cell(tx, g.lo[0], g.inv_h, g.dim[0]);
  const int cy = tgt_cell(ty, g.lo[1], g.inv_h, g.dim[1]);
  const int cz = tgt_cell(tz, g.lo[2], g.inv_h, g.dim[2]);
  double corr = 0.0;
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= g.dim[2]) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = cy + dy;
      if (y < 0 || y >= g.dim[1]) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = cx + dx;
        if (x < 0 || x >= g.dim[0]) continue;
        const int64_t cell = ((int64_t)z * g.dim[1] + y) * g.dim[0] + x;
        for (int32_t q = cell_start[cell]; q < cell_start[cell + 1]; ++q) {
          const int64_t j = cell_pts[q];
          const double ax = C[3 * j] - tx, ay = C[3 * j + 1] - ty, az = C[3 * j + 2] - tz;
          const double d = sqrt(ax * ax + ay * ay + az * az);
          double a = A[j];
          a = a < 0.0 ? 0.0 : a;
          const double th = df * sqrt(a / 3.141592653589793);
          if (!(d <= th)) continue;
          const double aj = A[j], sj = sigma[j];
          if (aj <= 0.0 || sj == 0.0) continue;
          const double inear = tgt_refined(PV + 9 * j, depth, order, tx, ty, tz);
          const double ifar = kKE * aj / (d > 1e-12 ? d : 1e-12);
          corr += sj * (inear - ifar);
        }
      }
    }
  }
  V[i] += corr;
}

template <typename T>
static int targets_run(const T* X, int64_t m, const T* C, const T* A, const T* sigma, int64_t n,
                       T* V, T* E, hipStream_t st) {
  if (m < 0 || n < 0) return fail(ED_ERR_ARG, "targets: bad sizes");
  if (m == 0) return ED_OK;
  if (!X || (!V && !E)) return fail(ED_ERR_ARG, "targets: NULL pointer");
  if (n == 0) {
    if (V) (void)hipMemsetAsync(V, 0, sizeof(T) * m, st);
    if (E) (void)hipMemsetAsync(E, 0, sizeof(T) * 3 * m, st);
    return ED_OK;
  }
  if (!C || !A || !sigma) return fail(ED_ERR_ARG, "targets: NULL source pointer");
  // target blocks x source splits: aim at >= 2048 workgroups
  const int64_t tb = (m + kTgtBlock - 1) / kTgtBlock;
  const int S = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)kTgtMaxSplit,
                                                              (2048 + tb - 1) / tb,
                                                              (n + 1023) / 1024}));
  const int64_t chunk = (n + S - 1) / S;
  std::vector<TargetItem> items;
  for (int64_t t0 = 0; t0 < m; t0 += kTgtBlock)
    for (int s = 0; s < S; ++s) {
      TargetItem it;
      it.t0 = t0;
      it.nt = (int32_t)std::min<int64_t>(kTgtBlock, m - t0);
      it.split = s;
      it.s0 = std::min<int64_t>(n, s * chunk);
      it.s1 = std::min<int64_t>(n, (s + 1) * chunk);
      items.push_back(it);
    }
  TargetItem* d_items = nullptr;
  T* part = nullptr;
  int rc = ED_OK;
  if (hipMalloc((void**)&d_items, sizeof(TargetItem) * items.size()) != hipSuccess ||
      hipMalloc((void**)&part, sizeof(T) * 4 * (size_t)S * m) != hipSuccess)
    rc = fail(ED_ERR_ALLOC, "targets: alloc");
  hipError_t e = hipSuccess;
  if (rc == ED_OK) {
    e = hipMemcpyAsync(d_items, items.data(), sizeof(TargetItem) * items.size(),
                       hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      const dim3 gi((unsigned)items.size()), gf((unsigned)((m + 255) / 256));
      if (V && E) {
        hipLaunchKernelGGL((k_targets_part<T, true, true>), gi, dim3(kTgtBlock), 0, st, d_items, X, C, A, sigma, m, part);
        hipLaunchKernelGGL((k_targets_finish<T, true, true>), gf, dim3(256), 0, st, m, S, (const T*)part, (T)kKE, V, E);
      } else if (V) {
        hipLaunchKernelGGL((k_targets_part<T, true, false>), gi, dim3(kTgtBlock), 0, st, d_items, X, C, A, sigma, m, part);
        hipLaunchKernelGGL((k_targets_finish<T, true, false>), gf, dim3(256), 0, st, m, S, (const T*)part, (T)kKE, V, E);
      } else {
        hipLaunchKernelGGL((k_targets_part<T, false, true>), gi, dim3(kTgtBlock), 0, st, d_items, X, C, A, sigma, m, part);
        hipLaunchKernelGGL((k_targets_finish<T, false, true>), gf, dim3(256), 0, st, m, S, (const T*)part, (T)kKE, V, E);
      }
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail(ED_ERR_LAUNCH, std::string("targets: ") + hipGetErrorString(e));
  }
  if (d_items) (void)hipFree(d_items);
  if (part) (void)hipFree(part);
  return rc;
}

}  // namespace ed

using namespace ed;

extern "C" int ed_targets_f64(const double* targets, int64_t m, const double* C, const double* A,
                              const double* sigma, int64_t n, double* V, double* E,
                              ed_stream stream) {
  return targets_run<double>(targets, m, C, A, sigma, n, V, E, (hipStream_t)stream);
}

extern "C" int ed_targets_f32(const float* targets, int64_t m, const float* C, const float* A,
                              const float* sigma, int64_t n, float* V, float* E,
                              ed_stream stream) {
  return targets_run<float>(targets, m, C, A, sigma, n, V, E, (hipStream_t)stream);
}

extern "C" int ed_near_targets_f64(const double* targets, int64_t m, const double* C,
                                   const double* A, const double* PV, const double* sigma,
                                   int64_t n, double factor, int32_t depth, int32_t order,
                                   double* V, ed_stream stream) {
  if (m < 0 || n < 0 || depth < 0 || depth > kTgtMaxDepth)
    return fail(ED_ERR_ARG, "near targets: bad sizes or depth (0..6)");
  if (m == 0 || n == 0) return ED_OK;
  if (!targets || !C || !A || !PV || !sigma || !V) return fail(ED_ERR_ARG, "near targets: NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  // the pair grid's cell edge is >= 2 * df' * R_max; df' = factor / 2 gives
  // an edge >= factor * R_max, the largest target-panel cutoff
  NearSearch ns;
  int rc = ns.build(C, A, n, 0.5 * factor, st);
  if (rc != ED_OK) return rc;
  if (!ns.any) {
    // fewer than two usable panels or a non-positive factor: one cell holding
    // every panel (the kernel's own test then decides, as the reference's)
    std::vector<int32_t> start = {0, (int32_t)n}, pts((size_t)n);
    for (int64_t j = 0; j < n; ++j) pts[(size_t)j] = (int32_t)j;
    NearGrid g{};
    g.inv_h = 0.0;
    g.dim[0] = g.dim[1] = g.dim[2] = 1;
    int32_t *ds = nullptr, *dp = nullptr;
    if (hipMalloc((void**)&ds, sizeof(int32_t) * 2) != hipSuccess ||
        hipMalloc((void**)&dp, sizeof(int32_t) * n) != hipSuccess) {
      if (ds) (void)hipFree(ds);
      return fail(ED_ERR_ALLOC, "near targets: alloc");
    }
    hipError_t e = hipMemcpy(ds, start.data(), sizeof(int32_t) * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, pts.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_targets_near, dim3((unsigned)((m + 127) / 128)), dim3(128), 0, st, m,
                         targets, C, A, PV, sigma, factor, depth, order, g, (const int32_t*)ds,
                         (const int32_t*)dp, V);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(ds);
    (void)hipFree(dp);
    if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("near targets: ") + hipGetErrorString(e));
    return ED_OK;
  }
  hipLaunchKernelGGL(k_targets_near, dim3((unsigned)((m + 127) / 128)), dim3(128), 0, st, m,
                     targets, C, A, PV, sigma, factor, depth, order, ns.g,
                     (const int32_t*)ns.d_start, (const int32_t*)ns.d_pts, V);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("near targets: ") + hipGetErrorString(e));
  return ED_OK;
}
