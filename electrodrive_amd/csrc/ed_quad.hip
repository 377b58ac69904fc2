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

#include <algorithm>
#include <cmath>
#include <vector>

#include "ed_internal.h"
#include "ed_near.h"

namespace ed {

#define ED_TRY(call)               \
  do {                             \
    const int rc_ = (call);        \
    if (rc_ != ED_OK) return rc_;  \
  } while (0)

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

// Near pairs.  Candidates come from a uniform cell grid built on the host
// (cell edge h >= the largest possible cutoff 2 * df * R_max, so a partner j
// of i always lies in one of the 27 cells around i's cell); the acceptance
// test itself is the reference's, evaluated with the same operations as a
// brute-force scan.  Phase 0 counts per row, phase 1 fills at the row offsets
// and sorts each row by j, so the output equals the i-major, j-ascending scan.
__device__ __forceinline__ int near_cell_coord(double x, double lo, double inv_h, int dim) {
  int c = (int)floor((x - lo) * inv_h);
  return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

// One wave per point i: the lanes share each candidate cell (64 candidates
// per step), accepted partners are counted with a ballot and, when filling,
// written at the row offset plus their rank among the accepted lanes.  The
// row is then put in ascending j: rows of up to kNearSortCap entries are
// collected in LDS and each lane places its entries by rank (j is unique in a
// row); longer rows fall back to an insertion sort by one lane.
constexpr int kNearWaves = 4;      // points per block
constexpr int kNearSortCap = 256;  // row entries sorted in LDS
__global__ __launch_bounds__(64 * kNearWaves) void k_near_pairs(const double* __restrict__ C,
                                                               const double* __restrict__ A,
                                                               int64_t n, double df, NearGrid g,
                                                               const int32_t* __restrict__ cell_start,
                                                               const int32_t* __restrict__ cell_pts,
                                                               int64_t* __restrict__ counts,
                                                               const int64_t* __restrict__ offs,
                                                               int64_t* __restrict__ pairs,
                                                               int32_t* __restrict__ cols32,
                                                               int64_t col_off) {
#pragma clang fp contract(off)
  __shared__ int32_t srt[kNearWaves][kNearSortCap];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * kNearWaves + wv;
  const bool valid = i < n;  // wave-uniform; every wave reaches the barrier
  const bool fill = pairs != nullptr || cols32 != nullptr;
  int64_t cnt = 0;  // wave-uniform
  int64_t w0 = 0;
  if (valid) {
    const double xi = C[3 * i], yi = C[3 * i + 1], zi = C[3 * i + 2];
    const double ri = near_radius(A[i]);
    w0 = offs ? offs[i] : 0;
    if (isfinite(xi) && isfinite(yi) && isfinite(zi) && isfinite(ri)) {
      const int cx = near_cell_coord(xi, g.lo[0], g.inv_h, g.dim[0]);
      const int cy = near_cell_coord(yi, g.lo[1], g.inv_h, g.dim[1]);
      const int cz = near_cell_coord(zi, g.lo[2], g.inv_h, g.dim[2]);
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
            const int32_t cs = cell_start[cell], ce = cell_start[cell + 1];
            for (int32_t q0 = cs; q0 < ce; q0 += 64) {
              const int32_t q = q0 + lane;
              bool ok = false;
              int64_t j = 0;
              if (q < ce) {
                j = cell_pts[q];
                ok = near_accept(xi, yi, zi, ri, C[3 * j], C[3 * j + 1], C[3 * j + 2], A[j], df);
              }
              const uint64_t m = __ballot(ok);
              if (ok && fill) {
                const int64_t pos = cnt + __popcll(m & ((1ull << lane) - 1));
                if (pos < kNearSortCap) {
                  srt[wv][pos] = (int32_t)j;
                } else {
                  if (pairs) {
                    pairs[2 * (w0 + pos)] = i;
                    pairs[2 * (w0 + pos) + 1] = j;
                  }
                  if (cols32) cols32[w0 + pos] = (int32_t)(j + col_off);
                }
              }
              cnt += __popcll(m);
            }
          }
        }
      }
    }
    if (counts && lane == 0) counts[i] = cnt;
  }
  if (!fill) return;  // uniform over the block
  __syncthreads();
  if (!valid || cnt == 0) return;
  if (cnt <= kNearSortCap) {
    for (int e = lane; e < cnt; e += 64) {
      const int32_t v = srt[wv][e];
      int r = 0;
      for (int k = 0; k < cnt; ++k) r += srt[wv][k] < v ? 1 : 0;
      if (pairs) {
        pairs[2 * (w0 + r)] = i;
        pairs[2 * (w0 + r) + 1] = v;
      }
      if (cols32) cols32[w0 + r] = (int32_t)(v + col_off);
    }
    return;
  }
  // long row: the LDS part joins the rest in global memory, then one lane
  // sorts the whole row
  for (int e = lane; e < kNearSortCap; e += 64) {
    const int32_t v = srt[wv][e];
    if (pairs) {
      pairs[2 * (w0 + e)] = i;
      pairs[2 * (w0 + e) + 1] = v;
    }
    if (cols32) cols32[w0 + e] = (int32_t)(v + col_off);
  }
  __threadfence_block();
  if (lane != 0) return;
  if (pairs) {
    int64_t* row = pairs + 2 * w0;
    for (int64_t a = 1; a < cnt; ++a) {
      const int64_t v = row[2 * a + 1];
      int64_t b = a - 1;
      while (b >= 0 && row[2 * b + 1] > v) {
        row[2 * (b + 1) + 1] = row[2 * b + 1];
        --b;
      }
      row[2 * (b + 1) + 1] = v;
    }
  }
  if (cols32) {
    int32_t* row = cols32 + w0;
    for (int64_t a = 1; a < cnt; ++a) {
      const int32_t v = row[a];
      int64_t b = a - 1;
      while (b >= 0 && row[b] > v) {
        row[b + 1] = row[b];
        --b;
      }
      row[b + 1] = v;
    }
  }
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

// Host cell grid of the near-pair search.  Points with a non-finite
// coordinate or radius can have no partner and are left out.  Returns false
// when no pair can exist (df or every radius zero / non-finite).
static bool near_grid_build(const double* C, const double* A, int64_t n, double df, NearGrid& g,
                            std::vector<int32_t>& start, std::vector<int32_t>& pts) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double rmax = 0.0;
  std::vector<char> ok((size_t)n, 0);
  int64_t n_ok = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double a = A[i] < 0.0 ? 0.0 : A[i];
    const double r = std::sqrt(a / 3.141592653589793);
    const double* c = C + 3 * i;
    if (!std::isfinite(c[0]) || !std::isfinite(c[1]) || !std::isfinite(c[2]) || !std::isfinite(r))
      continue;
    ok[i] = 1;
    ++n_ok;
    rmax = std::max(rmax, r);
    for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], c[d]); hi[d] = std::max(hi[d], c[d]); }
  }
  const double cut = 2.0 * df * rmax;
  if (n_ok < 2 || !(cut > 0.0) || !std::isfinite(cut)) return false;
  // cell edge: the cutoff with a relative margin for the rounding of the
  // cell coordinates, enlarged until the grid holds at most max(4n, 4096) cells
  double h = cut * (1.0 + 1e-6) + 1e-300;
  const double max_cells = std::max<double>(4.0 * (double)n, 4096.0);
  for (;;) {
    double cells = 1.0;
    for (int d = 0; d < 3; ++d) cells *= std::floor((hi[d] - lo[d]) / h) + 1.0;
    if (cells <= max_cells) break;
    h *= std::max(1.01, std::cbrt(cells / max_cells));
  }
  for (int d = 0; d < 3; ++d) {
    g.lo[d] = lo[d];
    g.dim[d] = (int32_t)(std::floor((hi[d] - lo[d]) / h) + 1.0);
  }
  g.inv_h = 1.0 / h;
  const int64_t ncell = (int64_t)g.dim[0] * g.dim[1] * g.dim[2];
  auto coord = [&](double x, int d) {
    int c = (int)std::floor((x - g.lo[d]) * g.inv_h);
    return c < 0 ? 0 : (c >= g.dim[d] ? g.dim[d] - 1 : c);
  };
  std::vector<int32_t> cell((size_t)n, -1);
  start.assign((size_t)ncell + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (!ok[i]) continue;
    const double* c = C + 3 * i;
    const int64_t id = ((int64_t)coord(c[2], 2) * g.dim[1] + coord(c[1], 1)) * g.dim[0] + coord(c[0], 0);
    cell[i] = (int32_t)id;
    ++start[id + 1];
  }
  for (int64_t k = 0; k < ncell; ++k) start[k + 1] += start[k];
  pts.assign((size_t)n_ok, 0);
  std::vector<int32_t> cur(start.begin(), start.end() - 1);
  for (int64_t i = 0; i < n; ++i)
    if (cell[i] >= 0) pts[cur[cell[i]]++] = (int32_t)i;
  return true;
}

NearSearch::~NearSearch() {
  if (d_start) (void)hipFree(d_start);
  if (d_pts) (void)hipFree(d_pts);
}

int NearSearch::build(const double* Cd, const double* Ad, int64_t n_, double df_, hipStream_t st) {
  C = Cd;
  A = Ad;
  n = n_;
  df = df_;
  any = false;
  if (n > (int64_t)INT32_MAX) return fail(ED_ERR_ARG, "near pairs: n exceeds int32 range");
  if (n <= 1) return ED_OK;
  std::vector<double> hc((size_t)n * 3), ha((size_t)n);
  hipError_t e = hipMemcpyAsync(hc.data(), C, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(ha.data(), A, sizeof(double) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(ED_ERR_DEVICE, std::string("near pairs: ") + hipGetErrorString(e));
  std::vector<int32_t> start, pts;
  if (!near_grid_build(hc.data(), ha.data(), n, df, g, start, pts)) return ED_OK;
  const int64_t ncell = (int64_t)start.size() - 1;
  if (hipMalloc((void**)&d_start, sizeof(int32_t) * (ncell + 1)) != hipSuccess ||
      hipMalloc((void**)&d_pts, sizeof(int32_t) * std::max<size_t>(pts.size(), 1)) != hipSuccess)
    return fail(ED_ERR_ALLOC, "near pairs: alloc");
  e = hipMemcpyAsync(d_start, start.data(), sizeof(int32_t) * (ncell + 1), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !pts.empty())
    e = hipMemcpyAsync(d_pts, pts.data(), sizeof(int32_t) * pts.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // host vectors go out of scope
  if (e != hipSuccess) return fail(ED_ERR_DEVICE, std::string("near pairs: ") + hipGetErrorString(e));
  any = true;
  return ED_OK;
}

int NearSearch::count(hipStream_t st, std::vector<int64_t>& ptr, double df_use) {
  ptr.assign((size_t)std::max<int64_t>(n, 0) + 1, 0);
  if (!any) return ED_OK;
  const double dfu = df_use > 0.0 ? std::min(df_use, df) : df;
  int64_t* counts = nullptr;
  if (hipMalloc((void**)&counts, sizeof(int64_t) * n) != hipSuccess)
    return fail(ED_ERR_ALLOC, "near pairs: alloc");
  const unsigned grid = (unsigned)((n + kNearWaves - 1) / kNearWaves);
  hipLaunchKernelGGL(k_near_pairs, dim3(grid), dim3(64 * kNearWaves), 0, st, C, A, n, dfu, g,
                     (const int32_t*)d_start, (const int32_t*)d_pts, counts,
                     (const int64_t*)nullptr, (int64_t*)nullptr, (int32_t*)nullptr, (int64_t)0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(ptr.data() + 1, counts, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(counts);
  if (e != hipSuccess) return fail(ED_ERR_DEVICE, std::string("near pairs: ") + hipGetErrorString(e));
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  return ED_OK;
}

int NearSearch::fill(hipStream_t st, const std::vector<int64_t>& ptr, int64_t* pairs,
                     int32_t* cols32, int64_t col_off, double df_use) {
  if (!any || ptr[n] == 0) return ED_OK;
  const double dfu = df_use > 0.0 ? std::min(df_use, df) : df;
  int64_t* offs = nullptr;
  if (hipMalloc((void**)&offs, sizeof(int64_t) * n) != hipSuccess)
    return fail(ED_ERR_ALLOC, "near pairs: alloc");
  hipError_t e = hipMemcpyAsync(offs, ptr.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    const unsigned grid = (unsigned)((n + kNearWaves - 1) / kNearWaves);
    hipLaunchKernelGGL(k_near_pairs, dim3(grid), dim3(64 * kNearWaves), 0, st, C, A, n, dfu, g,
                       (const int32_t*)d_start, (const int32_t*)d_pts, (int64_t*)nullptr,
                       (const int64_t*)offs, pairs, cols32, col_off);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(offs);
  if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("near pairs: ") + hipGetErrorString(e));
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
  NearSearch ns;
  std::vector<int64_t> ptr;
  ED_TRY(ns.build(C, A, n, df, st));
  ED_TRY(ns.count(st, ptr));
  const int64_t total = ptr[n];
  *n_pairs_out = total;
  // too small a buffer only reports the count (the caller retries)
  if (!pairs || total == 0 || capacity < total) return ED_OK;
  return ns.fill(st, ptr, pairs, nullptr, 0);
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
