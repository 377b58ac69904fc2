// Sparse approximate inverse (SAI) smoother of the native solver (no
// reference counterpart; the reference offers Jacobi and mass scaling only,
// core/bem_solver.py:36-117).
//
// Left SAI: row i of M minimises |m^T A[J, I] - e_i^T| over m supported on the
// pattern J(i) = {i} + panels within factor_pattern (R_i + R_j) of i, fitted
// on the rows I(i) = {i} + panels within factor_rows (R_i + R_j).  A is the
// collocation operator: K_E A_j / r_ij off the diagonal plus the near-field
// quadrature correction, self_integral_i A_i on it.  Everything runs in tree
// order; patterns never cross geometries of a batched plan.
//
// Layout: column-major ELL with kSaiJ slots per row (col[q * n + i],
// val[q * n + i], cnt[i]), so consecutive rows read consecutive addresses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ed_internal.h"
#include "ed_near.h"
#include "ed_plan.h"

namespace ed {

constexpr int kSaiJ = 32;   // pattern slots per row
constexpr int kSaiI = 128;  // fitting rows per row
constexpr int kSaiCand = 512;

__device__ __forceinline__ double sai_dist2(const double* C, int64_t a, int64_t b) {
  const double dx = C[3 * a] - C[3 * b], dy = C[3 * a + 1] - C[3 * b + 1], dz = C[3 * a + 2] - C[3 * b + 2];
  return dx * dx + dy * dy + dz * dz;
}

constexpr int kSaiThreads = 128;
constexpr int kSaiHash = 256;  // LDS hash slots for the fitting rows (>= 2 kSaiI)

// list[0] = i, then the (cap - 1) nearest candidates of the CSR row (ties by
// index; rows beyond kSaiCand candidates keep their first kSaiCand).  The
// candidates' distances are computed once into LDS (dsh, csh), then each
// candidate's rank is counted there.
__device__ void sai_select(int64_t i, const int32_t* __restrict__ col, int64_t beg, int64_t end,
                           const double* __restrict__ C, int cap, int* list, int* count,
                           double* dsh, int* csh) {
  const int c = (int)std::min<int64_t>(end - beg, kSaiCand);
  if (threadIdx.x == 0) list[0] = (int)i;
  if (c <= cap - 1) {
    for (int t = threadIdx.x; t < c; t += blockDim.x) list[1 + t] = col[beg + t];
    if (threadIdx.x == 0) *count = 1 + c;
  } else {
    for (int t = threadIdx.x; t < c; t += blockDim.x) {
      const int ct = col[beg + t];
      csh[t] = ct;
      dsh[t] = sai_dist2(C, i, ct);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < c; t += blockDim.x) {
      const int ct = csh[t];
      const double dt = dsh[t];
      int rank = 0;
      for (int u = 0; u < c; ++u) {
        const double du = dsh[u];
        rank += (du < dt || (du == dt && csh[u] < ct)) ? 1 : 0;
      }
      if (rank < cap - 1) list[1 + rank] = ct;
    }
    if (threadIdx.x == 0) *count = cap;
  }
  __syncthreads();
}

__device__ __forceinline__ int sai_hash(int key) {
  return (int)(((uint32_t)key * 2654435761u) >> 24) & (kSaiHash - 1);
}

// One 128-thread block per row (rows row0 + blockIdx.x of a geometry segment).
__global__ __launch_bounds__(kSaiThreads) void k_sai_build(
    int64_t n, int64_t row0, double factor_pattern,
    const int64_t* __restrict__ iptr, const int32_t* __restrict__ icol,
    const double* __restrict__ C, const double* __restrict__ A, const double* __restrict__ selfi,
    const int32_t* __restrict__ near_ptr, const int32_t* __restrict__ near_col,
    const double* __restrict__ near_w, double ridge, int32_t* __restrict__ out_col,
    double* __restrict__ out_val, int32_t* __restrict__ out_cnt) {
  __shared__ int Jl[kSaiJ];
  __shared__ int Il[kSaiI];
  // B holds the fitting rows of the row (at most ni_cap of them): sized at
  // launch from the longest fitting-row list, so short lists leave LDS for
  // more blocks per CU
  extern __shared__ double Bdyn[];
  auto B = reinterpret_cast<double(*)[kSaiJ + 1]>(Bdyn);
  __shared__ double G[kSaiJ][kSaiJ + 1];
  __shared__ double h[kSaiJ];
  __shared__ double dsh[kSaiCand];
  __shared__ int csh[kSaiCand];
  __shared__ int hk[kSaiHash], hv[kSaiHash];
  __shared__ int jc[kSaiCand];  // pattern candidates, ascending
  __shared__ int wcnt[kSaiThreads / 64];
  __shared__ int nJ, nI;
  const int tid = threadIdx.x;
  const int64_t loc = blockIdx.x;
  const int64_t i = row0 + loc;
  // Pattern candidates: the fitting-row candidates (factor_rows >= the
  // pattern factor) that pass the pattern's near-pair test, kept in the
  // row's ascending order by a ballot compaction (the first kSaiCand, as the
  // pattern's own search would give).
  {
    const int64_t ib = iptr[loc], ie = iptr[loc + 1];
    const double xi = C[3 * i], yi = C[3 * i + 1], zi = C[3 * i + 2];
    const double ri = near_radius(A[i]);
    const int lane = tid & 63, w = tid >> 6;
    int base = 0;
    for (int64_t q0 = ib; q0 < ie; q0 += kSaiThreads) {
      const int64_t q = q0 + tid;
      int j = 0;
      bool ok = false;
      if (q < ie) {
        j = icol[q];
        ok = near_accept(xi, yi, zi, ri, C[3 * (int64_t)j], C[3 * (int64_t)j + 1],
                         C[3 * (int64_t)j + 2], A[j], factor_pattern);
      }
      const uint64_t m = __ballot(ok);
      if (lane == 0) wcnt[w] = __popcll(m);
      __syncthreads();
      int pos = base + __popcll(m & ((1ull << lane) - 1));
      for (int u = 0; u < w; ++u) pos += wcnt[u];
      if (ok && pos < kSaiCand) jc[pos] = j;
      for (int u = 0; u < kSaiThreads / 64; ++u) base += wcnt[u];
      __syncthreads();
    }
    if (tid == 0) nJ = base;
    __syncthreads();
  }
  // every thread reads the candidate count before sai_select's thread 0
  // overwrites nJ with the pattern size
  const int n_cand = std::min(nJ, kSaiCand);
  __syncthreads();
  sai_select(i, jc, 0, n_cand, C, kSaiJ, Jl, &nJ, dsh, csh);
  sai_select(i, icol, iptr[loc], iptr[loc + 1], C, kSaiI, Il, &nI, dsh, csh);
  const int nj = nJ, ni = nI;
  // B[l][k] = A[J_k][I_l] (zero for k >= nj)
  for (int e = tid; e < ni * kSaiJ; e += blockDim.x) {
    const int l = e / kSaiJ, k = e - l * kSaiJ;
    double v = 0.0;
    if (k < nj) {
      const int a = Jl[k], b = Il[l];
      if (a == b) {
        v = selfi[a] * A[a];
      } else {
        const double r2 = fmax(sai_dist2(C, a, b), 1e-24);
        v = kKE * A[b] / sqrt(r2);
      }
    }
    B[l][k] = v;
  }
  // near-field corrections A[J_k][col] += w for the near entries of the rows
  // J_k: the fitting rows go into an LDS hash (panel -> position); thread k
  // adds the whole near row of J_k into column k of B, so entries that share
  // a (row, column) (duplicate pairs) are added by one thread in CSR order
  if (near_ptr) {
    for (int e = tid; e < kSaiHash; e += blockDim.x) hk[e] = -1;
    __syncthreads();
    for (int l = tid; l < ni; l += blockDim.x) {
      const int key = Il[l];
      int sl = sai_hash(key);
      while (true) {
        const int prev = atomicCAS(&hk[sl], -1, key);
        if (prev == -1) { hv[sl] = l; break; }
        sl = (sl + 1) & (kSaiHash - 1);
      }
    }
    __syncthreads();
    for (int k = tid; k < nj; k += blockDim.x) {
      const int q1 = near_ptr[Jl[k] + 1];
      for (int q = near_ptr[Jl[k]]; q < q1; ++q) {
        const int b = near_col[q];
        int sl = sai_hash(b);
        while (hk[sl] != -1 && hk[sl] != b) sl = (sl + 1) & (kSaiHash - 1);
        if (hk[sl] == b) B[hv[sl]][k] += near_w[q];
      }
    }
  }
  __syncthreads();
  // normal equations G = B^T B (2 x 4 register blocks per thread over the
  // full 32 x 32; entries beyond nj are zero), h = B^T e_i = B[0][:]
  {
    const int ra = (tid / 8) * 2, cb = (tid % 8) * 4;
    double acc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    for (int l = 0; l < ni; ++l) {
      const double a0 = B[l][ra], a1 = B[l][ra + 1];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double bv = B[l][cb + u];
        acc[0][u] += a0 * bv;
        acc[1][u] += a1 * bv;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      G[ra][cb + u] = acc[0][u];
      G[ra + 1][cb + u] = acc[1][u];
    }
  }
  for (int k = tid; k < nj; k += blockDim.x) h[k] = B[0][k];  // Il[0] == i
  __syncthreads();
  if (tid == 0) {
    double tr = 0.0;
    for (int a = 0; a < nj; ++a) tr += G[a][a];
    const double lam = ridge * (tr > 0.0 ? tr : 1.0) / nj;
    for (int a = 0; a < nj; ++a) G[a][a] += lam;
  }
  __syncthreads();
  // right-looking Cholesky G = L L^T (lower triangle, in place)
  for (int j = 0; j < nj; ++j) {
    if (tid == 0) {
      const double d = G[j][j];
      G[j][j] = d > 1e-300 ? sqrt(d) : 1e-150;
    }
    __syncthreads();
    const double d = G[j][j];
    for (int r = j + 1 + tid; r < nj; r += blockDim.x) G[r][j] /= d;
    __syncthreads();
    const int w = nj - j - 1;
    for (int e = tid; e < w * w; e += blockDim.x) {
      const int r = j + 1 + e / w, c = j + 1 + e % w;
      if (c <= r) G[r][c] -= G[r][j] * G[c][j];
    }
    __syncthreads();
  }
  // L y = h, then L^T m = y, column by column
  for (int r = 0; r < nj; ++r) {
    if (tid == 0) h[r] /= G[r][r];
    __syncthreads();
    for (int q = r + 1 + tid; q < nj; q += blockDim.x) h[q] -= G[q][r] * h[r];
    __syncthreads();
  }
  for (int r = nj - 1; r >= 0; --r) {
    if (tid == 0) h[r] /= G[r][r];
    __syncthreads();
    for (int q = tid; q < r; q += blockDim.x) h[q] -= G[r][q] * h[r];
    __syncthreads();
  }
  for (int q = tid; q < kSaiJ; q += blockDim.x) {
    const bool on = q < nj;
    out_col[(size_t)q * n + i] = on ? Jl[q] : (int32_t)i;
    out_val[(size_t)q * n + i] = on ? h[q] : 0.0;
  }
  if (tid == 0) out_cnt[i] = nj;
}

// z = M u (+ Z cy when k > 0)
__global__ void k_sai_apply(int64_t n, const int32_t* __restrict__ col,
                            const double* __restrict__ val, const int32_t* __restrict__ cnt,
                            const double* __restrict__ u, int k, const double* __restrict__ Z,
                            const double* __restrict__ cy, double* __restrict__ z) {
  __shared__ double c[32];
  if (threadIdx.x < k) c[threadIdx.x] = cy[threadIdx.x];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  const int m = cnt[i];
  for (int q = 0; q < m; ++q) s += val[(size_t)q * n + i] * u[col[(size_t)q * n + i]];
  for (int q = 0; q < k; ++q) s += c[q] * Z[(size_t)q * n + i];
  z[i] = s;
}

__global__ void k_sai_perm(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ src,
                           double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ void k_sai_unperm(int64_t n, const int32_t* __restrict__ perm,
                             const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}

void sai_apply_launch(ed_plan* P, const double* u, int k, const double* Z, const double* cy,
                      double* z, hipStream_t st) {
  const int64_t n = P->n;
  hipLaunchKernelGGL(k_sai_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     P->sai_col, P->sai_val, P->sai_cnt, u, k, Z, cy, z);
}

void sai_free(ed_plan* P) {
  if (P->sai_col) (void)hipFree(P->sai_col);
  if (P->sai_val) (void)hipFree(P->sai_val);
  if (P->sai_cnt) (void)hipFree(P->sai_cnt);
  P->sai_col = nullptr;
  P->sai_val = nullptr;
  P->sai_cnt = nullptr;
}

}  // namespace ed

using namespace ed;

extern "C" int ed_plan_build_sai(ed_plan* P, const double* self_integrals, double factor_pattern,
                                 double factor_rows, ed_stream stream) {
  if (!P || !self_integrals) return fail(ED_ERR_ARG, "ed_plan_build_sai: NULL argument");
  if (!(factor_pattern > 0.0) || !(factor_rows >= factor_pattern) || !std::isfinite(factor_rows))
    return fail(ED_ERR_ARG, "ed_plan_build_sai: need 0 < factor_pattern <= factor_rows < inf");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = P->n;
  gmres_invalidate_graph(P);  // a captured Arnoldi graph refers to the preconditioner arrays
  sai_free(P);
  double* selfi_t = nullptr;
  if (hipMalloc((void**)&P->sai_col, sizeof(int32_t) * kSaiJ * n) != hipSuccess ||
      hipMalloc((void**)&P->sai_val, sizeof(double) * kSaiJ * n) != hipSuccess ||
      hipMalloc((void**)&P->sai_cnt, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc((void**)&selfi_t, sizeof(double) * n) != hipSuccess) {
    sai_free(P);
    if (selfi_t) (void)hipFree(selfi_t);
    return fail(ED_ERR_ALLOC, "ed_plan_build_sai: alloc");
  }
  hipLaunchKernelGGL(k_sai_perm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, P->perm,
                     self_integrals, selfi_t);
  int rc = ED_OK;
  const auto& go = P->host.geom_pt_off;
  for (size_t gi = 0; gi + 1 < go.size() && rc == ED_OK; ++gi) {
    const int64_t b = go[gi], ng = go[gi + 1] - go[gi];
    if (ng <= 0) continue;
    // one search with the fitting-row factor; the kernel takes the pattern
    // candidates out of those rows
    NearSearch si;
    std::vector<int64_t> pi;
    int32_t* ci = nullptr;
    int64_t* dpi = nullptr;
    rc = si.build(P->cent64_t + 3 * b, P->area64_t + b, ng, factor_rows, st);
    if (rc == ED_OK) rc = si.count(st, pi);
    if (rc == ED_OK &&
        (hipMalloc((void**)&ci, sizeof(int32_t) * std::max<int64_t>(pi[ng], 1)) != hipSuccess ||
         hipMalloc((void**)&dpi, sizeof(int64_t) * (ng + 1)) != hipSuccess))
      rc = fail(ED_ERR_ALLOC, "ed_plan_build_sai: alloc");
    if (rc == ED_OK) rc = si.fill(st, pi, nullptr, ci, b);
    if (rc == ED_OK) {
      hipError_t e = hipMemcpyAsync(dpi, pi.data(), sizeof(int64_t) * (ng + 1), hipMemcpyHostToDevice, st);
      // the longest fitting-row list of the segment sizes B (row i + at most
      // kSaiI - 1 candidates)
      int64_t maxc = 0;
      for (int64_t k = 0; k < ng; ++k) maxc = std::max(maxc, pi[k + 1] - pi[k]);
      const int ni_cap = 1 + (int)std::min<int64_t>(maxc, kSaiI - 1);
      const size_t lds_b = sizeof(double) * (kSaiJ + 1) * (size_t)ni_cap;
      if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sai_build, dim3((unsigned)ng), dim3(kSaiThreads), lds_b, st, n, b,
                           factor_pattern, (const int64_t*)dpi, (const int32_t*)ci, P->cent64_t,
                           P->area64_t, (const double*)selfi_t,
                           P->n_near > 0 ? P->near_ptr : nullptr, P->near_col, P->near_w64,
                           1e-12, P->sai_col, P->sai_val, P->sai_cnt);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) rc = fail(ED_ERR_LAUNCH, std::string("ed_plan_build_sai: ") + hipGetErrorString(e));
    }
    if (ci) (void)hipFree(ci);
    if (dpi) (void)hipFree(dpi);
  }
  (void)hipFree(selfi_t);
  if (rc != ED_OK) sai_free(P);
  return rc;
}

// z = M_sai u in original panel order (device FP64); the smoother alone, for
// tests and for callers that run their own Krylov loop.
extern "C" int ed_plan_sai_apply(ed_plan* P, const double* u, double* z, ed_stream stream) {
  if (!P || !u || !z) return fail(ED_ERR_ARG, "ed_plan_sai_apply: NULL argument");
  if (!P->sai_cnt) return fail(ED_ERR_ARG, "ed_plan_sai_apply: sparse approximate inverse not built");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = P->n;
  double *ut = nullptr, *zt = nullptr;
  if (hipMalloc((void**)&ut, sizeof(double) * n) != hipSuccess ||
      hipMalloc((void**)&zt, sizeof(double) * n) != hipSuccess) {
    if (ut) (void)hipFree(ut);
    return fail(ED_ERR_ALLOC, "ed_plan_sai_apply: alloc");
  }
  const unsigned gn = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_sai_perm, dim3(gn), dim3(256), 0, st, n, P->perm, u, ut);
  sai_apply_launch(P, ut, 0, nullptr, nullptr, zt, st);
  hipLaunchKernelGGL(k_sai_unperm, dim3(gn), dim3(256), 0, st, n, P->perm, zt, z);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(ut);
  (void)hipFree(zt);
  if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("ed_plan_sai_apply: ") + hipGetErrorString(e));
  return ED_OK;
}

// ---------------------------------------------------------------------------
// Near-correction CSR built on the device from device pairs (no reference
// counterpart: the host path ed_plan_attach_near takes host arrays).  Pairs
// must be i-major (the order every near-pair builder produces) with valid
// indices; the result is identical to the host path's.
// ---------------------------------------------------------------------------
namespace ed {

__global__ void k_inv_perm(int64_t n, const int32_t* __restrict__ perm, int32_t* __restrict__ inv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) inv[perm[t]] = (int32_t)t;
}

__global__ void k_row_bounds(int64_t np, int64_t n, const int64_t* __restrict__ pairs,
                             int32_t* __restrict__ first, int32_t* __restrict__ last,
                             int32_t* __restrict__ bad) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int64_t i = pairs[2 * q], j = pairs[2 * q + 1];
  if (i < 0 || i >= n || j < 0 || j >= n) { atomicOr(bad, 1); return; }
  const int64_t ip = q > 0 ? pairs[2 * (q - 1)] : -1;
  const int64_t in = q + 1 < np ? pairs[2 * (q + 1)] : -1;
  if (q > 0 && ip > i) atomicOr(bad, 2);
  if (ip != i) first[i] = (int32_t)q;
  if (in != i) last[i] = (int32_t)(q + 1);
}

__global__ void k_row_counts(int64_t n, const int32_t* __restrict__ perm,
                             const int32_t* __restrict__ first, const int32_t* __restrict__ last,
                             int32_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) cnt[t] = last[perm[t]] - first[perm[t]];
}

__global__ void k_csr_fill(int64_t np, const int64_t* __restrict__ pairs,
                           const double* __restrict__ w, const int32_t* __restrict__ inv,
                           const int32_t* __restrict__ first, const int32_t* __restrict__ ptr,
                           int32_t* __restrict__ col, float* __restrict__ w32,
                           double* __restrict__ w64) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int64_t i = pairs[2 * q], j = pairs[2 * q + 1];
  const int64_t dst = ptr[inv[i]] + (q - first[i]);
  col[dst] = inv[j];
  w32[dst] = (float)w[q];
  w64[dst] = w[q];
}

}  // namespace ed

extern "C" int ed_plan_attach_near_device(ed_plan* P, const int64_t* pairs, const double* w,
                                          int64_t np, ed_stream stream) {
  if (!P) return fail(ED_ERR_ARG, "ed_plan_attach_near_device: NULL plan");
  if (np < 0 || (np > 0 && (!pairs || !w)) || np > (int64_t)INT32_MAX)
    return fail(ED_ERR_ARG, "ed_plan_attach_near_device: bad pairs");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = P->n;
  gmres_invalidate_graph(P);
  sai_free(P);
  (void)hipFree(P->near_ptr);
  (void)hipFree(P->near_col);
  (void)hipFree(P->near_w);
  (void)hipFree(P->near_w64);
  P->near_ptr = nullptr; P->near_col = nullptr; P->near_w = nullptr; P->near_w64 = nullptr;
  P->n_near = 0;
  int32_t *inv = nullptr, *first = nullptr, *last = nullptr, *cnt = nullptr, *bad = nullptr;
  int rc = ED_OK;
  const size_t m = (size_t)std::max<int64_t>(np, 1);
  if (hipMalloc((void**)&inv, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc((void**)&first, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc((void**)&last, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc((void**)&cnt, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc((void**)&bad, sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&P->near_ptr, sizeof(int32_t) * (n + 1)) != hipSuccess ||
      hipMalloc((void**)&P->near_col, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc((void**)&P->near_w, sizeof(float) * m) != hipSuccess ||
      hipMalloc((void**)&P->near_w64, sizeof(double) * m) != hipSuccess)
    rc = fail(ED_ERR_ALLOC, "ed_plan_attach_near_device: alloc");
  std::vector<int32_t> hcnt((size_t)n), hptr((size_t)n + 1, 0);
  int32_t hbad = 0;
  const unsigned gn = (unsigned)((n + 255) / 256), gp = (unsigned)((np + 255) / 256);
  hipError_t e = hipSuccess;
  if (rc == ED_OK) {
    e = hipMemsetAsync(first, 0, sizeof(int32_t) * n, st);
    if (e == hipSuccess) e = hipMemsetAsync(last, 0, sizeof(int32_t) * n, st);
    if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(int32_t), st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_inv_perm, dim3(gn), dim3(256), 0, st, n, P->perm, inv);
      if (np > 0)
        hipLaunchKernelGGL(k_row_bounds, dim3(gp), dim3(256), 0, st, np, n, pairs, first, last, bad);
      hipLaunchKernelGGL(k_row_counts, dim3(gn), dim3(256), 0, st, n, P->perm, first, last, cnt);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hcnt.data(), cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail(ED_ERR_DEVICE, std::string("ed_plan_attach_near_device: ") + hipGetErrorString(e));
  }
  if (rc == ED_OK && hbad)
    rc = fail(ED_ERR_ARG, (hbad & 1) ? "ed_plan_attach_near_device: index out of range"
                                     : "ed_plan_attach_near_device: pairs are not i-major");
  if (rc == ED_OK) {
    for (int64_t t = 0; t < n; ++t) hptr[t + 1] = hptr[t] + hcnt[t];
    e = hipMemcpyAsync(P->near_ptr, hptr.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && np > 0) {
      hipLaunchKernelGGL(k_csr_fill, dim3(gp), dim3(256), 0, st, np, pairs, w, inv, first,
                         P->near_ptr, P->near_col, P->near_w, P->near_w64);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail(ED_ERR_LAUNCH, std::string("ed_plan_attach_near_device: ") + hipGetErrorString(e));
  }
  (void)hipFree(inv); (void)hipFree(first); (void)hipFree(last); (void)hipFree(cnt); (void)hipFree(bad);
  if (rc != ED_OK) {
    // leave the plan with an empty (all-zero) correction
    (void)hipFree(P->near_ptr); (void)hipFree(P->near_col); (void)hipFree(P->near_w); (void)hipFree(P->near_w64);
    P->near_ptr = nullptr; P->near_col = nullptr; P->near_w = nullptr; P->near_w64 = nullptr;
    std::vector<int32_t> z((size_t)n + 1, 0);
    if (hipMalloc((void**)&P->near_ptr, sizeof(int32_t) * (n + 1)) == hipSuccess)
      (void)hipMemcpy(P->near_ptr, z.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice);
    return rc;
  }
  P->n_near = np;
  return ED_OK;
}
