// Dense single-layer collocation matvec (core/bem_kernel.py:157-230, :715-810):
//   V_i = sum_{j != i} K_E sigma_j A_j rsqrt(max(r_ij^2, 1e-24)) + K_ii sigma_i A_i
// with K_ii = self_integrals[i] (per-area) or 0.  Split form: two target rows
// per thread, source panels staged through LDS in tiles of 256, up to 8
// right-hand sides per pass (one r^2 and rsqrt per pair for all of them).
// Symmetric form (FP64, one right-hand side): each pair evaluated once for
// both rows.  The FP64 kernels refine the hardware
// FP64 reciprocal square root with one Newton step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>
#include <type_traits>

#include "ed_internal.h"
#include "ed_plan.h"

namespace ed {

constexpr int kDenseBlock = 256;
constexpr int kDenseRhs = 8;
constexpr int kDenseMaxSplit = 16;

template <typename T>
__device__ __forceinline__ T inv_sqrt(T r2);

// Hardware FP64 reciprocal square root (v_rsq_f64, an approximation) refined
// by one Newton step: measured max relative error of the dense matvec 1e-15
// against the FP64 oracle, far below what a 1e-8 residual needs.  (One
// instruction where an FP32 rsq needs two conversions around it.)
template <>
__device__ __forceinline__ double inv_sqrt<double>(double r2) {
  const double y = __builtin_amdgcn_rsq(r2);
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
__device__ __forceinline__ void dense_part_item(const DenseItem& it, const T* __restrict__ C,
                                                const T* __restrict__ A,
                                                const T* __restrict__ sigma, int64_t n,
                                                int64_t sig_ld, int r0, T* __restrict__ partial) {
  __shared__ T sx[kDenseBlock], sy[kDenseBlock], sz[kDenseBlock];
  __shared__ T sq[kDenseRhs][kDenseBlock];
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
__global__ __launch_bounds__(kDenseBlock) void k_dense_part(const DenseItem* __restrict__ items,
                                                            const T* __restrict__ C,
                                                            const T* __restrict__ A,
                                                            const T* __restrict__ sigma,
                                                            int64_t n, int64_t sig_ld, int r0,
                                                            T* __restrict__ partial) {
  dense_part_item<T, NR>(items[blockIdx.x], C, A, sigma, n, sig_ld, r0, partial);
}

// The same work items of one geometry [0, n) computed from the block index
// (item k = row block k / S, split k % S, as dense_schedule orders them):
// no schedule in device memory.
template <typename T, int NR>
__global__ __launch_bounds__(kDenseBlock) void k_dense_part_grid(int split, int64_t chunk,
                                                                 const T* __restrict__ C,
                                                                 const T* __restrict__ A,
                                                                 const T* __restrict__ sigma,
                                                                 int64_t n, int64_t sig_ld,
                                                                 int r0, T* __restrict__ partial) {
  const int64_t per_block = (int64_t)kDenseBlock * kDenseRows;
  DenseItem it;
  it.row0 = (int64_t)(blockIdx.x / split) * per_block;
  it.nrows = (int32_t)(n - it.row0 < per_block ? n - it.row0 : per_block);
  it.split = (int32_t)(blockIdx.x % split);
  it.src0 = (int64_t)it.split * chunk < n ? (int64_t)it.split * chunk : n;
  it.src1 = (int64_t)(it.split + 1) * chunk < n ? (int64_t)(it.split + 1) * chunk : n;
  dense_part_item<T, NR>(it, C, A, sigma, n, sig_ld, r0, partial);
}

// ---------------------------------------------------------------------------
// Symmetric form (FP64, one right-hand side).  The kernel 1/r is symmetric,
// so a tile of 512 rows a x 512 sources b of one geometry with b > a adds
// k_ij q_j to row i and k_ij q_i to row j: each pair is evaluated once, not
// twice.  Thread t holds rows t and t + 256 of the tile and at step s meets
// source (t + s) mod 512, so the 64 lanes of a wave touch 64 different
// sources per step; each wave adds its column contributions into its own LDS
// array (a plain read-modify-write, one lane per address per step), and the
// four arrays are added in a fixed order at the end.  Diagonal tiles (a = b)
// take row contributions only, skipping i = j.  Row and column partials go
// to one slot per tile; k_dense_sym_finish adds a row's slots in a fixed
// order (column slots of tiles c < a, then row slots of tiles b >= a).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sym_k(double xi, double yi, double zi, double x, double y,
                                        double z) {
  const double dx = xi - x, dy = yi - y, dz = zi - z;
  double r2 = dx * dx + dy * dy + dz * dz;
  r2 = r2 > 1e-24 ? r2 : 1e-24;
  return inv_sqrt<double>(r2);
}

__global__ __launch_bounds__(kDenseBlock) void k_dense_sym(const DenseSymItem* __restrict__ items,
                                                           const double* __restrict__ C,
                                                           const double* __restrict__ A,
                                                           const double* __restrict__ sigma,
                                                           double* __restrict__ rowpart,
                                                           double* __restrict__ colpart) {
  static_assert(kDenseSymTile == 2 * kDenseBlock, "two rows per thread");
  constexpr int NW = kDenseBlock / 64, TM = kDenseSymTile - 1;
  // sources as separate arrays: consecutive lanes read consecutive doubles
  __shared__ double sx[kDenseSymTile], sy[kDenseSymTile], sz[kDenseSymTile], sq[kDenseSymTile];
  __shared__ double cacc[NW][kDenseSymTile];
  const DenseSymItem it = items[blockIdx.x];
  const int t = threadIdx.x, w = t >> 6;
  for (int k = t; k < kDenseSymTile; k += kDenseBlock) {
    if (k < it.ns) {
      const int64_t j = it.s0 + k;
      sx[k] = C[3 * j]; sy[k] = C[3 * j + 1]; sz[k] = C[3 * j + 2];
      sq[k] = sigma[j] * A[j];
    } else {  // padding: far away, no charge
      sx[k] = sy[k] = sz[k] = 1e30;
      sq[k] = 0.0;
    }
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) cacc[ww][k] = 0.0;
  }
  double xi[2], yi[2], zi[2], qi[2], acc[2] = {0.0, 0.0};
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int r = t + u * kDenseBlock;
    if (r < it.nr) {
      const int64_t i = it.r0 + r;
      xi[u] = C[3 * i]; yi[u] = C[3 * i + 1]; zi[u] = C[3 * i + 2];
      qi[u] = sigma[i] * A[i];
    } else {  // padding row: far from every source and from the padding sources
      xi[u] = yi[u] = zi[u] = -1e30;
      qi[u] = 0.0;
    }
  }
  __syncthreads();
  if (it.diag) {
    for (int s = 0; s < kDenseSymTile; ++s) {
      const int k = (t + s) & TM;
      const double x = sx[k], y = sy[k], z = sz[k], q = sq[k];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        double kv = sym_k(xi[u], yi[u], zi[u], x, y, z);
        if (k == t + u * kDenseBlock) kv = 0.0;
        acc[u] += kv * q;
      }
    }
  } else {
    double* cw = cacc[w];
    int k = t;
    double x = sx[k], y = sy[k], z = sz[k], q = sq[k];
    for (int s = 0; s < kDenseSymTile; ++s) {
      // the next step's source (read-only) is fetched a step ahead
      const int kn = (k + 1) & TM;
      const double xn = sx[kn], yn = sy[kn], zn = sz[kn], qn = sq[kn];
      const double c = cw[k];
      const double k0 = sym_k(xi[0], yi[0], zi[0], x, y, z);
      const double k1 = sym_k(xi[1], yi[1], zi[1], x, y, z);
      acc[0] += k0 * q;
      acc[1] += k1 * q;
      cw[k] = c + (k0 * qi[0] + k1 * qi[1]);
      // the next step reads the column address another lane of this wave
      // wrote in this one: keep the accesses to cw in program order (the
      // wave's LDS operations execute in order)
      asm volatile("" ::: "memory");
      k = kn;
      x = xn; y = yn; z = zn; q = qn;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 2; ++u)
    rowpart[(size_t)blockIdx.x * kDenseSymTile + t + u * kDenseBlock] = acc[u];
  if (!it.diag)
    for (int k = t; k < kDenseSymTile; k += kDenseBlock) {
      double v = cacc[0][k];
#pragma unroll
      for (int ww = 1; ww < NW; ++ww) v += cacc[ww][k];
      colpart[(size_t)blockIdx.x * kDenseSymTile + k] = v;
    }
}

__global__ __launch_bounds__(kDenseBlock) void k_dense_sym_finish(
    const int64_t* __restrict__ trow0, const int32_t* __restrict__ tnr,
    const int32_t* __restrict__ tptr, const int32_t* __restrict__ tent,
    const double* __restrict__ rowpart, const double* __restrict__ colpart,
    const double* __restrict__ sigma, const double* __restrict__ selfi,
    const double* __restrict__ A, double ke, double* __restrict__ out) {
  const int rt = blockIdx.x;
  const int e0 = tptr[rt], e1 = tptr[rt + 1];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int r = threadIdx.x + u * kDenseBlock;
    if (r >= tnr[rt]) continue;
    const int64_t i = trow0[rt] + r;
    double acc = 0.0;
    for (int e = e0; e < e1; ++e) {
      const int32_t ent = tent[e];
      const double* part = (ent & 1) ? colpart : rowpart;
      acc += part[(size_t)(ent >> 1) * kDenseSymTile + r];
    }
    const double diag = selfi ? selfi[i] * A[i] : 0.0;
    double v = ke * acc + diag * sigma[i];
    if (!isfinite(v)) v = 0.0;  // core/bem_kernel.py:375-381
    out[i] = v;
  }
}

// Tiles of the symmetric form for geometry segments [off[g], off[g+1]).
void dense_sym_schedule(const std::vector<int64_t>& off, DenseSymSchedule& S) {
  S = DenseSymSchedule();
  S.tptr.push_back(0);
  for (size_t g = 0; g + 1 < off.size(); ++g) {
    const int64_t b = off[g], e = off[g + 1], ng = e - b;
    if (ng <= 0) continue;
    const int64_t T = (ng + kDenseSymTile - 1) / kDenseSymTile;
    const int64_t base = (int64_t)S.items.size();
    auto nrows = [&](int64_t a) { return (int32_t)std::min<int64_t>(kDenseSymTile, ng - a * kDenseSymTile); };
    auto idx = [&](int64_t a, int64_t c) { return base + a * T - a * (a - 1) / 2 + (c - a); };
    for (int64_t a = 0; a < T; ++a)
      for (int64_t c = a; c < T; ++c) {
        DenseSymItem it;
        it.r0 = b + a * kDenseSymTile;
        it.s0 = b + c * kDenseSymTile;
        it.nr = nrows(a);
        it.ns = nrows(c);
        it.diag = a == c ? 1 : 0;
        it.pad = 0;
        S.items.push_back(it);
      }
    for (int64_t a = 0; a < T; ++a) {
      S.trow0.push_back(b + a * kDenseSymTile);
      S.tnr.push_back(nrows(a));
      for (int64_t c = 0; c < a; ++c) S.tent.push_back((int32_t)(idx(c, a) * 2 + 1));
      for (int64_t c = a; c < T; ++c) S.tent.push_back((int32_t)(idx(a, c) * 2));
      S.tptr.push_back((int32_t)S.tent.size());
    }
  }
}

int dense_sym_upload(const DenseSymSchedule& S, DenseSymDevice& D) {
  D = DenseSymDevice();
  D.n_items = (int64_t)S.items.size();
  D.n_tiles = (int64_t)S.trow0.size();
  if (D.n_items == 0) return ED_OK;
  auto up = [](void** dst, const void* src, size_t bytes) {
    if (hipMalloc(dst, bytes) != hipSuccess) return false;
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up((void**)&D.items, S.items.data(), sizeof(DenseSymItem) * S.items.size()) ||
      !up((void**)&D.trow0, S.trow0.data(), sizeof(int64_t) * S.trow0.size()) ||
      !up((void**)&D.tnr, S.tnr.data(), sizeof(int32_t) * S.tnr.size()) ||
      !up((void**)&D.tptr, S.tptr.data(), sizeof(int32_t) * S.tptr.size()) ||
      !up((void**)&D.tent, S.tent.data(), sizeof(int32_t) * S.tent.size()) ||
      hipMalloc((void**)&D.part, sizeof(double) * 2 * kDenseSymTile * (size_t)D.n_items) != hipSuccess) {
    dense_sym_free(D);
    return fail(ED_ERR_ALLOC, "dense matvec: alloc");
  }
  return ED_OK;
}

void dense_sym_free(DenseSymDevice& D) {
  for (void* q : {(void*)D.items, (void*)D.trow0, (void*)D.tnr, (void*)D.tptr, (void*)D.tent, (void*)D.part})
    if (q) (void)hipFree(q);
  D = DenseSymDevice();
}

int dense_sym_run(const DenseSymDevice& D, const double* C, const double* A, const double* sigma,
                  const double* selfi, double* out, hipStream_t st) {
  if (D.n_items == 0) return ED_OK;
  double* rowpart = D.part;
  double* colpart = D.part + (size_t)kDenseSymTile * D.n_items;
  hipLaunchKernelGGL(k_dense_sym, dim3((unsigned)D.n_items), dim3(kDenseBlock), 0, st, D.items, C,
                     A, sigma, rowpart, colpart);
  hipLaunchKernelGGL(k_dense_sym_finish, dim3((unsigned)D.n_tiles), dim3(kDenseBlock), 0, st,
                     D.trow0, D.tnr, D.tptr, D.tent, (const double*)rowpart,
                     (const double*)colpart, sigma, selfi, A, (double)kKE, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("dense matvec: ") + hipGetErrorString(e));
  return ED_OK;
}

// row_nsplit NULL: every row has `split` partials (one geometry)
template <typename T, int NR>
__global__ void k_dense_finish(int64_t n, const int32_t* __restrict__ row_nsplit, int split,
                               const T* __restrict__ partial, const T* __restrict__ sigma,
                               int64_t sig_ld, const T* __restrict__ selfi,
                               const T* __restrict__ A, int r0, T ke, T* __restrict__ out,
                               int64_t out_ld) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int S = row_nsplit ? row_nsplit[i] : split;
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
    hipLaunchKernelGGL((k_dense_finish<T, NRV>), dim3(gn), dim3(256), 0, st, n, row_nsplit, 1, \
                       (const T*)partial, sigma, n, selfi, A, (int)r0, (T)kKE, out, n);        \
    break;
    switch (nr) {
      ED_DENSE_NR(1)
      ED_DENSE_NR(2)
      ED_DENSE_NR(3)
      ED_DENSE_NR(4)
      ED_DENSE_NR(5)
      ED_DENSE_NR(6)
      ED_DENSE_NR(7)
      default: ED_DENSE_NR(8)
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
  if constexpr (std::is_same<T, double>::value) {
    if (nrhs == 1) {  // symmetric form: every pair once
      DenseSymSchedule S;
      dense_sym_schedule({0, n}, S);
      DenseSymDevice D;
      int rc = dense_sym_upload(S, D);
      if (rc == ED_OK) rc = dense_sym_run(D, C, A, sigma, selfi, out, st);
      hipError_t e = hipSuccess;
      if (rc == ED_OK && (e = hipStreamSynchronize(st)) != hipSuccess)
        rc = fail(ED_ERR_DEVICE, std::string("dense matvec: ") + hipGetErrorString(e));
      dense_sym_free(D);
      return rc;
    }
  }
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

// ---------------------------------------------------------------------------
// No-allocation dense matvec (ed_dense_matvec_f64_noalloc): the split form's
// partial sums live in `scratch`, one copy of which each GPU holds until the
// process ends; the work items come from the block index.  Host code never
// passes `scratch` itself (a host-side stand-in): the device address of each
// GPU's copy is looked up once per device with that device current.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kNoAllocDoubles = int64_t(1) << 21;  // 16 MiB per GPU
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoAllocMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

static int noalloc_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return ed::fail(ED_ERR_DEVICE, std::string("dense noalloc: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return ed::fail(ED_ERR_DEVICE, "dense noalloc: device index");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return ed::fail(ED_ERR_DEVICE, std::string("dense noalloc: scratch address: ") + hipGetErrorString(e));
    g_noalloc_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_noalloc_ptr[dev];
  return ED_OK;
}

extern "C" int ed_dense_matvec_f64_noalloc(const double* C, const double* A, const double* sigma,
                                           const double* selfi, double* out, int64_t n,
                                           int64_t nrhs, ed_stream stream) {
  using namespace ed;
  if (n < 0 || nrhs < 1) return fail(ED_ERR_ARG, "dense noalloc: bad sizes");
  if (n == 0) return ED_OK;
  if (!C || !A || !sigma || !out) return fail(ED_ERR_ARG, "dense noalloc: NULL pointer");
  // the split of dense_schedule for one geometry
  const int S = (int)std::max<int64_t>(1, std::min<int64_t>(kDenseMaxSplit, (n + 2047) / 2048));
  const int64_t chunk = (n + S - 1) / S;
  if ((int64_t)S * n > kNoAllocDoubles)
    return fail(ED_ERR_ARG, "dense noalloc: n = " + std::to_string(n) + " exceeds the scratch limit "
                            "(split x n <= 2^21 doubles: n <= 131072)");
  const int64_t width = std::min<int64_t>(std::min<int64_t>(nrhs, kDenseRhs),
                                          kNoAllocDoubles / ((int64_t)S * n));
  double* part = nullptr;
  int rc = noalloc_scratch(&part);
  if (rc != ED_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t per_block = (int64_t)kDenseBlock * kDenseRows;
  const unsigned nb = (unsigned)(((n + per_block - 1) / per_block) * S);
  const unsigned gn = (unsigned)((n + 255) / 256);
  for (int64_t r0 = 0; r0 < nrhs; r0 += width) {
    const int nr = (int)std::min<int64_t>(width, nrhs - r0);
#define ED_NOALLOC_NR(NRV)                                                                     \
  case NRV:                                                                                     \
    hipLaunchKernelGGL((k_dense_part_grid<double, NRV>), dim3(nb), dim3(kDenseBlock), 0, st, S, \
                       chunk, C, A, sigma, n, n, (int)r0, part);                                \
    hipLaunchKernelGGL((k_dense_finish<double, NRV>), dim3(gn), dim3(256), 0, st, n,            \
                       (const int32_t*)nullptr, S, (const double*)part, sigma, n, selfi, A,     \
                       (int)r0, (double)kKE, out, n);                                           \
    break;
    switch (nr) {
      ED_NOALLOC_NR(1)
      ED_NOALLOC_NR(2)
      ED_NOALLOC_NR(3)
      ED_NOALLOC_NR(4)
      ED_NOALLOC_NR(5)
      ED_NOALLOC_NR(6)
      ED_NOALLOC_NR(7)
      default: ED_NOALLOC_NR(8)
    }
#undef ED_NOALLOC_NR
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ED_ERR_LAUNCH, std::string("dense noalloc: ") + hipGetErrorString(e));
  }
  return ED_OK;
}

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
