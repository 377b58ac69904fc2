// Restarted, left-(Jacobi-)preconditioned GMRES in FP64 on the device, with the
// plan's FMM operator (FP32 arithmetic) as A and, optionally, an outer FP64
// iterative refinement whose residual uses the dense FP64 operator plus the
// FP64 near correction (the accurate operator).
//
// Semantics follow core/bem_solver.py:125-920 (tol_abs = tol * ||b - A x0||,
// convergence on the TRUE residual at cycle end, codes 0/1/2/3/5/6), with two
// deliberate differences: orthogonalisation is classical Gram-Schmidt applied
// twice (CGS2, fused multi-dot kernels with fixed-order reductions) instead of
// MGS, and a cycle ends as soon as the preconditioned residual estimate meets
// its tolerance instead of always running `restart` steps (bem_solver.py:703-707).
// All vectors live in tree order; b / x / diag are permuted on entry and exit.
// A batched plan's operator is block diagonal over its geometries: without a
// coarse space the Arnoldi process runs on every geometry independently (one
// Krylov space, Hessenberg matrix and Givens sequence per geometry, all
// advanced by the same launches) and geometry g is converged at
// |r_g| <= tol |b_g|, so the step count is that of the slowest geometry,
// not that of the joint spectrum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "ed_plan.h"

namespace ed {

constexpr int kRedBlocks = 240;  // partial-sum blocks for multi-dots
constexpr int kRedThreads = 256;

struct GmresWS {
  int64_t n = 0;   // vector length: R stacked right-hand sides of n1 panels
  int64_t n1 = 0;  // panels of the plan
  int R = 1;
  int m = 0;
  double *V = nullptr, *w = nullptr, *vin = nullptr, *x = nullptr, *b = nullptr, *r = nullptr,
         *t = nullptr, *dinv = nullptr, *selfi_t = nullptr, *d = nullptr, *rhs = nullptr,
         *u2 = nullptr;
  // per segment: Hessenberg (m+1) x m, g (m+1), Givens cs / sn (m), CGS
  // coefficients h / h2 (m+2; h also holds y after the back substitution)
  double *H = nullptr, *g = nullptr, *cs = nullptr, *sn = nullptr, *h = nullptr, *h2 = nullptr;
  double* partial = nullptr;  // (m + 2) x tiles
  double* scal = nullptr;     // device scalars
  double* hscal = nullptr;    // pinned host mirror
  int* dj = nullptr;          // device Arnoldi index
  // Segments: the Arnoldi process runs independently on every geometry of a
  // batched plan (the operator is block diagonal) and every right-hand side,
  // sharing each launch.  Tiling k = 0: one segment per (right-hand side,
  // geometry); k = 1: one segment per right-hand side (single geometries, or
  // a coarse space that couples geometries).
  int Gmax = 1;
  int G_k[2] = {1, 1}, NT_k[2] = {0, 0};
  int4* tiles_k[2] = {nullptr, nullptr};  // (segment, lo, hi, 0)
  int* seg_t0_k[2] = {nullptr, nullptr};  // tile range of each segment
  double* segv = nullptr;   // device per-segment values (see SV_*), 8 x Gmax
  double* hsegv = nullptr;  // pinned host mirror
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  const double* graph_selfi = nullptr;
  int graph_mode = -1;  // preconditioner the captured graph was recorded with
  // per-segment residual and right-hand-side norms of the last solve (the
  // accurate operator's when it refined) and whether its segments were
  // (right-hand side, geometry) blocks (tiling 0) or whole right-hand sides
  std::vector<double> last_rn, last_bn;
  int last_tiling = -1;
  // per-poll trace of the last solve (kTraceCols doubles a row, see
  // ed_gmres_trace)
  std::vector<double> trace;
  // split dense FP64 operator for R > 1 right-hand sides (built on first use):
  // work items over the plan's geometries, splits per row, partial sums
  DenseItem* dn_items = nullptr;
  int32_t* dn_rns = nullptr;
  double* dn_part = nullptr;
  int64_t dn_n_items = 0;
  int dn_smax = 0;
};

// device scalar slots
enum { S_EST = 0, S_BREAK = 5, S_NSCAL = 8 };
// per-segment value rows of segv (each Gmax long)
enum { SV_NRM_R = 0, SV_NRM_Z = 1, SV_EST_TOL = 2, SV_ACTIVE = 3, SV_HNEXT = 4, SV_N = 8 };

void gmres_free(ed_plan* P) {
  GmresWS* W = (GmresWS*)P->gmres_ws;
  if (!W) return;
  if (W->gexec) (void)hipGraphExecDestroy(W->gexec);
  if (W->graph) (void)hipGraphDestroy(W->graph);
  void* ptrs[] = {W->V, W->w, W->vin, W->x, W->b, W->r, W->t, W->dinv, W->selfi_t, W->d, W->rhs,
                  W->u2, W->H, W->g, W->cs, W->sn, W->h, W->h2, W->partial, W->scal, W->dj,
                  W->tiles_k[0], W->tiles_k[1], W->seg_t0_k[0], W->seg_t0_k[1], W->segv,
                  W->dn_items, W->dn_rns, W->dn_part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (W->hscal) (void)hipHostFree(W->hscal);
  if (W->hsegv) (void)hipHostFree(W->hsegv);
  if (W->stream) (void)hipStreamDestroy(W->stream);
  delete W;
  P->gmres_ws = nullptr;
}

// Drops the captured Arnoldi graph (it holds the addresses of the operator
// and preconditioner arrays) but keeps the workspace allocations.
void gmres_invalidate_graph(ed_plan* P) {
  GmresWS* W = (GmresWS*)P->gmres_ws;
  if (!W) return;
  if (W->gexec) (void)hipGraphExecDestroy(W->gexec);
  if (W->graph) (void)hipGraphDestroy(W->graph);
  W->gexec = nullptr;
  W->graph = nullptr;
  W->graph_selfi = nullptr;
  W->graph_mode = -1;
}

#define ED_G(call)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(e_ == hipErrorOutOfMemory ? ED_ERR_ALLOC : ED_ERR_DEVICE,              \
                  std::string(#call) + ": " + hipGetErrorString(e_));                    \
  } while (0)
#define ED_T(x)                     \
  do {                              \
    int rc_ = (x);                  \
    if (rc_ != ED_OK) return rc_;   \
  } while (0)
#define ED_KCHECK(msg)                                                                  \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess)                                                               \
      return fail(ED_ERR_LAUNCH, std::string(msg) + ": " + hipGetErrorString(e_));      \
  } while (0)

// Tiles of at most `ts` entries inside each segment [off[g], off[g+1]).
static void make_tiles(const std::vector<int64_t>& off, int64_t ts, std::vector<int4>& tiles,
                       std::vector<int>& t0) {
  tiles.clear();
  t0.assign(1, 0);
  for (size_t g = 0; g + 1 < off.size(); ++g) {
    for (int64_t lo = off[g]; lo < off[g + 1]; lo += ts)
      tiles.push_back(make_int4((int)g, (int)lo, (int)std::min(lo + ts, off[g + 1]), 0));
    t0.push_back((int)tiles.size());
  }
}

static int ws_get(ed_plan* P, int m, int R, GmresWS** out) {
  GmresWS* W = (GmresWS*)P->gmres_ws;
  if (W && W->m == m && W->n1 == P->n && W->R == R) { *out = W; return ED_OK; }
  if ((int64_t)R * P->n >= ((int64_t)1 << 31))
    return fail(ED_ERR_ARG, "GMRES: right-hand sides x panels must stay below 2^31");
  gmres_free(P);
  W = new GmresWS();
  P->gmres_ws = W;
  W->n1 = P->n;
  W->R = R;
  W->n = (int64_t)R * P->n;
  W->m = m;
  const size_t n = (size_t)W->n;
  std::vector<int64_t> geo1 = P->host.geom_pt_off;
  if (geo1.size() < 2) geo1 = {0, P->n};
  // segment boundaries of the R stacked vectors
  std::vector<int64_t> geo = {0}, whole = {0};
  for (int r = 0; r < R; ++r) {
    for (size_t g = 1; g < geo1.size(); ++g) geo.push_back((int64_t)r * P->n + geo1[g]);
    whole.push_back((int64_t)(r + 1) * P->n);
  }
  // ~512 tiles over the whole vector, 256..4096 entries each
  const int64_t ts = std::min<int64_t>(4096, std::max<int64_t>(256, ((int64_t)n / 512 + 63) / 64 * 64));
  int ntmax = 1;
  for (int k = 0; k < 2; ++k) {
    std::vector<int4> tl;
    std::vector<int> t0;
    make_tiles(k == 0 ? geo : whole, ts, tl, t0);
    W->G_k[k] = (int)t0.size() - 1;
    W->NT_k[k] = (int)tl.size();
    ntmax = std::max(ntmax, W->NT_k[k]);
    ED_G(hipMalloc((void**)&W->tiles_k[k], sizeof(int4) * std::max<size_t>(tl.size(), 1)));
    ED_G(hipMalloc((void**)&W->seg_t0_k[k], sizeof(int) * t0.size()));
    if (!tl.empty())
      ED_G(hipMemcpy(W->tiles_k[k], tl.data(), sizeof(int4) * tl.size(), hipMemcpyHostToDevice));
    ED_G(hipMemcpy(W->seg_t0_k[k], t0.data(), sizeof(int) * t0.size(), hipMemcpyHostToDevice));
  }
  const int G = W->Gmax = std::max(W->G_k[0], W->G_k[1]);
  ED_G(hipMalloc((void**)&W->V, sizeof(double) * n * (m + 1)));
  double** vecs[] = {&W->w, &W->vin, &W->x, &W->b, &W->r, &W->t, &W->dinv, &W->selfi_t, &W->d,
                     &W->rhs, &W->u2};
  for (double** v : vecs) ED_G(hipMalloc((void**)v, sizeof(double) * n));
  ED_G(hipMalloc((void**)&W->H, sizeof(double) * G * (m + 1) * m));
  ED_G(hipMalloc((void**)&W->g, sizeof(double) * G * (m + 1)));
  ED_G(hipMalloc((void**)&W->cs, sizeof(double) * G * m));
  ED_G(hipMalloc((void**)&W->sn, sizeof(double) * G * m));
  ED_G(hipMalloc((void**)&W->h, sizeof(double) * G * (m + 2)));
  ED_G(hipMalloc((void**)&W->h2, sizeof(double) * G * (m + 2)));
  ED_G(hipMalloc((void**)&W->partial, sizeof(double) * (size_t)ntmax * (m + 2)));
  ED_G(hipMalloc((void**)&W->scal, sizeof(double) * S_NSCAL));
  ED_G(hipMalloc((void**)&W->segv, sizeof(double) * SV_N * G));
  ED_G(hipMalloc((void**)&W->dj, sizeof(int)));
  ED_G(hipHostMalloc((void**)&W->hscal, sizeof(double) * S_NSCAL, hipHostMallocDefault));
  ED_G(hipHostMalloc((void**)&W->hsegv, sizeof(double) * SV_N * G, hipHostMallocDefault));
  ED_G(hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking));
  *out = W;
  return ED_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_perm_in(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ src,
                          double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
__global__ void k_perm_out(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ src,
                           double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}
// Jacobi inverse: 1/d with non-finite or |d| < 1e-18 entries -> 1
// (core/bem_solver.py:67-78); identity when diag is NULL.
__global__ void k_dinv(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ diag,
                       double* __restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = diag ? diag[perm[i]] : 1.0;
  if (!isfinite(d) || fabs(d) < 1e-18) d = 1.0;
  dinv[i] = 1.0 / d;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
  return s;  // valid in thread 0
}

// Fixed-order reduction of `cnt` partial vectors; writes sqrt into scal slots.
__global__ void k_finish_norms(const double* __restrict__ partial, int cnt, double* __restrict__ scal,
                               int slot0) {
  __shared__ double sh[4];
  for (int k = 0; k < cnt; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < kRedBlocks; b += blockDim.x) v += partial[k * kRedBlocks + b];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) scal[slot0 + k] = sqrt(v);
  }
}

// ---- segmented Arnoldi kernels --------------------------------------------
// One block per tile (segment, lo, hi); per-tile partial sums are reduced per
// segment over its tile range in a fixed order, so every result is
// deterministic.  Segment values live in rows of segv (SV_*, stride G).

// r = b - t (t may be NULL), z = dinv .* r; partial[0][tile] = |r|^2,
// partial[1][tile] = |z|^2 (z, r optional outputs)
__global__ __launch_bounds__(kRedThreads) void k_resid_seg(const int4* __restrict__ tiles, int nt,
                                                           const double* __restrict__ b,
                                                           const double* __restrict__ t,
                                                           const double* __restrict__ dinv,
                                                           double* __restrict__ r,
                                                           double* __restrict__ z,
                                                           double* __restrict__ partial) {
  __shared__ double sh[4];
  const int4 tl = tiles[blockIdx.x];
  double s0 = 0.0, s1 = 0.0;
  for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) {
    const double ri = b[i] - (t ? t[i] : 0.0);
    if (r) r[i] = ri;
    const double zi = dinv[i] * ri;
    if (z) z[i] = zi;
    s0 += ri * ri;
    s1 += zi * zi;
  }
  double a = block_sum(s0, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = a;
  a = block_sum(s1, sh);
  if (threadIdx.x == 0) partial[nt + blockIdx.x] = a;
}

// partial[tile] = |x|^2 over the tile
__global__ __launch_bounds__(kRedThreads) void k_norm2_seg(const int4* __restrict__ tiles,
                                                           const double* __restrict__ x,
                                                           double* __restrict__ partial) {
  __shared__ double sh[4];
  const int4 tl = tiles[blockIdx.x];
  double s = 0.0;
  for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) s += x[i] * x[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[k * ldo + g] = sqrt(sum of partial[k][tiles of g]), k < cnt; one block
// per segment
__global__ void k_finish_seg_norms(const double* __restrict__ partial, int nt, int cnt,
                                   const int* __restrict__ seg_t0, double* __restrict__ out,
                                   int ldo) {
  __shared__ double sh[4];
  const int g = blockIdx.x;
  const int t0 = seg_t0[g], t1 = seg_t0[g + 1];
  for (int k = 0; k < cnt; ++k) {
    double v = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) v += partial[(size_t)k * nt + t];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) out[(size_t)k * ldo + g] = sqrt(v);
  }
}

// V0 = z / beta_g (0 for an inactive segment); g_g = [beta_g, 0...]; j = 0
__global__ void k_start_seg(const int4* __restrict__ tiles, int G, int m,
                            const double* __restrict__ z, double* __restrict__ V,
                            double* __restrict__ vin, double* __restrict__ gv,
                            const double* __restrict__ segv, int ldsv, int* dj) {
  const int4 tl = tiles[blockIdx.x];
  const double beta = segv[SV_NRM_Z * ldsv + tl.x];
  const bool act = segv[SV_ACTIVE * ldsv + tl.x] != 0.0 && beta > 0.0;
  const double inv = act ? 1.0 / beta : 0.0;
  for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) {
    const double v = z[i] * inv;
    V[i] = v;
    vin[i] = v;  // the first Arnoldi step's operator input
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) *dj = 0;
    for (int e = threadIdx.x; e < G * (m + 1); e += blockDim.x) {
      const int sg = e / (m + 1), k = e - sg * (m + 1);
      const bool a = segv[SV_ACTIVE * ldsv + sg] != 0.0;
      gv[e] = (k == 0 && a) ? segv[SV_NRM_Z * ldsv + sg] : 0.0;
    }
  }
}

// w = dinv .* w (optional); partial[k][tile] = V_k . w, k <= j
__global__ __launch_bounds__(kRedThreads) void k_dots_seg(const int4* __restrict__ tiles, int nt,
                                                          int64_t n, const double* __restrict__ dinv,
                                                          double* __restrict__ w,
                                                          const double* __restrict__ V,
                                                          const int* __restrict__ dj,
                                                          double* __restrict__ partial) {
  __shared__ double sh[4];
  const int j = *dj;
  const int4 tl = tiles[blockIdx.x];
  if (dinv)
    for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) w[i] *= dinv[i];
  __syncthreads();
  for (int k = 0; k <= j; ++k) {
    const double* vk = V + (size_t)k * n;
    double s = 0.0;
    for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) s += vk[i] * w[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(size_t)k * nt + blockIdx.x] = s;
  }
}

// h[g * ldh + k] = sum of partial[k][tiles of g]; one block per (segment g =
// blockIdx.x, row k = blockIdx.y), rows k > j exit
__global__ void k_finish_dots_seg(const double* __restrict__ partial, int nt,
                                  const int* __restrict__ seg_t0, const int* __restrict__ dj,
                                  double* __restrict__ h, int ldh) {
  __shared__ double sh[4];
  const int g = blockIdx.x, k = blockIdx.y;
  if (k > *dj) return;  // uniform per block
  const int t0 = seg_t0[g], t1 = seg_t0[g + 1];
  double v = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) v += partial[(size_t)k * nt + t];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) h[(size_t)g * ldh + k] = v;
}

// w -= sum_k c[g][k] V_k over the tile; then partial dots V_k . w (second CGS
// pass) or |w|^2 (norm pass)
__global__ __launch_bounds__(kRedThreads) void k_update_seg(const int4* __restrict__ tiles, int nt,
                                                            int64_t n, double* __restrict__ w,
                                                            const double* __restrict__ V,
                                                            const int* __restrict__ dj,
                                                            const double* __restrict__ c, int ldc,
                                                            int norm_pass,
                                                            double* __restrict__ partial) {
  __shared__ double sh[4];
  __shared__ double cc[256];
  const int j = *dj;
  const int4 tl = tiles[blockIdx.x];
  for (int k = threadIdx.x; k <= j; k += blockDim.x) cc[k] = c[(size_t)tl.x * ldc + k];
  __syncthreads();
  double nrm = 0.0;
  for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) {
    double wi = w[i];
    for (int k = 0; k <= j; ++k) wi -= cc[k] * V[(size_t)k * n + i];
    w[i] = wi;
    nrm += wi * wi;
  }
  __syncthreads();
  if (norm_pass) {
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = nrm;
    return;
  }
  for (int k = 0; k <= j; ++k) {
    const double* vk = V + (size_t)k * n;
    double s = 0.0;
    for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) s += vk[i] * w[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(size_t)k * nt + blockIdx.x] = s;
  }
}

// Per segment (one thread each): |w_g| from the norm-pass partials, column
// H_g[:, j] = h1 + h2 with H[j+1, j] = |w_g|, the Givens rotations
// (core/bem_solver.py:590-616) and est_g = |g_g[j+1]|.  Scalars: S_EST =
// max over active segments of est_g / est_tol_g, S_BREAK = 1 when an active
// segment breaks down.  j++.
constexpr int kGivLds = 4, kGivLen = 258;  // LDS path: segments, column length
__global__ void k_givens_seg(int G, int m, const double* __restrict__ partial,
                             const int* __restrict__ seg_t0, const double* __restrict__ h1,
                             const double* __restrict__ h2, double* __restrict__ H,
                             double* __restrict__ gv, double* __restrict__ cs,
                             double* __restrict__ sn, double* __restrict__ segv, int ldsv,
                             double* __restrict__ scal, int* dj) {
  __shared__ double smax[256];
  __shared__ int sbrk[256];
  const int j = *dj;
  const int ld = m + 1;
  // |w_g|^2 per segment: wave w sums the tiles of segments w, w + 4, ... across
  // its lanes (fixed order), into segv row SV_HNEXT
  {
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int sg = wv; sg < G; sg += nw) {
      double v = 0.0;
      for (int t = seg_t0[sg] + ln; t < seg_t0[sg + 1]; t += 64) v += partial[t];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (ln == 0) segv[SV_HNEXT * ldsv + sg] = v;
    }
  }
  __syncthreads();
  double rmax = 0.0;
  int brk = 0;
  if (G <= kGivLds && j + 2 <= kGivLen) {
    // few segments (one geometry): stage the column and the previous
    // rotations in LDS with all threads, so the serial rotation chain of
    // each segment runs on LDS instead of dependent global loads
    __shared__ double lcol[kGivLds][kGivLen], lc[kGivLds][kGivLen], ls[kGivLds][kGivLen];
    for (int e = threadIdx.x; e < G * (j + 1); e += blockDim.x) {
      const int sg = e / (j + 1), k = e - sg * (j + 1);
      lcol[sg][k] = h1[(size_t)sg * (m + 2) + k] + h2[(size_t)sg * (m + 2) + k];
      if (k < j) {
        lc[sg][k] = cs[(size_t)sg * m + k];
        ls[sg][k] = sn[(size_t)sg * m + k];
      }
    }
    __syncthreads();
    if (threadIdx.x < G) {
      const int sg = threadIdx.x;
      double* col = lcol[sg];
      double* g = gv + (size_t)sg * ld;
      const double hn = sqrt(segv[SV_HNEXT * ldsv + sg]);
      col[j + 1] = hn;
      segv[SV_HNEXT * ldsv + sg] = hn;
      const bool act = segv[SV_ACTIVE * ldsv + sg] != 0.0;
      if (act && (!isfinite(hn) || hn <= 0.0)) brk = 1;
      for (int k = 0; k < j; ++k) {
        const double t = lc[sg][k] * col[k] + ls[sg][k] * col[k + 1];
        col[k + 1] = -ls[sg][k] * col[k] + lc[sg][k] * col[k + 1];
        col[k] = t;
      }
      const double den = sqrt(col[j] * col[j] + col[j + 1] * col[j + 1]);
      double c = 1.0, s = 0.0;
      if (den >= 1e-14 && isfinite(den)) { c = col[j] / den; s = col[j + 1] / den; }
      cs[(size_t)sg * m + j] = c;
      sn[(size_t)sg * m + j] = s;
      col[j] = c * col[j] + s * col[j + 1];
      col[j + 1] = 0.0;
      const double gj = g[j];
      g[j] = c * gj;
      g[j + 1] = -s * gj;
      if (act) {
        const double et = segv[SV_EST_TOL * ldsv + sg];
        const double ratio = fabs(g[j + 1]) / (et > 0.0 ? et : 1e-300);
        rmax = isfinite(ratio) ? fmax(rmax, ratio) : INFINITY;
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < G * (j + 2); e += blockDim.x) {
      const int sg = e / (j + 2), k = e - sg * (j + 2);
      H[((size_t)sg * m + j) * ld + k] = lcol[sg][k];
    }
  } else
  for (int sg = threadIdx.x; sg < G; sg += blockDim.x) {
    const double v = segv[SV_HNEXT * ldsv + sg];
    const double* a1 = h1 + (size_t)sg * (m + 2);
    const double* a2 = h2 + (size_t)sg * (m + 2);
    double* col = H + ((size_t)sg * m + j) * ld;
    double* g = gv + (size_t)sg * ld;
    double* c_ = cs + (size_t)sg * m;
    double* s_ = sn + (size_t)sg * m;
    for (int k = 0; k <= j; ++k) col[k] = a1[k] + a2[k];
    const double hn = sqrt(v);
    col[j + 1] = hn;
    segv[SV_HNEXT * ldsv + sg] = hn;
    const bool act = segv[SV_ACTIVE * ldsv + sg] != 0.0;
    if (act && (!isfinite(hn) || hn <= 0.0)) brk = 1;
    for (int k = 0; k < j; ++k) {
      const double t = c_[k] * col[k] + s_[k] * col[k + 1];
      col[k + 1] = -s_[k] * col[k] + c_[k] * col[k + 1];
      col[k] = t;
    }
    const double den = sqrt(col[j] * col[j] + col[j + 1] * col[j + 1]);
    double c = 1.0, s = 0.0;
    if (den >= 1e-14 && isfinite(den)) { c = col[j] / den; s = col[j + 1] / den; }
    c_[j] = c;
    s_[j] = s;
    col[j] = c * col[j] + s * col[j + 1];
    col[j + 1] = 0.0;
    const double gj = g[j];
    g[j] = c * gj;
    g[j + 1] = -s * gj;
    if (act) {
      const double et = segv[SV_EST_TOL * ldsv + sg];
      const double ratio = fabs(g[j + 1]) / (et > 0.0 ? et : 1e-300);
      rmax = isfinite(ratio) ? fmax(rmax, ratio) : INFINITY;
    }
  }
  smax[threadIdx.x] = rmax;
  sbrk[threadIdx.x] = brk;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    int b = 0;
    for (int k = 0; k < (int)blockDim.x; ++k) {
      r = fmax(r, smax[k]);
      b |= sbrk[k];
    }
    scal[S_EST] = r;
    if (b) scal[S_BREAK] = 1.0;
    *dj = j + 1;
  }
}

// V_j = w / hnext_g over the tile (j already incremented; 0 after a
// breakdown), also written to vin, the next step's operator input (a fixed
// address, so a captured step needs no copy of V_j)
__global__ void k_scale_next_seg(const int4* __restrict__ tiles, int64_t n,
                                 const double* __restrict__ w, double* __restrict__ V,
                                 double* __restrict__ vin, const int* __restrict__ dj,
                                 const double* __restrict__ segv, int ldsv) {
  const int j = *dj;
  const int4 tl = tiles[blockIdx.x];
  const double hn = segv[SV_HNEXT * ldsv + tl.x];
  const double inv = (hn > 0.0 && isfinite(hn)) ? 1.0 / hn : 0.0;
  double* vj = V + (size_t)j * n;
  for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) {
    const double v = w[i] * inv;
    vj[i] = v;
    vin[i] = v;
  }
}

// y_g = R_g^{-1} g_g (back substitution, m_use = *dj), one thread per segment,
// stored in h
__global__ void k_backsolve_seg(int G, int m, const double* __restrict__ H,
                                const double* __restrict__ gv, const int* __restrict__ dj,
                                double* __restrict__ y) {
  const int mu = *dj;
  const int ld = m + 1;
  for (int sg = blockIdx.x * blockDim.x + threadIdx.x; sg < G; sg += gridDim.x * blockDim.x) {
    const double* Hs = H + (size_t)sg * m * ld;
    const double* g = gv + (size_t)sg * ld;
    double* ys = y + (size_t)sg * (m + 2);
    for (int i = mu - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < mu; ++k) s -= Hs[(size_t)k * ld + i] * ys[k];
      const double d = Hs[(size_t)i * ld + i];
      ys[i] = (d != 0.0 && isfinite(d)) ? s / d : 0.0;
    }
  }
}

// x += sum_{k < m_use} y_g[k] V_k over the tile
__global__ void k_xupdate_seg(const int4* __restrict__ tiles, int64_t n, int m,
                              double* __restrict__ x, const double* __restrict__ V,
                              const double* __restrict__ y, const int* __restrict__ dj) {
  __shared__ double yy[256];
  const int mu = *dj;
  const int4 tl = tiles[blockIdx.x];
  for (int k = threadIdx.x; k < mu; k += blockDim.x) yy[k] = y[(size_t)tl.x * (m + 2) + k];
  __syncthreads();
  for (int i = tl.y + threadIdx.x; i < tl.z; i += blockDim.x) {
    double s = x[i];
    for (int k = 0; k < mu; ++k) s += yy[k] * V[(size_t)k * n + i];
    x[i] = s;
  }
}

__global__ void k_axpy(int64_t n, double a, const double* __restrict__ xx, double* __restrict__ yy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) yy[i] += a * xx[i];
}

// out += sum_j w64 x_j over the tree-order near CSR (FP64)
__global__ void k_near_csr(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                           const double* __restrict__ w, const double* __restrict__ x,
                           double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  const int b = ptr[i], e = ptr[i + 1];
  for (int q = b; q < e; q += 4) {  // four loads in flight, clamped and masked
    int c[4];
    double wq[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = min(q + u, e - 1);
      c[u] = col[k];
      const double wv = w[k];
      wq[u] = q + u < e ? wv : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) s += wq[u] * xv[u];
  }
  out[i] += s;
}

// ---------------------------------------------------------------------------
// coarse space: M^-1 u = D^-1 (u - AZ c) + Z c,  c = E^-1 Z^T u,  E = Z^T A Z
// (Z orthonormalised at set-up; exact on span(Z), Jacobi elsewhere)
// ---------------------------------------------------------------------------
constexpr int kMaxCoarse = 32;

// partial[c][block] = sum_i Z_c[i] u[i] over the block's chunk
__global__ __launch_bounds__(kRedThreads) void k_coarse_dots(int64_t n, int k,
                                                             const double* __restrict__ Z,
                                                             const double* __restrict__ u,
                                                             double* __restrict__ partial) {
  __shared__ double sh[4];
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  for (int c = 0; c < k; ++c) {
    const double* zc = Z + (size_t)c * n;
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) s += zc[i] * u[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(size_t)c * kRedBlocks + blockIdx.x] = s;
  }
}

// cy = Einv (Z^T u), one block
__global__ void k_coarse_solve(int k, const double* __restrict__ partial,
                               const double* __restrict__ Einv, double* __restrict__ cy) {
  __shared__ double zu[kMaxCoarse];
  __shared__ double sh[4];
  for (int c = 0; c < k; ++c) {
    double v = 0.0;
    for (int b = threadIdx.x; b < kRedBlocks; b += blockDim.x) v += partial[(size_t)c * kRedBlocks + b];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) zu[c] = v;
  }
  __syncthreads();
  if (threadIdx.x < k) {
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += Einv[threadIdx.x * k + c] * zu[c];
    cy[threadIdx.x] = s;
  }
}

// out = dinv .* (u - AZ cy) + Z cy   (out may alias u)
__global__ void k_coarse_apply(int64_t n, int k, const double* __restrict__ u,
                               const double* __restrict__ dinv, const double* __restrict__ Z,
                               const double* __restrict__ AZ, const double* __restrict__ cy,
                               double* out) {
  __shared__ double c[kMaxCoarse];
  if (threadIdx.x < k) c[threadIdx.x] = cy[threadIdx.x];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double a = u[i], zc = 0.0;
    for (int q = 0; q < k; ++q) {
      a -= c[q] * AZ[(size_t)q * n + i];
      zc += c[q] * Z[(size_t)q * n + i];
    }
    out[i] = dinv[i] * a + zc;
  }
}

// out = Z cy
__global__ void k_coarse_combine(int64_t n, int k, const double* __restrict__ Z,
                                 const double* __restrict__ cy, double* __restrict__ out) {
  __shared__ double c[kMaxCoarse];
  if (threadIdx.x < k) c[threadIdx.x] = cy[threadIdx.x];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int q = 0; q < k; ++q) v += c[q] * Z[(size_t)q * n + i];
    out[i] = v;
  }
}

// u -= AZ cy (in place)
__global__ void k_coarse_sub(int64_t n, int k, const double* __restrict__ AZ,
                             const double* __restrict__ cy, double* __restrict__ u) {
  __shared__ double c[kMaxCoarse];
  if (threadIdx.x < k) c[threadIdx.x] = cy[threadIdx.x];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = u[i];
    for (int q = 0; q < k; ++q) v -= c[q] * AZ[(size_t)q * n + i];
    u[i] = v;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_norm2_part(int64_t n, const double* __restrict__ x,
                                                            double* __restrict__ partial) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * x[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Z_c -= sum_{q<c} (Z_q . Z_c) Z_q for one column (coefficients in cy), then
// the caller normalises; used twice per column at set-up (CGS2).
__global__ void k_coarse_orth(int64_t n, int c, double* __restrict__ Z, const double* __restrict__ cy) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = Z[(size_t)c * n + i];
    for (int q = 0; q < c; ++q) v -= cy[q] * Z[(size_t)q * n + i];
    Z[(size_t)c * n + i] = v;
  }
}

__global__ void k_coarse_coef(int c, const double* __restrict__ partial, double* __restrict__ cy) {
  __shared__ double sh[4];
  for (int q = 0; q < c; ++q) {
    double v = 0.0;
    for (int b = threadIdx.x; b < kRedBlocks; b += blockDim.x) v += partial[(size_t)q * kRedBlocks + b];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) cy[q] = v;
  }
}

__global__ void k_scale_col(int64_t n, double* __restrict__ x, const double* __restrict__ scal, int slot) {
  const double nr = scal[slot];
  const double inv = nr > 0.0 ? 1.0 / nr : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= inv;
}

// E (k x k, row-major E[r][c] = Z_r . AZ_c) from partial dots, then its
// inverse by Gauss-Jordan with partial pivoting (pivots below 1e-13 max|E|
// drop that direction).  One block.
__global__ void k_coarse_einv(int k, const double* __restrict__ partial, double* __restrict__ Einv) {
  __shared__ double M[kMaxCoarse][2 * kMaxCoarse];
  __shared__ double sh[4];
  __shared__ int piv;
  __shared__ double pmax;
  for (int e = 0; e < k * k; ++e) {
    double v = 0.0;
    for (int b = threadIdx.x; b < kRedBlocks; b += blockDim.x) v += partial[(size_t)e * kRedBlocks + b];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) M[e % k][e / k] = v;  // e = c * k + r holds Z_r . AZ_c
  }
  __syncthreads();
  for (int t = threadIdx.x; t < k * k; t += blockDim.x) M[t / k][k + t % k] = (t / k == t % k) ? 1.0 : 0.0;
  __syncthreads();
  double scale = 0.0;
  for (int r = 0; r < k; ++r)
    for (int c = 0; c < k; ++c) scale = fmax(scale, fabs(M[r][c]));
  for (int col = 0; col < k; ++col) {
    if (threadIdx.x == 0) {
      int best = col;
      for (int r = col + 1; r < k; ++r)
        if (fabs(M[r][col]) > fabs(M[best][col])) best = r;
      piv = best;
      pmax = M[best][col];
    }
    __syncthreads();
    const bool drop = fabs(pmax) <= 1e-13 * scale;
    if (!drop && piv != col && threadIdx.x < 2 * k) {
      const double a = M[col][threadIdx.x];
      M[col][threadIdx.x] = M[piv][threadIdx.x];
      M[piv][threadIdx.x] = a;
    }
    __syncthreads();
    if (drop) {
      if (threadIdx.x < 2 * k) M[col][threadIdx.x] = 0.0;
      __syncthreads();
      continue;
    }
    if (threadIdx.x < 2 * k) M[col][threadIdx.x] /= pmax;
    __syncthreads();
    for (int r = threadIdx.x; r < k; r += blockDim.x) {
      if (r == col) continue;
      const double f = M[r][col];
      for (int c = 0; c < 2 * k; ++c) M[r][c] -= f * M[col][c];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < k * k; t += blockDim.x) Einv[t] = M[t / k][k + t % k];
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
int dense_launch_f64(const double* C, const double* A, const double* sigma, const double* selfi,
                     double* out, int64_t n, int64_t nrhs, hipStream_t st);

// Trace rows: host time (s since the epoch), outer refinement step, cycle,
// Arnoldi steps so far, true residual norm (NaN when not measured),
// preconditioned residual norm or estimate, Krylov dimension, phase
// (0 initial residual, 1 Arnoldi poll, 2 cycle end, 3 refinement residual
// with the accurate operator).
constexpr int kTraceCols = 8;
enum TracePhase {
  kTraceInitial = 0,
  kTracePoll = 1,
  kTraceCycleEnd = 2,
  kTraceRefine = 3,
  kTraceUpdate = 4,
  kTraceApprox = 5
};

struct Solver {
  ed_plan* P;
  GmresWS* W;
  const ed_gmres_opts* o;
  const double* selfi_orig;
  int64_t n;   // stacked length R x n1
  unsigned gb;
  int64_t n1;  // panels
  int R;
  unsigned gn1;  // one block per 256 panels
  int total_iters = 0, cycles = 0, first_below = -1;
  int refine_step = 0;

  void trace(int phase, int dim, double rtrue, double rprec) {
    const double ts = std::chrono::duration<double>(
                          std::chrono::system_clock::now().time_since_epoch()).count();
    const double row[kTraceCols] = {ts, (double)refine_step, (double)cycles, (double)total_iters,
                                    rtrue, rprec, (double)dim, (double)phase};
    W->trace.insert(W->trace.end(), row, row + kTraceCols);
  }
  static double l2v(const std::vector<double>& v) {
    double q = 0.0;
    for (double e : v) q += e * e;
    return std::sqrt(q);
  }

  // the FMM operator on all R right-hand sides (one multi-RHS matvec)
  int approx_op(const double* in_tree, double* out_tree) {
    return fmm_launch(P, in_tree, 1, 1, selfi_orig, out_tree, 1, 1, W->stream, R);
  }
  // the operator of the residual updates: the FMM with its M2L in split FP16
  // (FP32-level accuracy; on a 16-bit plan through its split tables), the
  // plan's own form otherwise
  int update_op(const double* in_tree, double* out_tree) {
    const bool sixteen = P->m2l_precision == 1 || P->m2l_precision == 2;
    if (sixteen && P->t_m2l_h3) P->m2l_prec_override = 3;
    const int rc = approx_op(in_tree, out_tree);
    P->m2l_prec_override = -1;
    return rc;
  }
  bool update_op_ok() const {
    const bool sixteen = P->m2l_precision == 1 || P->m2l_precision == 2;
    return !sixteen || P->n_m2l_tiles == 0 || P->t_m2l_h3 != nullptr;
  }
  int accurate_op(const double* in_tree, double* out_tree) {
    if (R > 1) return accurate_op_multi(in_tree, out_tree);
    // block diagonal over the geometries of a batched plan: one symmetric-
    // form launch covers every geometry (schedule cached on the plan)
    if (!P->dense_sym) {
      DenseSymSchedule S;
      dense_sym_schedule(P->host.geom_pt_off, S);
      auto* D = new DenseSymDevice();
      const int rc = dense_sym_upload(S, *D);
      if (rc != ED_OK) {
        delete D;
        return rc;
      }
      P->dense_sym = D;
    }
    for (int r = 0; r < R; ++r) {
      const size_t o1 = (size_t)r * n1;
      ED_T(dense_sym_run(*P->dense_sym, P->cent64_t, P->area64_t, in_tree + o1, W->selfi_t,
                         out_tree + o1, W->stream));
      if (P->n_near > 0)
        hipLaunchKernelGGL(k_near_csr, dim3(gn1), dim3(256), 0, W->stream, n1, P->near_ptr,
                           P->near_col, P->near_w64, in_tree + o1, out_tree + o1);
    }
    ED_KCHECK("near csr");
    return ED_OK;
  }
  // R right-hand sides: the split form evaluates each pair once for all of
  // them (one r^2 and rsqrt, R FMAs), where the symmetric form would run R
  // times; then the FP64 near correction per right-hand side
  int accurate_op_multi(const double* in_tree, double* out_tree) {
    if (!W->dn_items) {
      std::vector<int64_t> geo = P->host.geom_pt_off;
      if (geo.size() < 2) geo = {0, n1};
      std::vector<DenseItem> items;
      std::vector<int32_t> rns;
      int smax = 1;
      dense_schedule(geo, items, rns, &smax);
      ED_G(hipMalloc((void**)&W->dn_items, sizeof(DenseItem) * std::max<size_t>(items.size(), 1)));
      ED_G(hipMalloc((void**)&W->dn_rns, sizeof(int32_t) * n1));
      ED_G(hipMalloc((void**)&W->dn_part, sizeof(double) * (size_t)smax * R * n1));
      if (!items.empty())
        ED_G(hipMemcpy(W->dn_items, items.data(), sizeof(DenseItem) * items.size(),
                       hipMemcpyHostToDevice));
      ED_G(hipMemcpy(W->dn_rns, rns.data(), sizeof(int32_t) * n1, hipMemcpyHostToDevice));
      W->dn_n_items = (int64_t)items.size();
      W->dn_smax = smax;
    }
    ED_T(dense_run<double>(W->dn_items, W->dn_n_items, W->dn_rns, W->dn_smax, P->cent64_t,
                           P->area64_t, in_tree, W->selfi_t, out_tree, n1, R, W->dn_part,
                           W->stream));
    if (P->n_near > 0)
      for (int r = 0; r < R; ++r) {
        const size_t o1 = (size_t)r * n1;
        hipLaunchKernelGGL(k_near_csr, dim3(gn1), dim3(256), 0, W->stream, n1, P->near_ptr,
                           P->near_col, P->near_w64, in_tree + o1, out_tree + o1);
      }
    ED_KCHECK("near csr");
    return ED_OK;
  }
  bool use_sai() const { return o->precond == 2; }
  bool general() const { return P->coarse_k > 0 || use_sai(); }
  int mode() const { return (use_sai() ? 4 : 0) + (o->precond ? 2 : 0) + (P->coarse_k > 0 ? 1 : 0); }
  // z = M^-1 u in place (coarse-corrected Jacobi), per right-hand side;
  // u may alias z
  void coarse_apply(const double* u, double* z) {
    const int k = P->coarse_k;
    const unsigned gb1 = (unsigned)std::min<int64_t>((n1 + 255) / 256, 4096);
    for (int r = 0; r < R; ++r) {
      const size_t o1 = (size_t)r * n1;
      hipLaunchKernelGGL(k_coarse_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, W->stream, n1, k,
                         P->coarse_z, u + o1, P->coarse_part);
      hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(256), 0, W->stream, k, P->coarse_part,
                         P->coarse_einv, P->coarse_cy);
      hipLaunchKernelGGL(k_coarse_apply, dim3(gb1), dim3(256), 0, W->stream, n1, k, u + o1,
                         W->dinv + o1, P->coarse_z, P->coarse_az, P->coarse_cy, z + o1);
    }
  }
  // z = M^-1 u with the SAI smoother, M^-1 u = S (u - AZ c) + Z c, per
  // right-hand side; u is overwritten (in-place subtraction) and must not
  // alias z
  void sai_precond(double* u, double* z) {
    const int k = P->coarse_k;
    const unsigned gb1 = (unsigned)std::min<int64_t>((n1 + 255) / 256, 4096);
    for (int r = 0; r < R; ++r) {
      const size_t o1 = (size_t)r * n1;
      if (k > 0) {
        hipLaunchKernelGGL(k_coarse_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, W->stream, n1,
                           k, P->coarse_z, u + o1, P->coarse_part);
        hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(256), 0, W->stream, k, P->coarse_part,
                           P->coarse_einv, P->coarse_cy);
        hipLaunchKernelGGL(k_coarse_sub, dim3(gb1), dim3(256), 0, W->stream, n1, k, P->coarse_az,
                           P->coarse_cy, u + o1);
      }
      sai_apply_launch(P, u + o1, k, P->coarse_z, P->coarse_cy, z + o1, W->stream);
    }
  }
  // active tiling: per (right-hand side, geometry) unless a coarse space
  // couples the geometries
  int tk() const { return P->coarse_k > 0 ? 1 : 0; }
  int G() const { return W->G_k[tk()]; }
  int NT() const { return W->NT_k[tk()]; }
  const int4* tiles() const { return W->tiles_k[tk()]; }
  const int* seg_t0() const { return W->seg_t0_k[tk()]; }

  // per-segment norms of (b - t) and M^-1 (b - t) into segv rows SV_NRM_R /
  // SV_NRM_Z and the host vectors nr, nz; r, z optional outputs
  int resid_norms(const double* b, const double* t, double* r, double* z, std::vector<double>& nr,
                  std::vector<double>& nz) {
    const bool gen = general() && z != nullptr;
    double* rr = (gen && r == nullptr) ? W->vin : r;
    const int nt = NT(), g = G(), ld = W->Gmax;
    hipLaunchKernelGGL(k_resid_seg, dim3(nt), dim3(kRedThreads), 0, W->stream, tiles(), nt, b, t,
                       W->dinv, rr, gen ? nullptr : z, W->partial);
    hipLaunchKernelGGL(k_finish_seg_norms, dim3(g), dim3(256), 0, W->stream, W->partial, nt, 2,
                       seg_t0(), W->segv + SV_NRM_R * ld, ld);
    if (gen) {
      if (use_sai()) {
        ED_G(hipMemcpyAsync(W->u2, rr, sizeof(double) * n, hipMemcpyDeviceToDevice, W->stream));
        sai_precond(W->u2, z);
      } else {
        coarse_apply(rr, z);
      }
      hipLaunchKernelGGL(k_norm2_seg, dim3(nt), dim3(kRedThreads), 0, W->stream, tiles(), z,
                         W->partial);
      hipLaunchKernelGGL(k_finish_seg_norms, dim3(g), dim3(256), 0, W->stream, W->partial, nt, 1,
                         seg_t0(), W->segv + SV_NRM_Z * ld, ld);
    }
    ED_KCHECK("resid");
    ED_G(hipMemcpyAsync(W->hsegv, W->segv, sizeof(double) * 2 * ld, hipMemcpyDeviceToHost,
                        W->stream));
    ED_G(hipStreamSynchronize(W->stream));
    nr.assign(W->hsegv + SV_NRM_R * ld, W->hsegv + SV_NRM_R * ld + g);
    nz.assign(W->hsegv + SV_NRM_Z * ld, W->hsegv + SV_NRM_Z * ld + g);
    return ED_OK;
  }
  int arnoldi_step() {
    const int m = W->m, nt = NT(), g = G(), ld = W->Gmax;
    if (use_sai()) {
      ED_T(approx_op(W->vin, W->u2));
      sai_precond(W->u2, W->w);
    } else {
      ED_T(approx_op(W->vin, W->w));
      if (P->coarse_k > 0) coarse_apply(W->w, W->w);
    }
    hipLaunchKernelGGL(k_dots_seg, dim3(nt), dim3(kRedThreads), 0, W->stream, tiles(), nt, n,
                       general() ? nullptr : W->dinv, W->w, W->V, W->dj, W->partial);
    hipLaunchKernelGGL(k_finish_dots_seg, dim3(g, m + 1), dim3(256), 0, W->stream, W->partial,
                       nt, seg_t0(), W->dj, W->h, m + 2);
    hipLaunchKernelGGL(k_update_seg, dim3(nt), dim3(kRedThreads), 0, W->stream, tiles(), nt, n,
                       W->w, W->V, W->dj, W->h, m + 2, 0, W->partial);
    // second Gram-Schmidt pass: coefficients h2
    hipLaunchKernelGGL(k_finish_dots_seg, dim3(g, m + 1), dim3(256), 0, W->stream, W->partial,
                       nt, seg_t0(), W->dj, W->h2, m + 2);
    hipLaunchKernelGGL(k_update_seg, dim3(nt), dim3(kRedThreads), 0, W->stream, tiles(), nt, n,
                       W->w, W->V, W->dj, W->h2, m + 2, 1, W->partial);
    hipLaunchKernelGGL(k_givens_seg, dim3(1), dim3(256), 0, W->stream, g, m, W->partial, seg_t0(),
                       W->h, W->h2, W->H, W->g, W->cs, W->sn, W->segv, ld, W->scal, W->dj);
    hipLaunchKernelGGL(k_scale_next_seg, dim3(nt), dim3(256), 0, W->stream, tiles(), n, W->w,
                       W->V, W->vin, W->dj, W->segv, ld);
    ED_KCHECK("arnoldi");
    return ED_OK;
  }

  // Restarted GMRES on approx_op for A d = rhs (tree order), x updated in
  // place, independently per segment: segment g stops at
  // |rhs_g - A x_g| <= tol_rel[g] |rhs_g - A x0_g| (tol_rel[g] >= 1 leaves it
  // untouched).  The code is 0 when every segment converged; r0 / r1 return
  // the per-segment initial and final true residual norms.
  int gmres(const double* rhs, double* x, const std::vector<double>& tol_rel, int maxiter,
            int* code, std::vector<double>* r0_out, std::vector<double>* r1_out) {
    const int m = W->m, g = G(), ld = W->Gmax, nt = NT();
    std::vector<double> nr, nz;
    ED_T(approx_op(x, W->t));
    ED_T(resid_norms(rhs, W->t, W->r, W->w, nr, nz));
    trace(kTraceInitial, 0, l2v(nr), l2v(nz));
    if (r0_out) *r0_out = nr;
    auto finish = [&](int c) {
      *code = c;
      if (r1_out) *r1_out = nr;
      return ED_OK;
    };
    std::vector<double> tol_abs(g);
    bool all_ok = true;
    for (int k = 0; k < g; ++k) {
      if (!std::isfinite(nr[k])) return finish(3);
      tol_abs[k] = std::max(tol_rel[k] * (nr[k] > 1e-300 ? nr[k] : 1.0), 1e-300);
      all_ok = all_ok && nr[k] <= tol_abs[k];
    }
    if (all_ok) return finish(0);
    int local = 0;
    bool use_graph = o->use_graph != 0;
    // target of each segment's preconditioned estimate |g_{j+1}|: tol * |M^-1 r0|
    std::vector<double> est_tol(g);
    for (int k = 0; k < g; ++k) est_tol[k] = tol_rel[k] * nz[k];
    while (local < maxiter) {
      bool any = false;
      for (int k = 0; k < g; ++k) {
        if (!std::isfinite(nz[k])) return finish(7);
        const bool act = nr[k] > tol_abs[k] && nz[k] > 0.0;
        W->hsegv[SV_EST_TOL * ld + k] = est_tol[k];
        W->hsegv[SV_ACTIVE * ld + k] = act ? 1.0 : 0.0;
        any = any || act;
      }
      if (!any) return finish(6);  // unconverged segments with a zero preconditioned residual
      ED_G(hipMemcpyAsync(W->segv + SV_EST_TOL * ld, W->hsegv + SV_EST_TOL * ld,
                          sizeof(double) * 2 * ld, hipMemcpyHostToDevice, W->stream));
      hipLaunchKernelGGL(k_start_seg, dim3(nt), dim3(256), 0, W->stream, tiles(), g, m, W->w, W->V,
                         W->vin, W->g, W->segv, ld, W->dj);
      ED_G(hipMemsetAsync(W->scal + S_BREAK, 0, sizeof(double), W->stream));
      ED_KCHECK("start");
      int j = 0;
      bool stop = false;
      const int ce = std::max(1, o->check_every);
      while (j < m && local < maxiter && !stop) {
        const int steps = std::min(std::min(ce, m - j), maxiter - local);
        for (int s = 0; s < steps; ++s) {
          if (use_graph) {
            if (!W->gexec || W->graph_selfi != selfi_orig || W->graph_mode != mode()) {
              if (W->gexec) { (void)hipGraphExecDestroy(W->gexec); W->gexec = nullptr; }
              if (W->graph) { (void)hipGraphDestroy(W->graph); W->graph = nullptr; }
              ED_G(hipStreamBeginCapture(W->stream, hipStreamCaptureModeThreadLocal));
              int rc = arnoldi_step();
              hipGraph_t gr = nullptr;
              hipError_t ec = hipStreamEndCapture(W->stream, &gr);
              if (rc != ED_OK) return rc;
              if (ec != hipSuccess) return fail(ED_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
              W->graph = gr;
              ED_G(hipGraphInstantiate(&W->gexec, gr, nullptr, nullptr, 0));
              W->graph_selfi = selfi_orig;
              W->graph_mode = mode();
            }
            ED_G(hipGraphLaunch(W->gexec, W->stream));
          } else {
            ED_T(arnoldi_step());
          }
        }
        j += steps;
        local += steps;
        total_iters += steps;
        ED_G(hipMemcpyAsync(W->hscal, W->scal, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, W->stream));
        ED_G(hipStreamSynchronize(W->stream));
        const double ratio = W->hscal[S_EST];
        // the estimate of |M^-1 r| (one block: exact; several: the worst
        // block's ratio to its target, times the first block's target)
        trace(kTracePoll, j, NAN, ratio * est_tol[0]);
        if (W->hscal[S_BREAK] != 0.0) stop = true;
        if (!std::isfinite(ratio)) return finish(5);
        if (ratio <= 1.0) {
          if (first_below < 0) first_below = total_iters;
          stop = true;
        }
      }
      // with a breakdown at step j the last column is still valid (hn = 0 only)
      hipLaunchKernelGGL(k_backsolve_seg, dim3(std::max(1, (g + 63) / 64)), dim3(64), 0,
                         W->stream, g, m, W->H, W->g, W->dj, W->h);
      hipLaunchKernelGGL(k_xupdate_seg, dim3(nt), dim3(256), 0, W->stream, tiles(), n, m, x, W->V,
                         W->h, W->dj);
      ED_KCHECK("update");
      ++cycles;
      ED_T(approx_op(x, W->t));
      ED_T(resid_norms(rhs, W->t, W->r, W->w, nr, nz));
      trace(kTraceCycleEnd, j, l2v(nr), l2v(nz));
      all_ok = true;
      for (int k = 0; k < g; ++k) {
        if (!std::isfinite(nr[k])) return finish(6);
        all_ok = all_ok && nr[k] <= tol_abs[k];
      }
      if (all_ok) return finish(0);
      if (W->hscal[S_BREAK] != 0.0 && j <= 1) return finish(2);
    }
    return finish(1);
  }
};

}  // namespace ed

using namespace ed;

extern "C" void ed_gmres_opts_default(ed_gmres_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->restart = 50;
  o->maxiter = 2000;
  o->tol = 1e-8;
  o->precond = 1;
  o->refine = 1;
  o->max_refine = 8;
  o->inner_tol = 1e-4;
  o->use_graph = 0;  // direct launches: measured faster than graph replay (5 k - 300 k panels)
  o->check_every = 4;
  o->residual_update = 1;
}

// R right-hand sides stacked (R x n, original order); every (right-hand
// side, geometry) segment runs its own Krylov space and stops at tol times its
// own |b| block, and all of them share each operator launch.
static int gmres_solve(ed_plan* P, const double* b, double* x, int R, const double* diag,
                       const double* selfi, const ed_gmres_opts* opts, ed_gmres_info* info,
                       ed_stream stream) {
  ed_gmres_opts o;
  ed_gmres_opts_default(&o);
  if (opts) o = *opts;
  if (o.restart <= 0 || o.restart > 255) return fail(ED_ERR_ARG, "restart must be in [1, 255]");
  if (o.maxiter <= 0) return fail(ED_ERR_ARG, "maxiter must be positive");
  if (o.precond < 0 || o.precond > 2) return fail(ED_ERR_ARG, "precond must be 0, 1 or 2");
  if (o.precond == 2 && !P->sai_cnt)
    return fail(ED_ERR_ARG, "precond 2 needs the sparse approximate inverse (ed_plan_build_sai)");
  std::memset(info, 0, sizeof(*info));
  info->first_iter_below_tol = -1;
  if (R > 1) ED_T(multi_reserve(P, R));
  GmresWS* W = nullptr;
  ED_T(ws_get(P, o.restart, R, &W));
  W->last_tiling = -1;
  W->trace.clear();
  const int64_t n1 = P->n, n = W->n;
  const unsigned gb = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  const unsigned gn = (unsigned)((n + 255) / 256), gn1 = (unsigned)((n1 + 255) / 256);
  hipStream_t caller = (hipStream_t)stream;
  // order the solver's stream after the caller's work
  hipEvent_t ev;
  ED_G(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  ED_G(hipEventRecord(ev, caller));
  ED_G(hipStreamWaitEvent(W->stream, ev, 0));
  for (int r = 0; r < R; ++r) {
    const size_t o1 = (size_t)r * n1;
    hipLaunchKernelGGL(k_perm_in, dim3(gn1), dim3(256), 0, W->stream, n1, P->perm, b + o1, W->b + o1);
    hipLaunchKernelGGL(k_perm_in, dim3(gn1), dim3(256), 0, W->stream, n1, P->perm, x + o1, W->x + o1);
    hipLaunchKernelGGL(k_dinv, dim3(gn1), dim3(256), 0, W->stream, n1, P->perm,
                       o.precond ? diag : nullptr, W->dinv + o1);
  }
  if (selfi)
    hipLaunchKernelGGL(k_perm_in, dim3(gn1), dim3(256), 0, W->stream, n1, P->perm, selfi, W->selfi_t);
  else
    ED_G(hipMemsetAsync(W->selfi_t, 0, sizeof(double) * n1, W->stream));
  ED_KCHECK("solver setup");
  Solver S{P, W, &o, selfi, n, gb, n1, R, gn1};
  const int G = S.G();
  int code = 1;
  std::vector<double> bn, dummy;
  ED_T(S.resid_norms(W->b, nullptr, nullptr, nullptr, bn, dummy));
  auto l2 = [](const std::vector<double>& v) {
    double q = 0.0;
    for (double e : v) q += e * e;
    return std::sqrt(q);
  };
  const double bn_all = l2(bn);
  if (!o.refine) {
    std::vector<double> r0, r1;
    ED_T(S.gmres(W->b, W->x, std::vector<double>(G, o.tol), o.maxiter, &code, &r0, &r1));
    info->resid0 = l2(r0);
    info->resid = l2(r1);
    W->last_rn = r1;
    info->tol_abs = o.tol * (info->resid0 > 1e-300 ? info->resid0 : 1.0);
  } else {
    // outer FP64 iterative refinement with the accurate operator; segment g
    // is done at |r_g| <= tol |b_g|.
    // Residual updates (o.residual_update): once an accurate residual r is
    // small, the next correction d is small too (|A d| ~ |r|), so
    // r' = r - A_u d with an operator A_u of relative error e is within
    // ~e |r| of the true b - A x'.  e is measured on the iterate when the
    // switch is considered (|A x - A_u x| / |A x|, A_u = the FMM with a
    // split-FP16 M2L), the switch is taken when kUpdSafety e |r| <= tol / 4,
    // and every update adds kUpdSafety e |A_u d| to a per-segment bound; a
    // segment is done at |r_upd| + bound <= tol |b|.
    constexpr double kUpdSafety = 10.0;
    std::vector<double> rn, zn, tol_abs(G), eps(G, 0.0), bound(G, 0.0), an, dn, dummy2;
    for (int k = 0; k < G; ++k) tol_abs[k] = o.tol * (bn[k] > 1e-300 ? bn[k] : 1.0);
    info->tol_abs = o.tol * (bn_all > 1e-300 ? bn_all : 1.0);
    const bool upd_on = o.residual_update != 0 && S.update_op_ok();
    bool updating = false;  // rn holds an updated residual
    code = 1;
    // Residuals b - A_u x with the update operator (error ~eps |b|, eps taken
    // as kUpdPrior before it is measured) drive the refinement, not as a
    // convergence test, while they exceed 10 kUpdPrior |b| and still fall by
    // half a step; the accurate residual comes below that (or when the
    // decrease stalls).
    constexpr double kUpdPrior = 1e-5;
    std::vector<double> rn_prev(G, INFINITY);
    for (int it = 0; it <= o.max_refine; ++it) {
      // The first correction only has to bring the residual down by the inner
      // tolerance, so its right-hand side may come from the approximate
      // operator; a residual that already looks converged is re-measured with
      // the accurate one, and every later step (and the final test) uses it.
      // With residual updates, steps whose residual is far above the update
      // operator's error take b - A_u x, the one accurate residual comes when
      // it is small enough to switch to updates, and an initial guess whose
      // residual exceeds |b| is dropped for x = 0 (r = b exactly).
      bool approx_resid = false;
      S.refine_step = it;
      if (it == 0 || (upd_on && !updating)) {
        if (upd_on) ED_T(S.update_op(W->x, W->t));
        else ED_T(S.approx_op(W->x, W->t));
        ED_T(S.resid_norms(W->b, W->t, W->r, nullptr, rn, zn));
        if (it == 0) info->resid0 = l2(rn);
        if (it == 0 && upd_on && !(l2(rn) < bn_all)) {
          ED_G(hipMemsetAsync(W->x, 0, sizeof(double) * n, W->stream));
          ED_G(hipMemcpyAsync(W->r, W->b, sizeof(double) * n, hipMemcpyDeviceToDevice,
                              W->stream));
          rn = bn;
        }
        approx_resid = true;
        for (int k = 0; k < G; ++k) {
          const double floor_k =
              upd_on ? std::max({10.0 * kUpdPrior * bn[k], 100.0 * tol_abs[k],
                                 0.25 * tol_abs[k] / (kUpdSafety * kUpdPrior)})
                     : std::max(100.0 * tol_abs[k], 10.0 * o.inner_tol * bn[k]);
          approx_resid = approx_resid && std::isfinite(rn[k]) && rn[k] > floor_k &&
                         rn[k] < 0.5 * rn_prev[k];
        }
        if (approx_resid) {
          rn_prev = rn;
          S.trace(kTraceApprox, 0, l2(rn), NAN);
        }
        // keep A_u x for the error measurement below (d is free here)
        if (upd_on && !approx_resid)
          ED_G(hipMemcpyAsync(W->d, W->t, sizeof(double) * n, hipMemcpyDeviceToDevice, W->stream));
      }
      if (!approx_resid && !updating) {
        ED_T(S.accurate_op(W->x, W->t));
        ED_T(S.resid_norms(W->b, W->t, W->r, nullptr, rn, zn));
        S.trace(kTraceRefine, 0, l2(rn), NAN);
        ++info->n_accurate;
        bool conv = true;
        for (int k = 0; k < G; ++k) conv = conv && rn[k] <= tol_abs[k];
        if (upd_on && !conv && it < o.max_refine) {
          // error of the update operator on this iterate (A_u x is in d)
          ED_T(S.resid_norms(W->t, W->d, nullptr, nullptr, dn, dummy2));
          ED_T(S.resid_norms(W->t, nullptr, nullptr, nullptr, an, dummy2));
          bool sw = true;
          for (int k = 0; k < G; ++k) {
            eps[k] = an[k] > 1e-300 ? dn[k] / an[k] : 0.0;
            info->update_err = std::max(info->update_err, eps[k]);
            sw = sw && std::isfinite(eps[k]) &&
                 (rn[k] <= tol_abs[k] || kUpdSafety * eps[k] * rn[k] <= 0.25 * tol_abs[k]);
          }
          updating = sw;
        }
      }
      info->refine_steps = it;
      bool finite = true, done = true;
      for (int k = 0; k < G; ++k) {
        finite = finite && std::isfinite(rn[k]);
        done = done && rn[k] + bound[k] <= tol_abs[k];
      }
      if (!finite) { code = 6; break; }
      if (!approx_resid && done) { code = 0; break; }
      if (it == o.max_refine || S.total_iters >= o.maxiter) {
        if (approx_resid || updating) {  // report an accurate residual
          ED_T(S.accurate_op(W->x, W->t));
          ED_T(S.resid_norms(W->b, W->t, W->r, nullptr, rn, zn));
          ++info->n_accurate;
          std::fill(bound.begin(), bound.end(), 0.0);
        }
        code = 1;
        break;
      }
      // inner solve  A_fmm d = r with d = 0; the inner solver reuses r as its
      // own scratch, so the outer residual is kept in `rhs`
      ED_G(hipMemsetAsync(W->d, 0, sizeof(double) * n, W->stream));
      ED_G(hipMemcpyAsync(W->rhs, W->r, sizeof(double) * n, hipMemcpyDeviceToDevice, W->stream));
      // each segment's inner solve only has to take its residual down to its
      // target (0.25 leaves room for the approximate operator's own error);
      // converged segments are left alone (tolerance 1)
      std::vector<double> itol(G);
      for (int k = 0; k < G; ++k)
        itol[k] = rn[k] + bound[k] <= tol_abs[k]
                      ? 1.0
                      : std::min(0.5, std::max(o.inner_tol, 0.25 * tol_abs[k] / rn[k]));
      // an inner solve that stalls on the approximate operator's rounding
      // floor is cut after 4 restart cycles; the refinement continues from it
      const int inner_max = std::min(o.maxiter - S.total_iters, 4 * o.restart);
      int icode = 1;
      ED_T(S.gmres(W->rhs, W->d, itol, inner_max, &icode, nullptr, nullptr));
      if (icode == 3 || icode == 5 || icode == 6 || icode == 7) { code = icode; break; }
      hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(256), 0, W->stream, n, 1.0, W->d, W->x);
      ED_KCHECK("refine update");
      if (updating) {
        // r <- r - A_u d (the old residual is in rhs)
        ED_T(S.update_op(W->d, W->t));
        ED_T(S.resid_norms(W->rhs, W->t, W->r, nullptr, rn, zn));
        ED_T(S.resid_norms(W->t, nullptr, nullptr, nullptr, an, dummy2));
        for (int k = 0; k < G; ++k) bound[k] += kUpdSafety * eps[k] * an[k];
        ++info->n_updated;
        S.refine_step = it + 1;
        S.trace(kTraceUpdate, 0, l2(rn), l2(bound));
      }
    }
    info->resid_bound = l2(bound);
    info->resid = l2(rn);
    W->last_rn = rn;
  }
  W->last_bn = bn;
  W->last_tiling = S.tk();
  info->code = code;
  info->iters = S.total_iters;
  info->cycles = S.cycles;
  info->first_iter_below_tol = S.first_below;
  info->rel_resid = info->resid / (bn_all > 1e-300 ? bn_all : 1.0);
  for (int r = 0; r < R; ++r)
    hipLaunchKernelGGL(k_perm_out, dim3(gn1), dim3(256), 0, W->stream, n1, P->perm,
                       W->x + (size_t)r * n1, x + (size_t)r * n1);
  ED_KCHECK("solver output");
  ED_G(hipEventRecord(ev, W->stream));
  ED_G(hipStreamWaitEvent(caller, ev, 0));
  ED_G(hipStreamSynchronize(W->stream));
  (void)hipEventDestroy(ev);
  return ED_OK;
}

extern "C" int ed_gmres_trace(const ed_plan* P, double* rows, int32_t cap, int32_t* n_rows) {
  if (!P || !n_rows) return fail(ED_ERR_ARG, "ed_gmres_trace: NULL argument");
  const GmresWS* W = static_cast<const GmresWS*>(P->gmres_ws);
  const int32_t nr = W ? (int32_t)(W->trace.size() / kTraceCols) : 0;
  *n_rows = nr;
  if (rows && cap > 0 && W)
    std::memcpy(rows, W->trace.data(), sizeof(double) * kTraceCols * (size_t)std::min(cap, nr));
  return ED_OK;
}

extern "C" int ed_gmres_solve(ed_plan* P, const double* b, double* x, const double* diag,
                              const double* selfi, const ed_gmres_opts* opts,
                              ed_gmres_info* info, ed_stream stream) {
  if (!P || !b || !x || !info) return fail(ED_ERR_ARG, "ed_gmres_solve: NULL argument");
  return gmres_solve(P, b, x, 1, diag, selfi, opts, info, stream);
}

extern "C" int ed_gmres_block_resid(ed_plan* P, double* resid, double* bnorm, int32_t cap,
                                    int32_t* n_blocks, int32_t* per_geometry) {
  if (!P || !n_blocks) return fail(ED_ERR_ARG, "ed_gmres_block_resid: NULL argument");
  const GmresWS* W = (const GmresWS*)P->gmres_ws;
  if (!W || W->last_tiling < 0 || W->last_rn.size() != W->last_bn.size())
    return fail(ED_ERR_ARG, "ed_gmres_block_resid: no solve has run on this plan");
  const int nb = (int)W->last_rn.size();
  *n_blocks = nb;
  if (per_geometry) *per_geometry = W->last_tiling == 0 ? 1 : 0;
  for (int k = 0; k < nb && k < cap; ++k) {
    if (resid) resid[k] = W->last_rn[k];
    if (bnorm) bnorm[k] = W->last_bn[k];
  }
  return ED_OK;
}

extern "C" int ed_gmres_solve_multi(ed_plan* P, const double* b, double* x, int32_t nrhs,
                                    const double* diag, const double* selfi,
                                    const ed_gmres_opts* opts, ed_gmres_info* info,
                                    ed_stream stream) {
  if (!P || !b || !x || !info) return fail(ED_ERR_ARG, "ed_gmres_solve_multi: NULL argument");
  if (nrhs < 1 || nrhs > 8) return fail(ED_ERR_ARG, "ed_gmres_solve_multi: nrhs must be in [1, 8]");
  return gmres_solve(P, b, x, nrhs, diag, selfi, opts, info, stream);
}

// Coarse space for the native solver (no reference counterpart).  Z holds k
// columns of length n (column-major, original panel order).  The columns are
// orthonormalised (CGS2), A Z is formed with the plan's FMM operator (with
// `self_integrals`), E = Z^T A Z is inverted on the device.  k = 0 removes it.
extern "C" int ed_plan_set_coarse(ed_plan* P, const double* Z, int32_t k,
                                  const double* self_integrals, ed_stream stream) {
  using namespace ed;
  if (!P || k < 0 || k > kMaxCoarse || (k > 0 && !Z))
    return fail(ED_ERR_ARG, "ed_plan_set_coarse: k must be in [0, 32] with Z given");
  gmres_invalidate_graph(P);  // a captured Arnoldi graph refers to the old coarse arrays
  void* olds[] = {P->coarse_z, P->coarse_az, P->coarse_einv, P->coarse_cy, P->coarse_part,
                  P->coarse_tmp};
  for (void* q : olds)
    if (q) (void)hipFree(q);
  P->coarse_z = P->coarse_az = P->coarse_einv = P->coarse_cy = P->coarse_part = nullptr;
  P->coarse_tmp = nullptr;
  P->coarse_k = 0;
  if (k == 0) return ED_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = P->n;
  const unsigned gn = (unsigned)((n + 255) / 256);
  ED_G(hipMalloc((void**)&P->coarse_z, sizeof(double) * n * k));
  ED_G(hipMalloc((void**)&P->coarse_az, sizeof(double) * n * k));
  ED_G(hipMalloc((void**)&P->coarse_einv, sizeof(double) * k * k));
  ED_G(hipMalloc((void**)&P->coarse_cy, sizeof(double) * kMaxCoarse));
  ED_G(hipMalloc((void**)&P->coarse_part, sizeof(double) * kRedBlocks * k * k));
  ED_G(hipMalloc((void**)&P->coarse_tmp, sizeof(double) * n));
  double* scal = nullptr;
  ED_G(hipMalloc((void**)&scal, sizeof(double) * 8));
  for (int c = 0; c < k; ++c)
    hipLaunchKernelGGL(k_perm_in, dim3(gn), dim3(256), 0, st, n, P->perm, Z + (size_t)c * n,
                       P->coarse_z + (size_t)c * n);
  for (int c = 0; c < k; ++c) {
    double* zc = P->coarse_z + (size_t)c * n;
    for (int pass = 0; pass < 2 && c > 0; ++pass) {
      hipLaunchKernelGGL(k_coarse_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, st, n, c,
                         P->coarse_z, zc, P->coarse_part);
      hipLaunchKernelGGL(k_coarse_coef, dim3(1), dim3(256), 0, st, c, P->coarse_part, P->coarse_cy);
      hipLaunchKernelGGL(k_coarse_orth, dim3(gn), dim3(256), 0, st, n, c, P->coarse_z, P->coarse_cy);
    }
    hipLaunchKernelGGL(k_norm2_part, dim3(kRedBlocks), dim3(kRedThreads), 0, st, n, zc,
                       P->coarse_part);
    hipLaunchKernelGGL(k_finish_norms, dim3(1), dim3(256), 0, st, P->coarse_part, 1, scal, 0);
    hipLaunchKernelGGL(k_scale_col, dim3(gn), dim3(256), 0, st, n, zc, scal, 0);
  }
  ED_KCHECK("coarse orthonormalisation");
  for (int c = 0; c < k; ++c) {
    int rc = fmm_launch(P, P->coarse_z + (size_t)c * n, 1, 1, self_integrals,
                        P->coarse_az + (size_t)c * n, 1, 1, st);
    if (rc != ED_OK) { (void)hipFree(scal); return rc; }
  }
  for (int c = 0; c < k; ++c)
    hipLaunchKernelGGL(k_coarse_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, st, n, k,
                       P->coarse_z, P->coarse_az + (size_t)c * n,
                       P->coarse_part + (size_t)c * k * kRedBlocks);
  hipLaunchKernelGGL(k_coarse_einv, dim3(1), dim3(64), 0, st, k, P->coarse_part, P->coarse_einv);
  ED_KCHECK("coarse operator");
  ED_G(hipStreamSynchronize(st));
  (void)hipFree(scal);
  P->coarse_k = k;
  return ED_OK;
}

// Coarse (Petrov-Galerkin) solve on span(Z): out = Z E^-1 Z^T u, E = Z^T A Z,
// original panel order (no reference counterpart).  With u = b it is the
// initial guess whose residual is orthogonal to span(Z).
extern "C" int ed_plan_coarse_solve(ed_plan* P, const double* u, double* out, ed_stream stream) {
  using namespace ed;
  if (!P || !u || !out) return fail(ED_ERR_ARG, "ed_plan_coarse_solve: NULL argument");
  if (P->coarse_k <= 0) return fail(ED_ERR_ARG, "ed_plan_coarse_solve: no coarse space set");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = P->n;
  const int k = P->coarse_k;
  const unsigned gn = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_perm_in, dim3(gn), dim3(256), 0, st, n, P->perm, u, P->coarse_tmp);
  hipLaunchKernelGGL(k_coarse_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, st, n, k,
                     P->coarse_z, P->coarse_tmp, P->coarse_part);
  hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(256), 0, st, k, P->coarse_part,
                     P->coarse_einv, P->coarse_cy);
  hipLaunchKernelGGL(k_coarse_combine, dim3(gn), dim3(256), 0, st, n, k, P->coarse_z,
                     P->coarse_cy, P->coarse_tmp);
  hipLaunchKernelGGL(k_perm_out, dim3(gn), dim3(256), 0, st, n, P->perm, P->coarse_tmp, out);
  ED_KCHECK("coarse solve");
  return ED_OK;
}
