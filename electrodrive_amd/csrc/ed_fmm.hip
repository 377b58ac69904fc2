// FMM matvec kernels for gfx950 (CDNA4, wave64).
//
// Data layout (tree order, Morton-sorted):
//   pts[i]       float4 (x, y, z, q) relative to the root centre, q = sigma A
//   M, L         n_nodes x nc (real-packed solid-harmonic coefficients)
// Stages of one matvec (one launch each, a level loop for M2M):
//   gather      q, sigma, self term into tree order
//   p2m         one wave per leaf: lanes stream points, LDS column reduction
//   m2m         per level: one block per parent, sum of T_oct * M_child
//   m2l         MFMA GEMM per offset table: Y[pair] = T_off * M[src] (FP32,
//               split FP16 or 16-bit MFMA), then a deterministic per-target sum
//               into each node's own part of its local expansion
//   down        one block per node: the L2L recursion over its own chain of
//               ancestors (all levels in one launch)
//   p2p         one wave per work item, partial sums per item: for one
//               right-hand side every unordered pair of distinct U-list
//               leaves once for both directions (symmetric items) plus a
//               leaf with itself; for several, <= 64 targets x a piece of
//               their leaf's U-list source stream
//   l2p_out     one thread per panel: L2P + P2P partials + near CSR + self
//               term, K_E scaling, scatter to the output order
// Several right-hand sides (up to 8) run through the same launches with the
// RHS as blockIdx.y (RhsStride); P2P then evaluates each r_ij once for all.
// No float sums by atomics: every sum has a fixed order.  (The one atomic is
// an unsigned max, whose result does not depend on order.)
#include <hip/hip_runtime.h>

#include <cmath>

#include "ed_plan.h"

namespace ed {

constexpr int kLeafBlock = 64;
// slots of the running multipole maximum (see fold_absmax)
constexpr int kMaxSlots = 64, kMaxSlotStride = 32;  // 32 x 4 B = 128 B

// The far-field passes up to the M2L GEMM raise their waves' issue priority:
// P2P runs beside them on a side stream, and its VALU-bound waves would
// otherwise take most issue cycles of every SIMD the two share, stretching
// the dependent chain.  The per-node reduce and the downward pass keep the
// default priority: by then the chain runs ahead of the P2P, which then sets
// the matvec (at 82k, 88.6 us with them at the default, 90.3 us raised).
__device__ __forceinline__ void chain_prio() { __builtin_amdgcn_s_setprio(3); }

// Several right-hand sides in one matvec: blockIdx.y is the RHS of every
// far-field launch and each per-RHS array is offset by these strides (all
// zero, R = 1 and qT = NULL for one right-hand side, whose charges sit in
// pts[i].w).
struct RhsStride {
  int64_t node = 0;  // M, L: n_nodes * nc
  int64_t pair = 0;  // ypair: n_pairs * ncp
  int64_t slot = 0;  // multipole-maximum slots
  int64_t part = 0;  // P2P partial rows
  int64_t n = 0;     // tree-order sigma and outputs
  const float* qT = nullptr;  // n x R charges, point-major
  int R = 1;
};

// Axis reflections.  A reflection R_s of the lattice axes (bit a of s: axis a
// negated) maps a translation operator T(o) to T(R_s o) = D T(o) D, D =
// diag(+-1), with coefficient k flipped when popcount(s & flip[k]) is odd
// (the real solid-harmonic coefficients are even or odd under each axis
// reflection): M2L offsets, M2M and L2L octants all use it.  A sign flip is
// exact.
__device__ __forceinline__ float refl(float v, int s, int fk) {
  return __int_as_float(__float_as_int(v) ^ ((__popc((unsigned)(s & fk)) & 1) << 31));
}

// ---------------------------------------------------------------------------
// Solid-harmonic streaming helpers (m-major recurrences; see ed_operators.cpp)
// ---------------------------------------------------------------------------
template <int P>
__device__ __forceinline__ void p2m_accum(float x, float y, float z, float q, float* Mr) {
  const float r2 = x * x + y * y + z * z;
  float pmr = 1.0f, pmi = 0.0f;
#pragma unroll
  for (int m = 0; m <= P; ++m) {
    if (m > 0) {
      const float f = -0.5f / (float)m;
      const float nr = f * (x * pmr - y * pmi);
      const float ni = f * (x * pmi + y * pmr);
      pmr = nr;
      pmi = ni;
    }
    float ar = 0.0f, ai = 0.0f, br = pmr, bi = pmi;  // R_{n-1}^m, R_n^m
#pragma unroll
    for (int n = m; n <= P; ++n) {
      if (m == 0) {
        Mr[n * n] += q * br;
      } else {
        Mr[n * n + 2 * m - 1] += q * br;
        Mr[n * n + 2 * m] -= q * bi;
      }
      if (n < P) {
        const float inv = 1.0f / (float)((n + 1) * (n + 1) - m * m);
        const float c = (float)(2 * n + 1) * z;
        const float nr = (c * br - r2 * ar) * inv;
        const float ni = (c * bi - r2 * ai) * inv;
        ar = br; ai = bi; br = nr; bi = ni;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// gather: tree-order q, sigma, self term; resets the multipole maximum
// ---------------------------------------------------------------------------
template <typename TS>
__global__ void k_gather(int n, const int32_t* __restrict__ perm, const TS* __restrict__ sigma,
                         int sigma_tree, const TS* __restrict__ selfi, float4* __restrict__ pts,
                         const float* __restrict__ area_t, float* __restrict__ sig_t,
                         float* __restrict__ selfc_t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  const float s = (float)sigma[sigma_tree ? i : o];
  sig_t[i] = s;
  pts[i].w = s * area_t[i];
  selfc_t[i] = selfi ? (float)selfi[o] : 0.0f;
}

// Near-field correction of each panel for one RHS (tree order):
// near_t[i] = sum of w_e sigma_t[col_e] over row i of the near CSR.  Four
// lanes a panel: lane r takes entries b + r, b + r + 4, ..., two of them in
// flight, and the four partial sums are added as (r0 + r1) + (r2 + r3).  It
// runs on the caller's stream after the far-field chain while the P2P
// finishes on the side stream, so it reads sigma through the permutation
// itself (the same float the gather writes) instead of waiting for the
// gather's output.
template <typename TS>
__global__ __launch_bounds__(256) void k_near(int n, const int32_t* __restrict__ near_ptr,
                                              const int32_t* __restrict__ near_col,
                                              const float* __restrict__ near_w,
                                              const int32_t* __restrict__ perm,
                                              const TS* __restrict__ sigma, int sigma_tree,
                                              float* __restrict__ near_t) {
  const int t = blockIdx.x * 256 + (int)threadIdx.x;
  const int i = min(t >> 2, n - 1), r = t & 3;
  const int b = near_ptr[i], e = near_ptr[i + 1];
  float acc = 0.0f;
  for (int k = b + r; k < e; k += 8) {
    const int k2 = min(k + 4, e - 1);
    const int c1 = near_col[k], c2 = near_col[k2];
    const float w1 = near_w[k], w2 = k + 4 < e ? near_w[k2] : 0.0f;
    const float s1 = (float)sigma[sigma_tree ? c1 : perm[c1]];
    const float s2 = (float)sigma[sigma_tree ? c2 : perm[c2]];
    acc += w1 * s1;
    acc += w2 * s2;
  }
  acc += __shfl_xor(acc, 1, 64);
  acc += __shfl_xor(acc, 2, 64);
  if (r == 0 && (t >> 2) < n) near_t[i] = acc;
}

// Largest |M_k * ascale_k| of the multipoles a wave wrote, folded into one
// of kMaxSlots running maxima of the matvec (the 16-bit M2L takes the
// largest; k_m2l_reduce resets them for the next matvec once the GEMM has
// read them, and plans start with them zero) as the bit pattern of a non-negative float:
// unsigned order is float order there, so the result does not depend on the
// order of the atomics.  Slots sit a cache line apart, so the atomics of
// different waves do not queue on one address.  NaN entries are skipped, as
// fmaxf would.  `bits` is NULL unless the M2L runs in a 16-bit form.
__device__ __forceinline__ void fold_absmax(float v, unsigned* bits, int slot) {
  v = v == v ? fabsf(v) : 0.0f;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0 && v > 0.0f)
    atomicMax(bits + (slot & (kMaxSlots - 1)) * kMaxSlotStride, __float_as_uint(v));
}

// The largest of the slots (every lane of the calling wave gets it).
__device__ __forceinline__ float slots_absmax(const unsigned* __restrict__ bits) {
  float v = __uint_as_float(bits[(threadIdx.x & (kMaxSlots - 1)) * kMaxSlotStride]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// y = T x for the NC outputs owned by this thread (k = t + W u), T column
// major (T[j NC + k]: input j, output k), x in LDS.  Each global round trip
// costs about half a microsecond here, so the table column is loaded in as
// few batches as the registers allow (one for NC <= 128 with one output per
// thread): every load of a batch is issued before the first use.  Indices
// are clamped instead of branched on.
template <int NC, int W>
__device__ __forceinline__ void op_apply(const float* __restrict__ T, const float* x, int t,
                                         float (&y)[(NC + W - 1) / W]) {
  constexpr int KO = (NC + W - 1) / W;
  constexpr int B = (128 / KO) < NC ? (128 / KO) : NC;  // rows per batch
  int kk[KO];
#pragma unroll
  for (int u = 0; u < KO; ++u) {
    kk[u] = min(t + W * u, NC - 1);
    y[u] = 0.0f;
  }
#pragma unroll 1
  for (int j0 = 0; j0 < NC; j0 += B) {
    float tv[B][KO];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = min(j0 + b, NC - 1);
#pragma unroll
      for (int u = 0; u < KO; ++u) tv[b][u] = T[(size_t)j * NC + kk[u]];
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const float xv = j0 + b < NC ? x[min(j0 + b, NC - 1)] : 0.0f;
#pragma unroll
      for (int u = 0; u < KO; ++u) y[u] = fmaf(tv[b][u], xv, y[u]);
    }
  }
}

// ---------------------------------------------------------------------------
// P2M: one 2-wave block per leaf at level >= 2 (lanes stream the points, the
// two waves' sums meet in LDS, then a column reduction in a fixed order).
// The charges come straight from the caller's sigma (q = (float)sigma[perm]
// * area, the same float k_gather writes for P2P), so P2M need not wait for
// k_gather, which runs on the side stream with P2P.  Only x, y, z of pts are
// used (k_gather may be writing .w meanwhile).
// ---------------------------------------------------------------------------
constexpr int kP2MThreads = 128;
template <int P, bool MULTI, typename TS>
__global__ __launch_bounds__(kP2MThreads) void k_p2m(const int32_t* __restrict__ leaves,
                                                     const int2* __restrict__ rng,
                                                     const float4* __restrict__ ctr,
                                                     const float4* __restrict__ pts,
                                                     const int32_t* __restrict__ perm,
                                                     const TS* __restrict__ sigma, int sigma_tree,
                                                     const float* __restrict__ area_t, int node2,
                                                     float* __restrict__ M,
                                                     const float* __restrict__ ascale,
                                                     unsigned* __restrict__ mmax_bits,
                                                     RhsStride rs) {
  chain_prio();
  constexpr int NC = (P + 1) * (P + 1);
  constexpr int W = kP2MThreads;
  __shared__ float red[64][NC + 1];
  const int ry = MULTI ? blockIdx.y : 0;
  if constexpr (MULTI) {
    M += ry * rs.node;
    if (mmax_bits) mmax_bits += ry * rs.slot;
  }
  const int leaf = leaves[blockIdx.x];
  if (leaf < node2) return;  // no multipole needed above level 2
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {
    const int2 r = rng[leaf];
    const float4 c = ctr[leaf];
    float acc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = 0.0f;
    const TS* sg = sigma + (MULTI ? (size_t)ry * rs.n : 0);
    for (int i = r.x + tid; i < r.y; i += W) {
      const float4 pq = pts[i];
      const float q = (float)sg[sigma_tree ? i : perm[i]] * area_t[i];
      p2m_accum<P>((pq.x - c.x) * c.w, (pq.y - c.y) * c.w, (pq.z - c.z) * c.w, q, acc);
    }
    // the two waves' sums per lane: wave 1 writes, wave 0 adds its own
    if (wave == 1) {
#pragma unroll
      for (int k = 0; k < NC; ++k) red[lane][k] = acc[k];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int k = 0; k < NC; ++k) red[lane][k] = acc[k] + red[lane][k];
    }
  }
  __syncthreads();
  float mx = 0.0f;
  for (int k = tid; k < NC; k += W) {
    // 64 lane sums in a fixed order, read 16 at a time (independent loads)
    float s = 0.0f;
#pragma unroll
    for (int t0 = 0; t0 < 64; t0 += 16) {
      float v[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) v[t] = red[t0 + t][k];
#pragma unroll
      for (int t = 0; t < 16; ++t) s += v[t];
    }
    M[(size_t)leaf * NC + k] = s;
    if (mmax_bits) mx = fmaxf(mx, fabsf(s * ascale[k]));
  }
  if (mmax_bits) fold_absmax(mx, mmax_bits, 2 * blockIdx.x + wave);
}

// ---------------------------------------------------------------------------
// M2M, one launch per level: one 4-wave block per parent.  The children's
// multipoles are staged in LDS; the two 128-thread halves apply the octant
// operators of children h, h + 2, ... (M_p += T_oct(c) M_c, one output per
// thread), and the two half sums are added in a fixed order.
// ---------------------------------------------------------------------------
template <int P, bool MULTI>
__global__ __launch_bounds__(256) void k_m2m(const int32_t* __restrict__ parents,
                                             const int32_t* __restrict__ first_child,
                                             const int32_t* __restrict__ n_child,
                                             const int32_t* __restrict__ octant,
                                             const float* __restrict__ T, float* __restrict__ M,
                                             const float* __restrict__ ascale,
                                             unsigned* __restrict__ mmax_bits, RhsStride rs) {
  chain_prio();
  constexpr int NC = (P + 1) * (P + 1);
  constexpr int W = 128, KO = (NC + W - 1) / W;
  __shared__ float mc[8][NC];
  __shared__ float part[2][NC];
  __shared__ int octs[8];
  if constexpr (MULTI) {
    M += blockIdx.y * rs.node;
    if (mmax_bits) mmax_bits += blockIdx.y * rs.slot;
  }
  const int par = parents[blockIdx.x];
  const int c0 = first_child[par], nch = n_child[par];
  for (int e = threadIdx.x; e < nch * NC; e += blockDim.x) {
    const int c = e / NC;
    mc[c][e - c * NC] = M[(size_t)(c0 + c) * NC + (e - c * NC)];
  }
  if (threadIdx.x < nch) octs[threadIdx.x] = octant[c0 + threadIdx.x];
  __syncthreads();
  const int h = threadIdx.x >> 7, t = threadIdx.x & (W - 1);
  float acc[KO];
#pragma unroll
  for (int u = 0; u < KO; ++u) acc[u] = 0.0f;
  for (int c = h; c < nch; c += 2) {
    float y[KO];
    op_apply<NC, W>(T + (size_t)octs[c] * NC * NC, mc[c], t, y);
#pragma unroll
    for (int u = 0; u < KO; ++u) acc[u] += y[u];
  }
#pragma unroll
  for (int u = 0; u < KO; ++u)
    if (t + W * u < NC) part[h][t + W * u] = acc[u];
  __syncthreads();
  float mx = 0.0f;
  for (int k = threadIdx.x; k < NC; k += blockDim.x) {
    const float v = part[0][k] + part[1][k];
    M[(size_t)par * NC + k] = v;
    if (mmax_bits) mx = fmaxf(mx, fabsf(v * ascale[k]));
  }
  if (mmax_bits) fold_absmax(mx, mmax_bits, blockIdx.x * 4 + (threadIdx.x >> 6));  // whole waves
}

// ---------------------------------------------------------------------------
// M2M with the operator in LDS (orders whose (p+1)^2 x (p+1)^2 table fits,
// p <= 12).  The eight octant operators are reflections of the one for
// octant 7 (child on the + side of every axis): T(o) = D T(7) D with s = ~o &
// 7, so a block stages that one table once and flips signs instead of
// reading a child's table from L2 per child (8 x 58 KB per parent at p = 10,
// which kept the L2-bound per-level M2M at ~10 us).
// ---------------------------------------------------------------------------
template <int P>
struct OctLds {
  static constexpr int NC = (P + 1) * (P + 1);
  static constexpr int E4 = (NC * NC + 3) / 4;  // float4s of the padded table
  static constexpr bool ON = (size_t)E4 * 16 + (size_t)NC * 40 <= 150 * 1024;
};

// Ts <- canonical octant table (column-major, T[j NC + k], padded to E4
// float4s, 16-byte aligned in global memory and LDS) by nt = 256 threads,
// 16 float4 loads in flight per thread.
template <int E4>
__device__ __forceinline__ void stage_oct_table(const float4* __restrict__ T7, float4* Ts, int t) {
  constexpr int NT = 256, U = 16;
  for (int e0 = 0; e0 < E4; e0 += U * NT) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + t + u * NT;
      v[u] = T7[min(e, E4 - 1)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + t + u * NT;
      if (e < E4) Ts[e] = v[u];
    }
  }
}

// M2M, one launch per level, one 4-wave block per parent: the children's
// multipoles (reflected to octant 7) sit in LDS as xs[j][8]; thread (h, k)
// sums T(7)[j][k] xs[j][c] over its half of j for all eight children at
// once (one table read, two broadcast float4 reads, eight FMAs per j), the
// per-child results are reflected back and added in child order, then the
// two halves in a fixed order.
template <int P, bool MULTI>
__global__ __launch_bounds__(256) void k_m2m_lds(const int32_t* __restrict__ parents,
                                                 const int32_t* __restrict__ first_child,
                                                 const int32_t* __restrict__ n_child,
                                                 const int32_t* __restrict__ octant,
                                                 const float4* __restrict__ T7,
                                                 const uint8_t* __restrict__ flip,
                                                 float* __restrict__ M,
                                                 const float* __restrict__ ascale,
                                                 unsigned* __restrict__ mmax_bits, RhsStride rs) {
  chain_prio();
  constexpr int NC = (P + 1) * (P + 1), W = 128, KO = (NC + W - 1) / W, JH = (NC + 1) / 2;
  __shared__ __align__(16) float Ts[OctLds<P>::E4 * 4];
  __shared__ __align__(16) float xs[NC][8];
  __shared__ float part[2][NC];
  __shared__ int sgn[8];
  __shared__ int fl[NC];
  if constexpr (MULTI) {
    M += blockIdx.y * rs.node;
    if (mmax_bits) mmax_bits += blockIdx.y * rs.slot;
  }
  const int par = parents[blockIdx.x];
  const int c0 = first_child[par], nch = n_child[par];
  if (threadIdx.x < 8) sgn[threadIdx.x] = threadIdx.x < nch ? (~octant[c0 + threadIdx.x]) & 7 : 0;
  for (int k = threadIdx.x; k < NC; k += blockDim.x) fl[k] = flip[k];
  stage_oct_table<OctLds<P>::E4>(T7, reinterpret_cast<float4*>(Ts), threadIdx.x);
  __syncthreads();
  for (int e = threadIdx.x; e < 8 * NC; e += blockDim.x) {
    const int c = e / NC, j = e - c * NC;
    xs[j][c] = c < nch ? refl(M[(size_t)(c0 + c) * NC + j], sgn[c], fl[j]) : 0.0f;
  }
  __syncthreads();
  const int h = threadIdx.x >> 7, t = threadIdx.x & (W - 1);
  const int jb = h * JH, je = h ? NC : JH;
#pragma unroll
  for (int u = 0; u < KO; ++u) {
    const int k = min(t + W * u, NC - 1);
    float acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = 0.0f;
#pragma unroll 4
    for (int j = jb; j < je; ++j) {
      const float tv = Ts[j * NC + k];
      const float4 a = *reinterpret_cast<const float4*>(&xs[j][0]);
      const float4 b = *reinterpret_cast<const float4*>(&xs[j][4]);
      acc[0] = fmaf(tv, a.x, acc[0]); acc[1] = fmaf(tv, a.y, acc[1]);
      acc[2] = fmaf(tv, a.z, acc[2]); acc[3] = fmaf(tv, a.w, acc[3]);
      acc[4] = fmaf(tv, b.x, acc[4]); acc[5] = fmaf(tv, b.y, acc[5]);
      acc[6] = fmaf(tv, b.z, acc[6]); acc[7] = fmaf(tv, b.w, acc[7]);
    }
    float y = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) y += refl(acc[c], sgn[c], fl[k]);
    if (t + W * u < NC) part[h][k] = y;
  }
  __syncthreads();
  float mx = 0.0f;
  for (int k = threadIdx.x; k < NC; k += blockDim.x) {
    const float v = part[0][k] + part[1][k];
    M[(size_t)par * NC + k] = v;
    if (mmax_bits) mx = fmaxf(mx, fabsf(v * ascale[k]));
  }
  if (mmax_bits) fold_absmax(mx, mmax_bits, blockIdx.x * 4 + (threadIdx.x >> 6));  // whole waves
}

// ---------------------------------------------------------------------------
// M2L (VALU, CSR by target): L_own[t] = inv_s * sum_{s in V(t)} T_off M[s]
// (the node's own part; k_m2l_down adds the ancestors').  Kept as the plain
// reference path of the GEMM form below.
// ---------------------------------------------------------------------------
template <int P, bool MULTI>
__global__ __launch_bounds__(128) void k_m2l_valu(int node0, const int32_t* __restrict__ v_ptr,
                                                  const int32_t* __restrict__ v_src,
                                                  const int32_t* __restrict__ v_tab,
                                                  const float4* __restrict__ ctr,
                                                  const float* __restrict__ T,
                                                  const float* __restrict__ M,
                                                  float* __restrict__ L, RhsStride rs) {
  chain_prio();
  constexpr int NC = (P + 1) * (P + 1);
  __shared__ float ms[NC];
  if constexpr (MULTI) {
    M += blockIdx.y * rs.node;
    L += blockIdx.y * rs.node;
  }
  const int t = node0 + blockIdx.x;
  const int b = v_ptr[t], e = v_ptr[t + 1];
  constexpr int KO = (NC + 127) / 128;  // outputs per thread
  float acc[KO];
#pragma unroll
  for (int u = 0; u < KO; ++u) acc[u] = 0.0f;
  for (int q = b; q < e; ++q) {
    const int s = v_src[q];
    __syncthreads();
    for (int j = threadIdx.x; j < NC; j += blockDim.x) ms[j] = M[(size_t)s * NC + j];
    __syncthreads();
    const float* Tq = T + (size_t)v_tab[q] * NC * NC;
#pragma unroll
    for (int u = 0; u < KO; ++u) {
      const int k = threadIdx.x + 128 * u;
      if (k < NC)
        for (int j = 0; j < NC; ++j) acc[u] += Tq[(size_t)j * NC + k] * ms[j];
    }
  }
  const float is = ctr[t].w;
#pragma unroll
  for (int u = 0; u < KO; ++u) {
    const int k = threadIdx.x + 128 * u;
    if (k < NC) L[(size_t)t * NC + k] = acc[u] * is;
  }
}

// ---------------------------------------------------------------------------
// M2L (MFMA): one block of 4 waves per tile of <= 128 pairs sharing one offset
// table.  Y[pair][0..NCP) = T_tab (NCP x KP, row-major, zero padded) * M[src].
// v_mfma_f32_32x32x2f32 lane layout: A[i][k] from lane (i + 32 k),
// B[k][j] from lane (j + 32 k), D[i][j] in lane j + 32*((i/4)%2), register
// 4*(i/8) + i%4.
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int P>
struct M2LShape {
  static constexpr int NC = (P + 1) * (P + 1);
  static constexpr int NCP = (NC + 31) / 32 * 32;  // output rows, padded
  static constexpr int KP = (NC + 1) / 2 * 2;      // inner dim, padded to K=2
  static constexpr int KS = KP + 1;                // LDS row stride
  static constexpr int COLS = 64;                   // pairs per tile (block)
  static constexpr int THREADS = 256;               // 4 waves: 2 column tiles x 2 row halves
  // row parts per column tile: at most 2 row blocks (64 accumulators) per wave
  static constexpr int NRBT = NCP / 32;
  static constexpr int RSPLIT = NRBT <= 2 ? 1 : (NRBT % 2 == 0 ? NRBT / 2 : NRBT);
  static constexpr bool ENABLED = P <= 12;          // table fits the 160 KiB LDS
};

// 16-bit MFMA (v_mfma_f32_32x32x16_{bf16,f16}, gfx950): K steps of 16.
template <int P>
struct M2LShapeH {
  static constexpr int NC = (P + 1) * (P + 1);
  static constexpr int NCP = (NC + 31) / 32 * 32;
  static constexpr int KP = (NC + 15) / 16 * 16;
  static constexpr int KS = KP + 8;  // row stride in elements (16-byte aligned)
  static constexpr int COLS = 128;
  static constexpr int THREADS = 256;
};

// 16-bit inputs are scaled by one power of two per matvec so that the largest
// scaled multipole is 2^14: FP16 keeps its ~11 decades of range where the
// values are (the corrections of iterative refinement are orders of magnitude
// below the first right-hand side) and the scale is undone exactly on Y.
constexpr int kM2LScaleExp = 14;

// Reflection-folded coefficients of a 16-bit M2L tile, in LDS: row s (one per
// axis reflection s = 0..7) of sa holds refl(ascale[k] * sc, s, flip[k]) (0
// past NC), so one multiply scales and reflects a gathered multipole
// coefficient; row s of so holds refl(1, s, flip[i]), the sign of result row
// i.  Both products are exact sign flips of the unreflected ones.  Rows are
// padded by 4 floats so that lanes of different reflections start on
// different LDS banks.  load() fetches each thread's share before the scale
// is known; store() writes the rows once it is.
template <int P>
struct M2LRefl {
  using S = M2LShapeH<P>;
  static constexpr int SA = S::KP + 4, SO = S::NCP + 4;
  static constexpr int NA = (8 * S::KP + 255) / 256, NO = (8 * S::NCP + 255) / 256;
  float av[NA];
  int fa[NA], fo[NO];
  __device__ __forceinline__ void load(const float* __restrict__ ascale,
                                       const uint8_t* __restrict__ flip) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int idx = (int)threadIdx.x + 256 * u, k = idx % S::KP;
      av[u] = (idx < 8 * S::KP && k < S::NC) ? ascale[k] : 0.0f;
      fa[u] = flip[min(k, S::NC - 1)];
    }
#pragma unroll
    for (int u = 0; u < NO; ++u) fo[u] = flip[((int)threadIdx.x + 256 * u) % S::NCP];
  }
  __device__ __forceinline__ void store(float sc, float* __restrict__ sa,
                                        float* __restrict__ so) const {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int idx = (int)threadIdx.x + 256 * u, s = idx / S::KP;
      if (idx < 8 * S::KP) sa[s * SA + idx - s * S::KP] = refl(av[u] * sc, s, fa[u]);
    }
#pragma unroll
    for (int u = 0; u < NO; ++u) {
      const int idx = (int)threadIdx.x + 256 * u, s = idx / S::NCP;
      if (idx < 8 * S::NCP) so[s * SO + idx - s * S::NCP] = refl(1.0f, s, fo[u]);
    }
  }
};

// M2L tables exist for offsets with non-negative components only: a pair
// whose offset is a reflection of a table's offset applies D T(o) D (refl).
template <int P, bool MULTI>
__global__ __launch_bounds__(M2LShape<P>::THREADS) void k_m2l_mfma(const int32_t* __restrict__ tile_tab,
                                                  const int32_t* __restrict__ tile_beg,
                                                  const int32_t* __restrict__ tile_cnt,
                                                  const int32_t* __restrict__ pair_src,
                                                  const uint8_t* __restrict__ pair_sgn,
                                                  const uint8_t* __restrict__ flip,
                                                  const float* __restrict__ Tm,
                                                  const float* __restrict__ ascale,
                                                  const float* __restrict__ M,
                                                  float* __restrict__ Y, RhsStride rs) {
  chain_prio();
  using S = M2LShape<P>;
  constexpr int NC = S::NC, NCP = S::NCP, KP = S::KP, KS = S::KS;
  constexpr int NT = S::THREADS, NW = NT / 64, KH = KP / 2;
  __shared__ float ta[NCP * KS];
  __shared__ uint8_t fl[NCP];
  if constexpr (MULTI) {
    M += blockIdx.y * rs.node;
    Y += blockIdx.y * rs.pair;
  }
  const int tile = blockIdx.x;
  const int tab = tile_tab[tile], beg = tile_beg[tile], cnt = tile_cnt[tile];
  if (cnt == 0) return;  // padding tile of the XCD-aware order
  // table -> LDS with 16-byte loads (NCP * KP is a multiple of 4 and every
  // table starts 16-byte aligned), unrolled so the loads are in flight together
  const float4* T4 = reinterpret_cast<const float4*>(Tm + (size_t)tab * NCP * KP);
  constexpr int N4 = NCP * KP / 4;
  static_assert((NCP * KP) % 4 == 0, "table size must be a multiple of 4 floats");
#pragma unroll 8
  for (int t = threadIdx.x; t < N4; t += NT) {
    const float4 v = T4[t];
    const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 4 * t + q, i = e / KP, k = e - i * KP;
      ta[i * KS + k] = vv[q];
    }
  }
  for (int k = threadIdx.x; k < NCP; k += NT) fl[k] = flip[k];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 31, lk = lane >> 5;
  constexpr int RS = S::RSPLIT;
  constexpr int NRB = NCP / 32 / RS;  // row blocks per wave
  const int ntiles = (cnt + 31) / 32;
  // work item = (column tile, row part); waves take items round robin
  for (int item = wave; item < ntiles * RS; item += NW) {
    const int ct = item / RS, rb0 = (item % RS) * NRB;
    const int col = ct * 32 + li;
    // B[k][col] for k = 2s + lk, straight from the (L2-resident) multipoles.
    // Every load is unconditional (a clamped, in-bounds address; masked after)
    // so the compiler issues all of them before the first wait.
    const bool cv = col < cnt;
    const int srow = cv ? pair_src[beg + min(col, cnt - 1)] : pair_src[beg];
    const int sg = pair_sgn[beg + min(col, cnt - 1)];
    const float* mrow = M + (size_t)srow * NC;
    float bv[KH];
#pragma unroll
    for (int s = 0; s < KH; ++s) {
      const int k = 2 * s + lk;
      const int kk = min(k, NC - 1);
      const float v = refl(mrow[kk] * ascale[kk], sg, fl[kk]);
      bv[s] = (cv && k < NC) ? v : 0.0f;
    }
    f32x16 acc[NRB];
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[rb][r] = 0.0f;
#pragma unroll
    for (int s = 0; s < KH; ++s) {
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) {
        const float a = ta[((rb0 + rb) * 32 + li) * KS + 2 * s + lk];
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bv[s], acc[rb], 0, 0, 0);
      }
    }
    if (col < cnt) {
      float* yrow = Y + (size_t)(beg + col) * NCP;
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = (rb0 + rb) * 32 + q * 8 + lk * 4;
          float4 v = make_float4(refl(acc[rb][4 * q + 0], sg, fl[row]),
                                 refl(acc[rb][4 * q + 1], sg, fl[row + 1]),
                                 refl(acc[rb][4 * q + 2], sg, fl[row + 2]),
                                 refl(acc[rb][4 * q + 3], sg, fl[row + 3]));
          *reinterpret_cast<float4*>(yrow + row) = v;
        }
    }
  }
}

// 16-bit M2L GEMM: table and gathered (scaled) multipoles rounded to bf16 or
// fp16, FP32 accumulation.  Lane layout of v_mfma_f32_32x32x16: A[i][k] for
// k = 8*(lane/32) .. +7 from lane i + 32*(k/8), B likewise with j = lane%32;
// D as in the fp32 kernel.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

template <typename T16> struct Mfma16;
template <> struct Mfma16<__bf16> {
  typedef bf16x8 vec;
  static __device__ __forceinline__ f32x16 run(vec a, vec b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma16<_Float16> {
  typedef f16x8 vec;
  static __device__ __forceinline__ f32x16 run(vec a, vec b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};

template <int P, typename T16, bool MULTI>
__global__ __launch_bounds__(256) void k_m2l_mfma16(const int32_t* __restrict__ tile_tab,
                                                    const int32_t* __restrict__ tile_beg,
                                                    const int32_t* __restrict__ tile_cnt,
                                                    const int32_t* __restrict__ pair_src,
                                                    const uint8_t* __restrict__ pair_sgn,
                                                    const uint8_t* __restrict__ flip,
                                                    const T16* __restrict__ Tm,
                                                    const float* __restrict__ ascale,
                                                    const float* __restrict__ M,
                                                    const unsigned* __restrict__ mmax_bits,
                                                    float* __restrict__ Y, RhsStride rs) {
  chain_prio();
  using S = M2LShapeH<P>;
  using V = typename Mfma16<T16>::vec;
  if constexpr (MULTI) {
    M += blockIdx.y * rs.node;
    Y += blockIdx.y * rs.pair;
    mmax_bits += blockIdx.y * rs.slot;
  }
  constexpr int NC = S::NC, NCP = S::NCP, KP = S::KP, KS = S::KS, NSV = KP / 16;
  __shared__ __attribute__((aligned(16))) T16 ta[NCP * KS];
  using R = M2LRefl<P>;
  __shared__ __attribute__((aligned(16))) float sa[8 * R::SA];  // reflection-folded ascale x scale
  __shared__ __attribute__((aligned(16))) float so[8 * R::SO];  // result-row signs
  if (tile_cnt[blockIdx.x] == 0) return;  // padding tile of the XCD-aware order
  const int tile = blockIdx.x;
  const int tab = tile_tab[tile], beg = tile_beg[tile], cnt = tile_cnt[tile];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col0 = wave * 32;
  const int li = lane & 31, lk = lane >> 5;
  const int col = col0 + li;
  const bool cv = col < cnt;
  // Every global load before the first use (as in k_m2l_mfma16x3): the
  // pair's source and sign, the scale slots, ascale / flip, the table (-> LDS,
  // 16 bytes per load, all in flight at p <= 10) and this lane's multipole
  // column, which goes into the B fragment straight from registers.
  const int sg = pair_sgn[beg + min(col, cnt - 1)];
  const int src = pair_src[beg + min(col, cnt - 1)];
  const unsigned slot_bits = mmax_bits[(threadIdx.x & (kMaxSlots - 1)) * kMaxSlotStride];
  R rf;
  rf.load(ascale, flip);
  const uint4* T8 = reinterpret_cast<const uint4*>(Tm + (size_t)tab * NCP * KP);
  constexpr int N8 = NCP * KP / 8, TU = 16;
  constexpr bool kWhole = N8 % (TU * 256) == 0;
  float mv[NSV][8];  // k = 16 sv + 8 lk + j
  for (int t0 = 0; t0 < N8; t0 += TU * 256) {
    uint4 tv[TU];
#pragma unroll
    for (int u = 0; u < TU; ++u) tv[u] = T8[min(t0 + (int)threadIdx.x + 256 * u, N8 - 1)];
    if (t0 == 0 && col0 < cnt) {
      const float* mrow = M + (size_t)src * NC;
#pragma unroll
      for (int sv = 0; sv < NSV; ++sv)
#pragma unroll
        for (int j = 0; j < 8; ++j) mv[sv][j] = mrow[min(16 * sv + 8 * lk + j, NC - 1)];
    }
#pragma unroll
    for (int u = 0; u < TU; ++u) {
      const int t = t0 + (int)threadIdx.x + 256 * u;
      if (kWhole || t < N8) {
        const int e = 8 * t, i = e / KP, k = e - i * KP;
        *reinterpret_cast<uint4*>(ta + i * KS + k) = tv[u];
      }
    }
  }
  // power-of-two input scale from the largest |M * ascale| of this matvec
  float gmax = __uint_as_float(slot_bits);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) gmax = fmaxf(gmax, __shfl_xor(gmax, o, 64));
  const float sc = gmax > 0.0f && isfinite(gmax) ? exp2f((float)kM2LScaleExp - ceilf(log2f(gmax))) : 1.0f;
  const float inv_sc = 1.0f / sc;
  rf.store(sc, sa, so);
  __syncthreads();
  if (col0 >= cnt) return;  // no barrier below
  constexpr int NRB = NCP / 32;
  f32x16 acc[NRB];
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[rb][r] = 0.0f;
#pragma unroll
  for (int sv = 0; sv < NSV; ++sv) {
    // a column past cnt computes from the clamped pair's multipoles and is
    // not stored (the columns of the product are independent)
    const float4 s0 = *reinterpret_cast<const float4*>(sa + sg * R::SA + 16 * sv + 8 * lk);
    const float4 s1 = *reinterpret_cast<const float4*>(sa + sg * R::SA + 16 * sv + 8 * lk + 4);
    const float sk[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    V bv;
#pragma unroll
    for (int j = 0; j < 8; ++j) bv[j] = (T16)(mv[sv][j] * sk[j]);
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) {
      const V a = *reinterpret_cast<const V*>(ta + (rb * 32 + li) * KS + 16 * sv + 8 * lk);
      acc[rb] = Mfma16<T16>::run(a, bv, acc[rb]);
    }
  }
  if (cv) {
    float* yrow = Y + (size_t)(beg + col) * NCP;
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = rb * 32 + q * 8 + lk * 4;
        const float4 sr = *reinterpret_cast<const float4*>(so + sg * R::SO + row);
        float4 v = make_float4(acc[rb][4 * q + 0] * inv_sc * sr.x, acc[rb][4 * q + 1] * inv_sc * sr.y,
                               acc[rb][4 * q + 2] * inv_sc * sr.z, acc[rb][4 * q + 3] * inv_sc * sr.w);
        *reinterpret_cast<float4*>(yrow + row) = v;
      }
  }
}

// Split-FP16 M2L with FP32-level accuracy: table and (scaled) multipoles are
// each a sum hi + lo of two FP16 numbers (22 significant bits), and
// A B ~ A_hi B_hi + A_hi B_lo + A_lo B_hi is three v_mfma_f32_32x32x16_f16
// with FP32 accumulation (the dropped A_lo B_lo is ~2^-22 of the product).
// Both table halves sit in LDS; each lane forms its column's B hi / lo in
// registers from the gathered FP32 multipoles.  Four waves x 32 columns x
// all rows, the same tile schedule and Y layout as the 16-bit kernel.  Each
// lane issues all loads of its multipole column (KP / 2 values) before the
// table copy, so their latency hides behind it.  (k_m2l_mfma16x3 for one
// right-hand side, k_m2l_mfma16x3_multi for several.)

// The KP / 2 multipole values of one lane's column: k = 16 sv + 8 lk + j
// (addresses clamped to the row).
template <int P>
__device__ __forceinline__ void m2l16x3_load(const float* __restrict__ mrow, int lk,
                                             float (&mv)[M2LShapeH<P>::KP / 16][8]) {
  using S = M2LShapeH<P>;
#pragma unroll
  for (int sv = 0; sv < S::KP / 16; ++sv)
#pragma unroll
    for (int j = 0; j < 8; ++j) mv[sv][j] = mrow[min(16 * sv + 8 * lk + j, S::NC - 1)];
}

// Power-of-two input scale of the 16-bit M2L from the largest |M * ascale|
// of one RHS.
__device__ __forceinline__ float m2l_scale(const unsigned* __restrict__ slots) {
  const float gmax = slots_absmax(slots);
  return gmax > 0.0f && isfinite(gmax) ? exp2f((float)kM2LScaleExp - ceilf(log2f(gmax))) : 1.0f;
}

// One wave's 32 columns x all rows of the split-FP16 product for one RHS,
// stored to its Y rows (lanes with cv false compute from the clamped pair's
// multipoles and store nothing: the columns are independent).  sa_l / so_l
// are the lane's reflection rows of M2LRefl; sa_l holds ascale x sc0, and for
// another RHS with scale sc the inputs take the power of two rel = sc / sc0
// on top (RESCALE), which rounds exactly like ascale x sc.
template <int P, bool RESCALE>
__device__ __forceinline__ void m2l16x3_column(const _Float16* __restrict__ ta,
                                               const float* __restrict__ sa_l,
                                               const float* __restrict__ so_l,
                                               const float (&mv)[M2LShapeH<P>::KP / 16][8],
                                               float rel, float inv_sc, bool cv, int li, int lk,
                                               float* __restrict__ yrow) {
  using S = M2LShapeH<P>;
  constexpr int NCP = S::NCP, KS = S::KS, NSV = S::KP / 16, NRB = NCP / 32;
  f32x16 acc[NRB];
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[rb][q] = 0.0f;
#pragma unroll
  for (int sv = 0; sv < NSV; ++sv) {
    const float4 s0 = *reinterpret_cast<const float4*>(sa_l + 16 * sv + 8 * lk);
    const float4 s1 = *reinterpret_cast<const float4*>(sa_l + 16 * sv + 8 * lk + 4);
    const float sk[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    f16x8 bh, bl;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float v = mv[sv][j] * sk[j];
      if (RESCALE) v *= rel;
      const _Float16 h = (_Float16)v;
      bh[j] = h;
      bl[j] = (_Float16)(v - (float)h);
    }
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) {
      const int at = (rb * 32 + li) * KS + 16 * sv + 8 * lk;
      const f16x8 ah = *reinterpret_cast<const f16x8*>(ta + at);
      const f16x8 al = *reinterpret_cast<const f16x8*>(ta + NCP * KS + at);
      acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[rb], 0, 0, 0);
      acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[rb], 0, 0, 0);
      acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[rb], 0, 0, 0);
    }
  }
  if (!cv) return;
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = rb * 32 + q * 8 + lk * 4;
      const float4 sr = *reinterpret_cast<const float4*>(so_l + row);
      float4 v = make_float4(acc[rb][4 * q + 0] * inv_sc * sr.x, acc[rb][4 * q + 1] * inv_sc * sr.y,
                             acc[rb][4 * q + 2] * inv_sc * sr.z, acc[rb][4 * q + 3] * inv_sc * sr.w);
      *reinterpret_cast<float4*>(yrow + row) = v;
    }
}

template <int P>
__global__ __launch_bounds__(256) void k_m2l_mfma16x3(const int32_t* __restrict__ tile_tab,
                                                      const int32_t* __restrict__ tile_beg,
                                                      const int32_t* __restrict__ tile_cnt,
                                                      const int32_t* __restrict__ pair_src,
                                                      const uint8_t* __restrict__ pair_sgn,
                                                      const uint8_t* __restrict__ flip,
                                                      const _Float16* __restrict__ Tm,
                                                      const float* __restrict__ ascale,
                                                      const float* __restrict__ M,
                                                      const unsigned* __restrict__ mmax_bits,
                                                      float* __restrict__ Y) {
  chain_prio();
  using S = M2LShapeH<P>;
  constexpr int NC = S::NC, NCP = S::NCP, KP = S::KP, KS = S::KS, NSV = KP / 16;
  __shared__ __attribute__((aligned(16))) _Float16 ta[2 * NCP * KS];  // hi, then lo
  using R = M2LRefl<P>;
  __shared__ __attribute__((aligned(16))) float sa[8 * R::SA];  // reflection-folded ascale x scale
  __shared__ __attribute__((aligned(16))) float so[8 * R::SO];  // result-row signs
  if (tile_cnt[blockIdx.x] == 0) return;  // padding tile of the XCD-aware order
  const int tile = blockIdx.x;
  const int tab = tile_tab[tile], beg = tile_beg[tile], cnt = tile_cnt[tile];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col0 = wave * 32;
  const int li = lane & 31, lk = lane >> 5;
  const int col = col0 + li;
  const bool cv = col < cnt;
  // Every global load of the block is issued before the first use, the
  // table first: its two halves -> LDS, 8 elements (16 bytes) per load, at
  // most 16 in flight per thread (all of them at p <= 10); rows are KP (a
  // multiple of 16) elements long, so a chunk never crosses a row.  (In
  // batches of 4 behind the scale lookup this took five dependent round
  // trips.)
  const int sg = pair_sgn[beg + min(col, cnt - 1)];
  const int src = pair_src[beg + min(col, cnt - 1)];
  const unsigned slot_bits = mmax_bits[(threadIdx.x & (kMaxSlots - 1)) * kMaxSlotStride];
  R rf;
  rf.load(ascale, flip);
  const uint4* T8 = reinterpret_cast<const uint4*>(Tm + (size_t)tab * 2 * NCP * KP);
  constexpr int N8 = NCP * KP / 8, TU = 16;
  constexpr bool kWhole = (2 * N8) % (TU * 256) == 0;  // no partial batch (p = 10)
  float mv[NSV][8];  // this lane's multipole values: k = 16 sv + 8 lk + j
  for (int t0 = 0; t0 < 2 * N8; t0 += TU * 256) {
    uint4 tv[TU];
#pragma unroll
    for (int u = 0; u < TU; ++u) tv[u] = T8[min(t0 + (int)threadIdx.x + 256 * u, 2 * N8 - 1)];
    if (t0 == 0 && col0 < cnt) {  // the multipole gathers behind the first batch
      const float* mrow = M + (size_t)src * NC;
#pragma unroll
      for (int sv = 0; sv < NSV; ++sv)
#pragma unroll
        for (int j = 0; j < 8; ++j) mv[sv][j] = mrow[min(16 * sv + 8 * lk + j, NC - 1)];
    }
#pragma unroll
    for (int u = 0; u < TU; ++u) {
      const int t = t0 + (int)threadIdx.x + 256 * u;
      if (kWhole || t < 2 * N8) {
        const int hh = t >= N8 ? 1 : 0;
        const int e = 8 * (t - hh * N8), i = e / KP, k = e - i * KP;
        *reinterpret_cast<uint4*>(ta + hh * NCP * KS + i * KS + k) = tv[u];
      }
    }
  }
  // power-of-two input scale from the largest |M * ascale| of this matvec
  // (the slots, loaded above, reduced across the wave)
  float gmax = __uint_as_float(slot_bits);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) gmax = fmaxf(gmax, __shfl_xor(gmax, o, 64));
  const float sc = gmax > 0.0f && isfinite(gmax) ? exp2f((float)kM2LScaleExp - ceilf(log2f(gmax))) : 1.0f;
  const float inv_sc = 1.0f / sc;
  rf.store(sc, sa, so);
  __syncthreads();
  if (col0 >= cnt) return;  // no barrier below
  m2l16x3_column<P, false>(ta, sa + sg * R::SA, so + sg * R::SO, mv, 1.0f, inv_sc, cv, li, lk,
                           Y + (size_t)(beg + col) * NCP);
}

// Several right-hand sides: one block per tile runs the tile for every RHS
// against the table it loaded once (the first RHS's scale folded into the
// LDS coefficients, the others' as an exact power of two on top).  The
// one-RHS kernel above keeps its own body: this helper-based form measured
// 1-2 us slower at 82k when it served one RHS too.
template <int P>
__global__ __launch_bounds__(256) void k_m2l_mfma16x3_multi(const int32_t* __restrict__ tile_tab,
                                                      const int32_t* __restrict__ tile_beg,
                                                      const int32_t* __restrict__ tile_cnt,
                                                      const int32_t* __restrict__ pair_src,
                                                      const uint8_t* __restrict__ pair_sgn,
                                                      const uint8_t* __restrict__ flip,
                                                      const _Float16* __restrict__ Tm,
                                                      const float* __restrict__ ascale,
                                                      const float* __restrict__ M,
                                                      const unsigned* __restrict__ mmax_bits,
                                                      float* __restrict__ Y, RhsStride rs) {
  chain_prio();
  using S = M2LShapeH<P>;
  constexpr int NC = S::NC, NCP = S::NCP, KP = S::KP, KS = S::KS, NSV = KP / 16;
  __shared__ __attribute__((aligned(16))) _Float16 ta[2 * NCP * KS];  // hi, then lo
  using R = M2LRefl<P>;
  __shared__ __attribute__((aligned(16))) float sa[8 * R::SA];  // reflection-folded ascale x sc0
  __shared__ __attribute__((aligned(16))) float so[8 * R::SO];  // result-row signs
  if (tile_cnt[blockIdx.x] == 0) return;  // padding tile of the XCD-aware order
  const int tile = blockIdx.x;
  const int tab = tile_tab[tile], beg = tile_beg[tile], cnt = tile_cnt[tile];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col0 = wave * 32;
  const int li = lane & 31, lk = lane >> 5;
  const int col = col0 + li;
  const bool cv = col < cnt;
  // every global load before the first use, as in k_m2l_mfma16x3
  const int sg = pair_sgn[beg + min(col, cnt - 1)];
  const int src = pair_src[beg + min(col, cnt - 1)];
  const unsigned slot_bits = mmax_bits[(threadIdx.x & (kMaxSlots - 1)) * kMaxSlotStride];
  R rf;
  rf.load(ascale, flip);
  // both table halves -> LDS, 8 elements (16 bytes) per load, 16 in flight
  // per thread; rows are KP (a multiple of 16) elements long
  const uint4* T8 = reinterpret_cast<const uint4*>(Tm + (size_t)tab * 2 * NCP * KP);
  constexpr int N8 = NCP * KP / 8, TU = 16;
  constexpr bool kWhole = (2 * N8) % (TU * 256) == 0;
  float mv[NSV][8];  // this lane's multipole values of RHS r: k = 16 sv + 8 lk + j
  for (int t0 = 0; t0 < 2 * N8; t0 += TU * 256) {
    uint4 tv[TU];
#pragma unroll
    for (int u = 0; u < TU; ++u) tv[u] = T8[min(t0 + (int)threadIdx.x + 256 * u, 2 * N8 - 1)];
    if (t0 == 0 && col0 < cnt) m2l16x3_load<P>(M + (size_t)src * NC, lk, mv);
#pragma unroll
    for (int u = 0; u < TU; ++u) {
      const int t = t0 + (int)threadIdx.x + 256 * u;
      if (kWhole || t < 2 * N8) {
        const int hh = t >= N8 ? 1 : 0;
        const int e = 8 * (t - hh * N8), i = e / KP, k = e - i * KP;
        *reinterpret_cast<uint4*>(ta + hh * NCP * KS + i * KS + k) = tv[u];
      }
    }
  }
  float gmax = __uint_as_float(slot_bits);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) gmax = fmaxf(gmax, __shfl_xor(gmax, o, 64));
  const float sc0 = gmax > 0.0f && isfinite(gmax) ? exp2f((float)kM2LScaleExp - ceilf(log2f(gmax))) : 1.0f;
  rf.store(sc0, sa, so);
  __syncthreads();
  if (col0 >= cnt) return;  // no barrier below
  m2l16x3_column<P, false>(ta, sa + sg * R::SA, so + sg * R::SO, mv, 1.0f, 1.0f / sc0, cv, li, lk,
                           Y + (size_t)(beg + col) * NCP);
#pragma unroll 1
  for (int r = 1; r < rs.R; ++r) {
    // keeps the table reads inside the loop: hoisted out of it they would
    // hold every A fragment of the tile in registers
    asm volatile("" ::: "memory");
    m2l16x3_load<P>(M + r * rs.node + (size_t)src * NC, lk, mv);
    const float sc = m2l_scale(mmax_bits + r * rs.slot);
    m2l16x3_column<P, true>(ta, sa + sg * R::SA, so + sg * R::SO, mv, sc / sc0, 1.0f / sc, cv, li, lk,
                            Y + r * rs.pair + (size_t)(beg + col) * NCP);
  }
}

// Own part of each node's local expansion from the GEMM rows:
// Lown[t] = inv_s(t) binv * (sum of t's rows Y in a fixed order).  The row
// indices of a batch are loaded first (uniform, scalar), then the rows.
template <int P, bool MULTI>
__global__ __launch_bounds__(128) void k_m2l_reduce(int node0, const int32_t* __restrict__ tptr,
                                                    const int32_t* __restrict__ trow,
                                                    const float4* __restrict__ ctr,
                                                    const float* __restrict__ binv,
                                                    const float* __restrict__ Y,
                                                    float* __restrict__ Lown,
                                                    unsigned* __restrict__ mmax_bits,
                                                    RhsStride rs) {
  if constexpr (MULTI) chain_prio();  // one RHS: default priority (see chain_prio)
  using S = M2LShape<P>;
  constexpr int NC = S::NC, NCP = S::NCP, W = 128, KO = (NC + W - 1) / W, RB = 16;
  if constexpr (MULTI) {
    Y += blockIdx.y * rs.pair;
    Lown += blockIdx.y * rs.node;
    if (mmax_bits) mmax_bits += blockIdx.y * rs.slot;
  }
  // the GEMM (previous launch) has read the scale slots: clear them for the
  // next matvec's P2M / M2M
  if (mmax_bits && blockIdx.x == 0 && threadIdx.x < kMaxSlots)
    mmax_bits[threadIdx.x * kMaxSlotStride] = 0u;
  const int t = node0 + blockIdx.x;
  const int b = tptr[t], e = tptr[t + 1];
  const float is = ctr[t].w;
  float sum[KO];
#pragma unroll
  for (int u = 0; u < KO; ++u) sum[u] = 0.0f;
  for (int r0 = b; r0 < e; r0 += RB) {
    int rows[RB];
#pragma unroll
    for (int v = 0; v < RB; ++v) rows[v] = trow[min(r0 + v, e - 1)];
    float yv[RB][KO];
#pragma unroll
    for (int v = 0; v < RB; ++v)
#pragma unroll
      for (int u = 0; u < KO; ++u)
        yv[v][u] = Y[(size_t)rows[v] * NCP + min((int)threadIdx.x + W * u, NC - 1)];
#pragma unroll
    for (int v = 0; v < RB; ++v)
#pragma unroll
      for (int u = 0; u < KO; ++u)
        if (r0 + v < e) sum[u] += yv[v][u];
  }
#pragma unroll
  for (int u = 0; u < KO; ++u) {
    const int k = threadIdx.x + W * u;
    if (k < NC) Lown[(size_t)t * NC + k] = sum[u] * (is * binv[k]);
  }
}

// Complete local expansions in one launch: one block per node t at level >=
// 2 runs the L2L recursion down its own chain of ancestors (level 2 to t),
// x = Lown(a_top), then x = T_oct(a) x + Lown(a), so that no block waits for
// another (a node's chain is at most the tree depth long).
constexpr int kMaxChain = 24;
template <int P, bool MULTI>
__global__ __launch_bounds__(128) void k_m2l_down(int node0, const int32_t* __restrict__ parent,
                                                  const int32_t* __restrict__ octant,
                                                  const float* __restrict__ Lown,
                                                  const float* __restrict__ T,
                                                  float* __restrict__ L, RhsStride rs) {
  if constexpr (MULTI) chain_prio();  // one RHS: default priority (see chain_prio)
  constexpr int NC = (P + 1) * (P + 1), W = 128, KO = (NC + W - 1) / W;
  __shared__ float xs[2][NC];
  __shared__ int chain[kMaxChain], octs[kMaxChain];
  __shared__ int nchain;
  if constexpr (MULTI) {
    Lown += blockIdx.y * rs.node;
    L += blockIdx.y * rs.node;
  }
  const int t = node0 + blockIdx.x;
  if (threadIdx.x == 0) {
    int d = 0;
    for (int a = t; a >= node0 && d < kMaxChain; a = parent[a]) {
      chain[d] = a;
      octs[d] = octant[a];
      ++d;
    }
    nchain = d;
  }
  __syncthreads();
  const int nq = nchain;
  if (nq == 1) {  // level 2: nothing to inherit
    for (int k = threadIdx.x; k < NC; k += W) L[(size_t)t * NC + k] = Lown[(size_t)t * NC + k];
    return;
  }
  int buf = 0;
  for (int k = threadIdx.x; k < NC; k += W) xs[0][k] = Lown[(size_t)chain[nq - 1] * NC + k];
  __syncthreads();
  for (int q = nq - 2; q >= 0; --q) {
    const int a = chain[q];
    float own[KO];
#pragma unroll
    for (int u = 0; u < KO; ++u) own[u] = Lown[(size_t)a * NC + min((int)threadIdx.x + W * u, NC - 1)];
    float tx[KO];
    op_apply<NC, W>(T + (size_t)octs[q] * NC * NC, xs[buf], threadIdx.x, tx);
    if (q == 0) {
#pragma unroll
      for (int u = 0; u < KO; ++u) {
        const int k = threadIdx.x + W * u;
        if (k < NC) L[(size_t)t * NC + k] = own[u] + tx[u];
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < KO; ++u) {
      const int k = threadIdx.x + W * u;
      if (k < NC) xs[buf ^ 1][k] = own[u] + tx[u];
    }
    __syncthreads();
    buf ^= 1;
  }
}

// ---------------------------------------------------------------------------
// P2P: one wave per work item (ed::P2PItem), four independent waves a
// block.  A chunk of n_t <= 64 targets of one leaf meets `cnt` consecutive
// points of the leaf's source stream: the
// leaf itself first (with the j != i exclusion), then its other U-list
// leaves.  Chunks of more than 32 targets (one slice) read each source at a
// wave-uniform address, so the loads are scalar and the sources sit in SGPRs:
// no LDS staging and no barriers.  Smaller chunks split the lanes into
// S = min(64 / n_t, 8) slices, slice s taking every S-th point, and add the
// slices in a fixed order at the end.  Every item writes its own row of
// partial sums and k_l2p_out adds a panel's rows in item order, so the
// result does not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int kP2PUnroll = 8;
constexpr int kP2PWaves = 4;  // independent items per block

// r^2 + 1e-24 in place of max(r^2, 1e-24): equal for every r^2 above 3e-17
// (the addend is below half an ulp there), and both give 1e12 for r = 0.
__device__ __forceinline__ float p2p_r2(float4 xi, float4 s) {
  const float dx = xi.x - s.x, dy = xi.y - s.y, dz = xi.z - s.z;
  return fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 1e-24f)));
}

// acc + sum of q_j rsqrt(r_ij^2) over j in [j, end), j wave-uniform.  SELF
// (the chunk's own leaf): r^2 = inf for j == me, and rsq(inf) = 0.
template <bool SELF>
__device__ __forceinline__ float p2p_uniform(const float4* __restrict__ pts, int j, int end,
                                             float4 xi, int me, float acc) {
  constexpr int U = kP2PUnroll;
  for (; j + U <= end; j += U) {
    float4 s[U];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] = pts[j + u];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float r2 = p2p_r2(xi, s[u]);
      if (SELF) r2 = (j + u == me) ? INFINITY : r2;
      acc = fmaf(s[u].w, __builtin_amdgcn_rsqf(r2), acc);
    }
  }
  for (; j < end; ++j) {
    const float4 s = pts[j];
    float r2 = p2p_r2(xi, s);
    if (SELF) r2 = (j == me) ? INFINITY : r2;
    acc = fmaf(s.w, __builtin_amdgcn_rsqf(r2), acc);
  }
  return acc;
}

// The same over the points j0 + sl + t S of [j0, end) for slice sl of S; the
// step count is uniform, indices past the end are clamped and get r^2 = inf.
template <bool SELF>
__device__ __forceinline__ float p2p_sliced(const float4* __restrict__ pts, int j0, int end,
                                            int sl, int S, float4 xi, int me, float acc) {
  constexpr int U = kP2PUnroll / 2;
  const int nsteps = (end - j0 + S - 1) / S;
  const int jb = j0 + sl;
  int t = 0;
  for (; t + U <= nsteps; t += U) {
    float4 s[U];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] = pts[min(jb + (t + u) * S, end - 1)];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = jb + (t + u) * S;
      float r2 = p2p_r2(xi, s[u]);
      r2 = (j >= end || (SELF && j == me)) ? INFINITY : r2;
      acc = fmaf(s[u].w, __builtin_amdgcn_rsqf(r2), acc);
    }
  }
  for (; t < nsteps; ++t) {
    const int j = jb + t * S;
    const float4 s = pts[min(j, end - 1)];
    float r2 = p2p_r2(xi, s);
    r2 = (j >= end || (SELF && j == me)) ? INFINITY : r2;
    acc = fmaf(s.w, __builtin_amdgcn_rsqf(r2), acc);
  }
  return acc;
}

// One one-directional item.
__device__ __forceinline__ void p2p_one_item(const ed::P2PItem& it, const int2* __restrict__ seg,
                                             const float4* __restrict__ pts,
                                             float* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int ntc = it.ntc, S = it.slices;
  const int tl = lane % ntc, sl = lane / ntc;
  const bool act = sl < S;
  const int me = it.tgt0 + tl;
  const float4 xi = pts[me];
  float acc = 0.0f;
  int k = it.seg, off = it.off, left = it.cnt;
  while (left > 0) {
    const int2 r = seg[k];
    const int j0 = r.x + off;
    const int n = min(r.y - j0, left);
    const bool self = r.x <= it.tgt0 && it.tgt0 < r.y;
    if (S == 1)
      acc = self ? p2p_uniform<true>(pts, j0, j0 + n, xi, me, acc)
                 : p2p_uniform<false>(pts, j0, j0 + n, xi, me, acc);
    else
      acc = self ? p2p_sliced<true>(pts, j0, j0 + n, act ? sl : 0, S, xi, me, acc)
                 : p2p_sliced<false>(pts, j0, j0 + n, act ? sl : 0, S, xi, me, acc);
    left -= n;
    ++k;
    off = 0;
  }
  if (S > 1) {  // slices of target tl sit in lanes tl + q n_t, q < S
    float v = 0.0f;
    for (int q = 0; q < S; ++q) v += __shfl(acc, tl + q * ntc, 64);
    acc = v;
  }
  if (lane < ntc) part[it.out + lane] = acc;
}

// Column sums of a wave's 64 x 16 block: v[u] in lane l is row l, column u;
// lane l returns the sum over all 64 rows of column (l >> 2) & 15, added in a
// fixed order.  Six pairing steps, each between lanes that differ in one
// deciding bit of the column index: v_permlane32_swap (lane bit 5 picks
// columns u or u + 8), v_permlane16_swap (bit 4: + 4), DPP row_mirror (bit 3:
// + 2) and row_half_mirror (bit 2: + 1) each halve the columns a lane holds
// while adding its partner's copy; the last two (DPP quad_perm) add the four
// lanes that hold the same column.  35 VALU instructions for 1,024 entries.
__device__ __forceinline__ float p2p_colsum16(const float (&v)[16], int lane) {
  float w[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[u]),
                                                    __float_as_uint(v[u + 8]), false, false);
    w[u] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
  float x[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(w[u]),
                                                    __float_as_uint(w[u + 4]), false, false);
    x[u] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
  const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0;
  constexpr int kRowMirror = 0x140, kRowHalfMirror = 0x141, kQuadXor1 = 0xB1, kQuadXor2 = 0x4E;
  float y[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const float keep = b3 ? x[u + 2] : x[u], send = b3 ? x[u] : x[u + 2];
    y[u] = keep + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(send), kRowMirror,
                                                          0xF, 0xF, false));
  }
  const float keep = b2 ? y[1] : y[0], send = b2 ? y[0] : y[1];
  float z = keep + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(send), kRowHalfMirror,
                                                           0xF, 0xF, false));
  z += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(z), kQuadXor1, 0xF, 0xF, false));
  z += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(z), kQuadXor2, 0xF, 0xF, false));
  return z;
}

// One symmetric item: lane l holds the target its chunk's segments give it
// (lanes past nl carry charge 0 and write nothing); sources stream in groups
// of 16 at wave-uniform addresses (pts[n_pts] is the zero-charge sentinel).  Per pair one r^2 and one rsqrt k_ij: the
// target side adds q_j k_ij per lane, the source side sums q_i k_ij over the
// lanes (p2p_colsum16 per group).  Targets and sources are different
// leaves, so there is no j == i exclusion.
__device__ __forceinline__ void p2p_sym_item(const ed::P2PSymItem& it,
                                             const ed::P2PSymSeg* __restrict__ tseg,
                                             const float4* __restrict__ pts, int n_pts,
                                             float* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  int panel = 0, slot = -1;
  for (int k = 0; k < it.nseg; ++k) {  // segments in lane order: the last one at or below l
    const ed::P2PSymSeg g = tseg[it.seg0 + k];
    if (lane >= g.lane0) {
      panel = g.panel0 + (lane - g.lane0);
      slot = g.slot0 + (lane - g.lane0) + it.q * g.stride;
    }
  }
  if (lane >= it.nl) {
    panel = tseg[it.seg0].panel0;
    slot = -1;
  }
  const float4 xi = pts[panel];
  const float qi = slot >= 0 ? xi.w : 0.0f;
  const int s0 = it.s0, s1 = it.s0 + it.cnt;
  float acc = 0.0f;
  // one group of 16 sources: full groups read 16 consecutive points (no
  // masks: contiguous scalar loads), the last group reads the zero-charge
  // sentinel pts[n] past s1 (its column sums are not stored)
  auto group = [&](int j0, bool tail) {
    float4 sp[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) sp[u] = pts[tail && j0 + u >= s1 ? n_pts : j0 + u];
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const float k = __builtin_amdgcn_rsqf(p2p_r2(xi, sp[u]));
      acc = fmaf(sp[u].w, k, acc);
      v[u] = qi * k;
    }
    const float csum = p2p_colsum16(v, lane);
    const int u = (lane >> 2) & 15;
    if ((lane & 3) == 0 && j0 + u < s1) part[it.outS + (j0 - s0) + u] = csum;
  };
  int j0 = s0;
  for (; j0 + 16 <= s1; j0 += 16) group(j0, false);
  if (j0 < s1) group(j0, true);
  if (slot >= 0) part[slot] = acc;
}

// One-right-hand-side P2P: item blocks [0, nb_sym) run the symmetric items
// (the heavier ones first), the rest the one-directional items, each block
// walking them with a grid stride.  When the far-field chain is the longer
// branch the grid is capped at kP2PBlocksPerCU blocks per compute unit (the
// plan decides, see ed_plan.h): with every SIMD's register file full of P2P
// waves the chain's kernels waited for P2P waves to retire (P2M 8.3 us
// alone, 17 us beside the uncapped P2P).  82k sphere matvec: 94.6 us
// uncapped, 87.0 us at five blocks per CU (three: 92.6, four: 88.6, six:
// 97.5, eight: 98.7); the 300k torus, where the P2P is the longer branch,
// takes 244 us uncapped and 292 us capped at five.
// (No waves-per-EU cap: at 8 waves the compiler holds the SGPR budget so
// low that the 64 source SGPRs of a group spill to VGPR lanes; uncapped it
// needs 52 VGPRs, which still gives 8 waves per SIMD.)
__global__ __launch_bounds__(64 * kP2PWaves) void k_p2p_sym(
    const ed::P2PSymItem* __restrict__ sym, const ed::P2PSymSeg* __restrict__ tseg, int n_sym,
    int nb_sym,
    const ed::P2PItem* __restrict__ items, int n_items, const int2* __restrict__ seg,
    const float4* __restrict__ pts, int n_pts, float* __restrict__ part) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nb_all = nb_sym + (n_items + kP2PWaves - 1) / kP2PWaves;
  // a capped grid walks the item blocks (each wave on its own; no barriers)
  for (int blk = blockIdx.x; blk < nb_all; blk += gridDim.x) {
    if (blk < nb_sym) {
      const int item = blk * kP2PWaves + wave;
      if (item < n_sym) p2p_sym_item(sym[item], tseg, pts, n_pts, part);
    } else {
      const int item = (blk - nb_sym) * kP2PWaves + wave;
      if (item < n_items) p2p_one_item(items[item], seg, pts, part);
    }
  }
}

// P2P for R right-hand sides: the same items and schedule; r^2 and its
// rsqrt are computed once per interaction and weight the R charges of the
// source (qT, point-major, read at wave-uniform addresses with the point).
// Per RHS the sums run in the same order as the one-RHS kernel, so each
// RHS's partials equal that kernel's bit for bit.  Partial rows of RHS r sit
// at part + r * part_stride.
template <bool SELF, int R>
__device__ __forceinline__ void p2p_uniform_multi(const float4* __restrict__ pts,
                                                  const float* __restrict__ qT, int j, int end,
                                                  float4 xi, int me, float (&acc)[R]) {
  constexpr int U = 4;
  for (; j + U <= end; j += U) {
    float4 s[U];
    float q[U][R];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      s[u] = pts[j + u];
#pragma unroll
      for (int r = 0; r < R; ++r) q[u][r] = qT[(size_t)(j + u) * R + r];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float r2 = p2p_r2(xi, s[u]);
      if (SELF) r2 = (j + u == me) ? INFINITY : r2;
      const float k = __builtin_amdgcn_rsqf(r2);
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = fmaf(q[u][r], k, acc[r]);
    }
  }
  for (; j < end; ++j) {
    const float4 s = pts[j];
    float r2 = p2p_r2(xi, s);
    if (SELF) r2 = (j == me) ? INFINITY : r2;
    const float k = __builtin_amdgcn_rsqf(r2);
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = fmaf(qT[(size_t)j * R + r], k, acc[r]);
  }
}

template <bool SELF, int R>
__device__ __forceinline__ void p2p_sliced_multi(const float4* __restrict__ pts,
                                                 const float* __restrict__ qT, int j0, int end,
                                                 int sl, int S, float4 xi, int me,
                                                 float (&acc)[R]) {
  const int nsteps = (end - j0 + S - 1) / S;
  const int jb = j0 + sl;
  for (int t = 0; t < nsteps; ++t) {
    const int j = jb + t * S;
    const int jc = min(j, end - 1);
    const float4 s = pts[jc];
    float r2 = p2p_r2(xi, s);
    r2 = (j >= end || (SELF && j == me)) ? INFINITY : r2;
    const float k = __builtin_amdgcn_rsqf(r2);
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = fmaf(qT[(size_t)jc * R + r], k, acc[r]);
  }
}

template <int R>
__global__ __launch_bounds__(64 * kP2PWaves) void k_p2p_multi(const ed::P2PItem* __restrict__ items,
                                                              int n_items,
                                                              const int2* __restrict__ seg,
                                                              const float4* __restrict__ pts,
                                                              const float* __restrict__ qT,
                                                              float* __restrict__ part,
                                                              int64_t part_stride) {
  const int item = blockIdx.x * kP2PWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= n_items) return;  // whole waves only
  const ed::P2PItem it = items[item];
  const int lane = threadIdx.x & 63;
  const int ntc = it.ntc, S = it.slices;
  const int tl = lane % ntc, sl = lane / ntc;
  const bool act = sl < S;
  const int me = it.tgt0 + tl;
  const float4 xi = pts[me];
  float acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0f;
  int k = it.seg, off = it.off, left = it.cnt;
  while (left > 0) {
    const int2 rg = seg[k];
    const int j0 = rg.x + off;
    const int n = min(rg.y - j0, left);
    const bool self = rg.x <= it.tgt0 && it.tgt0 < rg.y;
    if (S == 1) {
      if (self) p2p_uniform_multi<true, R>(pts, qT, j0, j0 + n, xi, me, acc);
      else p2p_uniform_multi<false, R>(pts, qT, j0, j0 + n, xi, me, acc);
    } else {
      if (self) p2p_sliced_multi<true, R>(pts, qT, j0, j0 + n, act ? sl : 0, S, xi, me, acc);
      else p2p_sliced_multi<false, R>(pts, qT, j0, j0 + n, act ? sl : 0, S, xi, me, acc);
    }
    left -= n;
    ++k;
    off = 0;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    float a = acc[r];
    if (S > 1) {  // slices of target tl sit in lanes tl + q n_t, q < S
      float v = 0.0f;
      for (int q = 0; q < S; ++q) v += __shfl(a, tl + q * ntc, 64);
      a = v;
    }
    if (lane < ntc) part[r * part_stride + it.out + lane] = a;
  }
}

// gather for R right-hand sides (sigma R x n, original or tree order): sigma
// of each RHS into sig_m (R x n), the charges point-major into qT (n x R), the
// self term once, and every RHS's multipole-maximum slots reset.
template <typename TS>
__global__ void k_gather_multi(int n, int R, const int32_t* __restrict__ perm,
                               const TS* __restrict__ sigma, int sigma_tree,
                               const TS* __restrict__ selfi, const float* __restrict__ area_t,
                               float* __restrict__ sig_m, float* __restrict__ qT,
                               float* __restrict__ selfc_t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (i >= n) return;
  const int o = perm[i];
  const float s = (float)sigma[(size_t)r * n + (sigma_tree ? i : o)];
  sig_m[(size_t)r * n + i] = s;
  qT[(size_t)i * R + r] = s * area_t[i];
  if (r == 0) selfc_t[i] = selfi ? (float)selfi[o] : 0.0f;
}

// Part Q (of four) of the L2P sum: the orders m with l2p_owner(m) == Q (a
// snake over m, so that the four parts hold about equal numbers of terms:
// at p = 10 {0,7,8}, {1,6,9}, {2,5,10}, {3,4}); every part runs the diagonal
// recurrence R_m^m itself.
constexpr int l2p_owner(int m) { return (m % 8) < 4 ? m % 8 : 7 - m % 8; }

template <int P, int Q>
__device__ __forceinline__ float l2p_part(float x, float y, float z, const float* __restrict__ Lc) {
  const float r2 = x * x + y * y + z * z;
  float pmr = 1.0f, pmi = 0.0f;
  float acc0 = 0.0f, acc1 = 0.0f;
#pragma unroll
  for (int m = 0; m <= P; ++m) {
    if (m > 0) {
      const float f = -0.5f / (float)m;
      const float nr = f * (x * pmr - y * pmi);
      const float ni = f * (x * pmi + y * pmr);
      pmr = nr;
      pmi = ni;
    }
    if (l2p_owner(m) == Q) {  // folded per m by the full unroll
      float ar = 0.0f, ai = 0.0f, br = pmr, bi = pmi;
#pragma unroll
      for (int n = m; n <= P; ++n) {
        if (m == 0) {
          acc0 += Lc[n * n] * br;
        } else {
          acc1 += Lc[n * n + 2 * m - 1] * br - Lc[n * n + 2 * m] * bi;
        }
        if (n < P) {
          const float inv = 1.0f / (float)((n + 1) * (n + 1) - m * m);
          const float c = (float)(2 * n + 1) * z;
          const float nr = (c * br - r2 * ar) * inv;
          const float ni = (c * bi - r2 * ai) * inv;
          ar = br; ai = bi; br = nr; bi = ni;
        }
      }
    }
  }
  return acc0 + 2.0f * acc1;
}

// Per panel (tree order): far field by L2P from its leaf's local expansion,
// the P2P sum, the near-field correction (CSR row), the self term; output
// in tree or original order.  Non-finite results become 0
// (fmm3d/bem_fmm.py:741-749).  A block of four waves takes 64 panels: wave q
// evaluates part q of the L2P (l2p_part), the near-correction entries and
// partial rows q, q + 4, ... of each panel (a quarter of the dependent loads
// per lane and four times the waves of one thread per panel), and wave 0
// adds the four parts from LDS in order (0 + 1) + (2 + 3).
template <int P, typename TO, bool MULTI>
__global__ __launch_bounds__(256) void k_l2p_out(
    int n, const int32_t* __restrict__ pt_node, const float4* __restrict__ ctr,
    const float4* __restrict__ pts, const float* __restrict__ L,
    const int2* __restrict__ p2p_ref, const float* __restrict__ p2p_part,
    const int2* __restrict__ node_rng, const int2* __restrict__ p2s_leaf,
    const int32_t* __restrict__ near_ptr, const int32_t* __restrict__ near_col,
    const float* __restrict__ near_w, const float* __restrict__ sig_t,
    const float* __restrict__ near_t, const float* __restrict__ selfc_t,
    const int32_t* __restrict__ perm, int out_tree, float ke, TO* __restrict__ out,
    RhsStride rs) {
  constexpr int NC = (P + 1) * (P + 1);
  __shared__ float part4[4][64];
  const int ry = MULTI ? blockIdx.y : 0;
  if constexpr (MULTI) {
    L += ry * rs.node;
    p2p_part += ry * rs.part;
    sig_t += ry * rs.n;
    out += ry * rs.n;
  }
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = blockIdx.x * 64 + lane;
  const int i = min(i0, n - 1);  // lanes past n compute a copy and write nothing
  const int node = pt_node[i];
  const float4 xi = pts[i];
  float corr = 0.0f;
  // several RHS: this wave's near-field entries nb + q, nb + q + 4, ... in
  // groups of four with clamped indices and zero weights past the row end
  // (every load of a group issued before the first wait); one RHS: k_near's
  // sum, added once below
  const int nb = MULTI ? near_ptr[i] : 0, ne = MULTI ? near_ptr[i + 1] : 0;
  for (int e0 = nb + q; e0 < ne; e0 += 16) {
    int c[4];
    float wq[4], sv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = min(e0 + 4 * u, ne - 1);
      c[u] = near_col[e];
      const float wv = near_w[e];
      wq[u] = e0 + 4 * u < ne ? wv : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) sv[u] = sig_t[c[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) corr += wq[u] * sv[u];
  }
  const float4 c = ctr[node];
  const float x = (xi.x - c.x) * c.w, y = (xi.y - c.y) * c.w, z = (xi.z - c.z) * c.w;
  const float* Lc = L + (size_t)node * NC;
  float far;
  switch (q) {
    case 0: far = l2p_part<P, 0>(x, y, z, Lc); break;
    case 1: far = l2p_part<P, 1>(x, y, z, Lc); break;
    case 2: far = l2p_part<P, 2>(x, y, z, Lc); break;
    default: far = l2p_part<P, 3>(x, y, z, Lc); break;
  }
  float near = 0.0f;
  if constexpr (MULTI) {  // the one-directional schedule: rows of 64 per item
    const int2 pr = p2p_ref[i];
    for (int k = q; k < pr.y; k += 16) {  // four loads in flight, added in order
      float a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = k + 4 * u < pr.y ? p2p_part[pr.x + 64 * (k + 4 * u)] : 0.0f;
#pragma unroll
      for (int u = 0; u < 4; ++u) near += a[u];
    }
  } else {  // the symmetric schedule: K rows of n_L slots per leaf, in order
    const int2 lr = p2s_leaf[node], nr = node_rng[node];
    const int nl = nr.y - nr.x;
    const float* pp = p2p_part + lr.x + (i - nr.x);
    for (int k = q; k < lr.y; k += 32) {  // eight loads in flight, added in order
      float a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = k + 4 * u < lr.y ? pp[(size_t)(k + 4 * u) * nl] : 0.0f;
#pragma unroll
      for (int u = 0; u < 8; ++u) near += a[u];
    }
  }
  part4[q][lane] = ke * (far + near) + corr;
  __syncthreads();
  if (q != 0 || i0 >= n) return;
  const float qi = MULTI ? rs.qT[(size_t)i * rs.R + ry] : xi.w;
  float v = ((part4[0][lane] + part4[1][lane]) + (part4[2][lane] + part4[3][lane])) +
            selfc_t[i] * qi;
  if constexpr (!MULTI) v += near_t[i];
  if (!isfinite(v)) v = 0.0f;
  out[out_tree ? i : perm[i]] = (TO)v;
}

// ---------------------------------------------------------------------------
// Host launcher
// ---------------------------------------------------------------------------
#define ED_LAUNCH_CHECK(msg)                                                   \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) return fail(ED_ERR_LAUNCH, std::string(msg) + ": " + \
                                                         hipGetErrorString(e_)); \
  } while (0)

// Stage boundary markers for ed_fmm_profile (no-op unless profiling).
#define ED_MARK(k)                                                   \
  do {                                                               \
    if (Pn->prof_events) (void)hipEventRecord(Pn->prof_events[k], st); \
  } while (0)

#define ED_HIP_LAUNCH(call, msg)                                              \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) return fail(ED_ERR_LAUNCH, std::string(msg) + ": " + \
                                                         hipGetErrorString(e_)); \
  } while (0)

static void p2p_launch(ed_plan* Pn, int R, const RhsStride& rs, hipStream_t s) {
  const unsigned nb = (unsigned)((Pn->n_p2p_items + kP2PWaves - 1) / kP2PWaves);
  if (R == 1) {  // the symmetric schedule
    const int nb_sym = (int)((Pn->n_p2s_sym + kP2PWaves - 1) / kP2PWaves);
    const int nb_one = (int)((Pn->n_p2s_items + kP2PWaves - 1) / kP2PWaves);
    const int cap = Pn->p2p_blocks_per_cu * std::max(Pn->n_cu, 1);
    const int grid = cap > 0 ? std::min(nb_sym + nb_one, cap) : nb_sym + nb_one;
    if (grid > 0)
      hipLaunchKernelGGL(k_p2p_sym, dim3((unsigned)grid), dim3(64 * kP2PWaves), 0, s,
                         (const ed::P2PSymItem*)Pn->p2s_sym, (const ed::P2PSymSeg*)Pn->p2s_tseg,
                         (int)Pn->n_p2s_sym,
                         nb_sym,
                         (const ed::P2PItem*)Pn->p2s_items, (int)Pn->n_p2s_items, Pn->p2s_seg,
                         Pn->pts, (int)Pn->n, Pn->p2s_part);
    return;
  }
  const auto* items = (const ed::P2PItem*)Pn->p2p_items;
  const int ni = (int)Pn->n_p2p_items;
  float* part = Pn->multi.part;
  switch (R) {
#define ED_P2P_R(RR)                                                                        \
  case RR:                                                                                  \
    hipLaunchKernelGGL(k_p2p_multi<RR>, dim3(nb), dim3(64 * kP2PWaves), 0, s, items, ni,    \
                       Pn->p2p_seg, Pn->pts, rs.qT, part, rs.part);                         \
    break;
    ED_P2P_R(2) ED_P2P_R(3) ED_P2P_R(4) ED_P2P_R(5) ED_P2P_R(6) ED_P2P_R(7) ED_P2P_R(8)
#undef ED_P2P_R
    default: break;  // R is checked by the caller
  }
}

// One matvec for R right-hand sides (R = 1: the plan's own buffers; R > 1:
// the multi-RHS buffers, blockIdx.y = RHS in every far-field launch).
template <int P, bool MULTI>
static int fmm_run(ed_plan* Pn, const void* sigma, int sig64, int sig_tree, const void* selfi,
                   void* out, int out64, int out_tree, int R, hipStream_t st) {
  constexpr int NC = (P + 1) * (P + 1);
  const int n = (int)Pn->n;
  constexpr bool multi = MULTI;
  RhsStride rs;
  if (multi) {
    rs.node = Pn->n_nodes * NC;
    rs.pair = Pn->multi.pair_stride;
    rs.slot = (int64_t)kMaxSlots * kMaxSlotStride;
    rs.part = 64 * std::max<int64_t>(Pn->n_p2p_items, 1);
    rs.n = n;
    rs.qT = Pn->multi.qT;
    rs.R = R;
  }
  float* Mb = multi ? Pn->multi.M : Pn->M;
  float* Lb = multi ? Pn->multi.L : Pn->L;
  float* Yb = multi ? Pn->multi.ym2m : Pn->ym2m;   // own parts of the local expansions
  float* Yp = multi ? Pn->multi.ypair : Pn->ypair;
  float* sigb = multi ? Pn->multi.sig : Pn->sig_t;
  float* partb = multi ? Pn->multi.part : Pn->p2s_part;
  const unsigned RY = (unsigned)R;
  ED_MARK(0);
  const int gb = (n + 255) / 256;
  // M2L form: -1 VALU, 0 FP32 MFMA, 1 / 2 / 3 BF16 / FP16 / split FP16 MFMA;
  // the 16-bit forms scale their inputs by the largest |M * ascale|, which
  // P2M and M2M fold into mmax as they write M
  // (m2l_prec_override 3 on a 16-bit plan: split FP16 through its t_m2l_h3)
  const bool split_override = Pn->m2l_prec_override == 3 && Pn->t_m2l_h3 != nullptr;
  const int prec = Pn->n_m2l_tiles > 0 ? (split_override ? 3 : Pn->m2l_precision) : -1;
  const uint16_t* th = split_override ? Pn->t_m2l_h3 : Pn->t_m2l_h;
  unsigned* mmax = (prec >= 1 && prec <= 3) ? (multi ? Pn->multi.mmax : Pn->m2l_mmax) : nullptr;
  // gather (tree-order sigma, P2P charges, self term) and P2P depend on
  // nothing else of the matvec: off the profiling path they run on the side
  // stream beside the far-field passes, whose P2M reads sigma itself.  (P2P
  // forked after P2M instead became the longer branch: 0.125 vs 0.114 ms per
  // matvec at 82k panels.)
  const bool fork = Pn->prof_events == nullptr && Pn->n_p2p_items > 0;
  auto gather = [&](hipStream_t gs) -> int {
    if (multi) {
      if (sig64)
        hipLaunchKernelGGL(k_gather_multi<double>, dim3(gb, RY), dim3(256), 0, gs, n, R, Pn->perm,
                           (const double*)sigma, sig_tree, (const double*)selfi, Pn->area_t,
                           sigb, Pn->multi.qT, Pn->selfc_t);
      else
        hipLaunchKernelGGL(k_gather_multi<float>, dim3(gb, RY), dim3(256), 0, gs, n, R, Pn->perm,
                           (const float*)sigma, sig_tree, (const float*)selfi, Pn->area_t, sigb,
                           Pn->multi.qT, Pn->selfc_t);
    } else if (sig64) {
      hipLaunchKernelGGL(k_gather<double>, dim3(gb), dim3(256), 0, gs, n, Pn->perm,
                         (const double*)sigma, sig_tree, (const double*)selfi, Pn->pts,
                         Pn->area_t, Pn->sig_t, Pn->selfc_t);
    } else {
      hipLaunchKernelGGL(k_gather<float>, dim3(gb), dim3(256), 0, gs, n, Pn->perm,
                         (const float*)sigma, sig_tree, (const float*)selfi, Pn->pts,
                         Pn->area_t, Pn->sig_t, Pn->selfc_t);
    }
    ED_LAUNCH_CHECK("gather");
    return ED_OK;
  };
  auto near = [&](hipStream_t ns) -> int {
    if (!multi) {
      const dim3 gn((unsigned)((4 * (int64_t)n + 255) / 256));
      if (sig64)
        hipLaunchKernelGGL(k_near<double>, gn, dim3(256), 0, ns, n, Pn->near_ptr, Pn->near_col,
                           Pn->near_w, Pn->perm, (const double*)sigma, sig_tree, Pn->near_t);
      else
        hipLaunchKernelGGL(k_near<float>, gn, dim3(256), 0, ns, n, Pn->near_ptr, Pn->near_col,
                           Pn->near_w, Pn->perm, (const float*)sigma, sig_tree, Pn->near_t);
      ED_LAUNCH_CHECK("near");
    }
    return ED_OK;
  };
  // the join: a signal (stream memory write and wait) when the P2P is the
  // longer branch (the plan leaves its grid uncapped), else an event, and an
  // event under graph capture, where a replay could not advance the sequence
  // number the wait compares against.  300k-panel torus, P2P longer: 245.5
  // -> 241.3 us per matvec with the signal; 82k sphere, far-field chain
  // longer: 87.8 -> 88.5 us, so it keeps the event there.
  bool use_sig = false;
  uint64_t seq = 0;
  if (fork) {
    if (Pn->sync_join && Pn->p2p_blocks_per_cu == 0) {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      ED_HIP_LAUNCH(hipStreamIsCapturing(st, &cs), "p2p fork");
      use_sig = cs == hipStreamCaptureStatusNone;
    }
    ED_HIP_LAUNCH(hipEventRecord(Pn->ev_fork, st), "p2p fork");
    ED_HIP_LAUNCH(hipStreamWaitEvent(Pn->side, Pn->ev_fork, 0), "p2p fork");
    if (int rc = gather(Pn->side); rc != ED_OK) return rc;
    p2p_launch(Pn, R, rs, Pn->side);
    ED_LAUNCH_CHECK("p2p");
    if (use_sig) {
      seq = Pn->sync_seq + 1;
      ED_HIP_LAUNCH(hipStreamWriteValue64(Pn->side, Pn->sync_join, seq, 0), "p2p join");
      Pn->sync_seq = seq;
    } else {
      ED_HIP_LAUNCH(hipEventRecord(Pn->ev_join, Pn->side), "p2p join");
    }
  } else {
    if (int rc = gather(st); rc != ED_OK) return rc;
    if (int rc = near(st); rc != ED_OK) return rc;
  }
  ED_MARK(1);
  const int nleaf = (int)Pn->n_leaves;
  const auto& H = Pn->host;
  const int depth = H.depth;
  if (depth >= 2) {
    const int node2 = (int)H.level_begin[2];
    if (sig64)
      hipLaunchKernelGGL((k_p2m<P, MULTI, double>), dim3(nleaf, RY), dim3(kP2MThreads), 0, st,
                         Pn->leaves, Pn->node_rng, Pn->node_ctr, Pn->pts, Pn->perm,
                         (const double*)sigma, sig_tree, Pn->area_t, node2, Mb, Pn->m2l_ascale,
                         mmax, rs);
    else
      hipLaunchKernelGGL((k_p2m<P, MULTI, float>), dim3(nleaf, RY), dim3(kP2MThreads), 0, st,
                         Pn->leaves, Pn->node_rng, Pn->node_ctr, Pn->pts, Pn->perm,
                         (const float*)sigma, sig_tree, Pn->area_t, node2, Mb, Pn->m2l_ascale,
                         mmax, rs);
    ED_LAUNCH_CHECK("p2m");
    // upward pass: parents at level l (children at l+1), deepest first
    for (int l = depth - 1; l >= 2; --l) {
      const int64_t pb = Pn->parents_begin[l], pe = Pn->parents_begin[l + 1];
      if (pe > pb) {
        if constexpr (OctLds<P>::ON)
          hipLaunchKernelGGL((k_m2m_lds<P, MULTI>), dim3((unsigned)(pe - pb), RY), dim3(256), 0,
                             st, Pn->parents_list + pb, Pn->first_child, Pn->n_child,
                             Pn->node_octant, Pn->t_m2m7, Pn->m2l_flip, Mb,
                             Pn->m2l_ascale, mmax, rs);
        else
          hipLaunchKernelGGL((k_m2m<P, MULTI>), dim3((unsigned)(pe - pb), RY), dim3(256), 0, st,
                             Pn->parents_list + pb, Pn->first_child, Pn->n_child,
                             Pn->node_octant, Pn->t_m2m, Mb, Pn->m2l_ascale, mmax, rs);
        ED_LAUNCH_CHECK("m2m");
      }
    }
    ED_MARK(2);
    // M2L for all nodes at levels >= 2
    const int node0 = node2;
    const int nm = (int)(Pn->n_nodes - node0);
    bool valu = prec < 0;
    if constexpr (!M2LShape<P>::ENABLED) valu = true;  // tables beyond LDS: VALU form only
    if (valu) {
      // each node's own part into Yb
      hipLaunchKernelGGL((k_m2l_valu<P, MULTI>), dim3(nm, RY), dim3(128), 0, st, node0, Pn->v_ptr,
                         Pn->v_src, Pn->v_tab, Pn->node_ctr, Pn->t_m2l, Mb, Yb, rs);
      ED_LAUNCH_CHECK("m2l_valu");
    } else if constexpr (M2LShape<P>::ENABLED) {
      const dim3 tiles((unsigned)Pn->n_m2l_tiles, RY);
      if (prec == 3) {
        if constexpr (MULTI)  // one block per tile runs every RHS (table loaded once)
          hipLaunchKernelGGL((k_m2l_mfma16x3_multi<P>), dim3((unsigned)Pn->n_m2l_tiles),
                             dim3(256), 0, st, Pn->m2l_tile_tab, Pn->m2l_tile_beg,
                             Pn->m2l_tile_cnt, Pn->m2l_pair_src, Pn->m2l_pair_sgn, Pn->m2l_flip,
                             (const _Float16*)th, Pn->m2l_ascale, Mb, mmax, Yp, rs);
        else
          hipLaunchKernelGGL((k_m2l_mfma16x3<P>), dim3((unsigned)Pn->n_m2l_tiles), dim3(256), 0,
                             st, Pn->m2l_tile_tab, Pn->m2l_tile_beg, Pn->m2l_tile_cnt,
                             Pn->m2l_pair_src, Pn->m2l_pair_sgn, Pn->m2l_flip,
                             (const _Float16*)th, Pn->m2l_ascale, Mb, mmax, Yp);
      }
      else if (prec == 1)
        hipLaunchKernelGGL((k_m2l_mfma16<P, __bf16, MULTI>), tiles, dim3(256), 0, st, Pn->m2l_tile_tab,
                           Pn->m2l_tile_beg, Pn->m2l_tile_cnt, Pn->m2l_pair_src,
                           Pn->m2l_pair_sgn, Pn->m2l_flip, (const __bf16*)Pn->t_m2l_h,
                           Pn->m2l_ascale, Mb, mmax, Yp, rs);
      else if (prec == 2)
        hipLaunchKernelGGL((k_m2l_mfma16<P, _Float16, MULTI>), tiles, dim3(256), 0, st,
                           Pn->m2l_tile_tab, Pn->m2l_tile_beg, Pn->m2l_tile_cnt,
                           Pn->m2l_pair_src, Pn->m2l_pair_sgn, Pn->m2l_flip,
                           (const _Float16*)Pn->t_m2l_h, Pn->m2l_ascale, Mb, mmax, Yp, rs);
      else if constexpr (M2LShape<P>::ENABLED)
        hipLaunchKernelGGL((k_m2l_mfma<P, MULTI>), tiles, dim3(M2LShape<P>::THREADS), 0, st,
                           Pn->m2l_tile_tab, Pn->m2l_tile_beg, Pn->m2l_tile_cnt,
                           Pn->m2l_pair_src, Pn->m2l_pair_sgn, Pn->m2l_flip, Pn->t_m2l_mfma,
                           Pn->m2l_ascale, Mb, Yp, rs);
      ED_LAUNCH_CHECK("m2l_mfma");
    }
    ED_MARK(3);
    if constexpr (M2LShape<P>::ENABLED) {
      if (!valu) {
        hipLaunchKernelGGL((k_m2l_reduce<P, MULTI>), dim3(nm, RY), dim3(128), 0, st, node0,
                           Pn->m2l_tgt_ptr, Pn->m2l_tgt_row, Pn->node_ctr, Pn->m2l_binv, Yp, Yb,
                           mmax, rs);
        ED_LAUNCH_CHECK("m2l_reduce");
      }
    }
    ED_MARK(4);
    // complete local expansions (own part + the ancestors', translated)
    hipLaunchKernelGGL((k_m2l_down<P, MULTI>), dim3(nm, RY), dim3(128), 0, st, node0,
                       Pn->node_parent, Pn->node_octant, Yb, Pn->t_l2l, Lb, rs);
    ED_LAUNCH_CHECK("m2l_down");
    ED_MARK(5);
  } else {
    ED_MARK(2);
    ED_MARK(3);
    ED_MARK(4);
    ED_MARK(5);
  }
  if (fork) {
    // the near-field correction after the far-field chain, while the P2P
    // (the longer branch) finishes on the side stream
    if (int rc = near(st); rc != ED_OK) return rc;
    if (use_sig)
      ED_HIP_LAUNCH(hipStreamWaitValue64(st, Pn->sync_join, seq, hipStreamWaitValueGte, ~0ull),
                    "p2p join");
    else
      ED_HIP_LAUNCH(hipStreamWaitEvent(st, Pn->ev_join, 0), "p2p join");
  } else if (Pn->n_p2p_items > 0) {
    p2p_launch(Pn, R, rs, st);
    ED_LAUNCH_CHECK("p2p");
  }
  const float ke = (float)kKE;
  const unsigned gb64 = (unsigned)((n + 63) / 64);  // k_l2p_out: 64 panels a block
  ED_MARK(7);
  if (out64)
    hipLaunchKernelGGL((k_l2p_out<P, double, MULTI>), dim3(gb64, RY), dim3(256), 0, st, n, Pn->pt_node,
                       Pn->node_ctr, Pn->pts, Lb, Pn->p2p_ref, partb, Pn->node_rng,
                       Pn->p2s_leaf, Pn->near_ptr, Pn->near_col,
                       Pn->near_w, sigb, Pn->near_t, Pn->selfc_t, Pn->perm, out_tree, ke,
                       (double*)out, rs);
  else
    hipLaunchKernelGGL((k_l2p_out<P, float, MULTI>), dim3(gb64, RY), dim3(256), 0, st, n, Pn->pt_node,
                       Pn->node_ctr, Pn->pts, Lb, Pn->p2p_ref, partb, Pn->node_rng,
                       Pn->p2s_leaf, Pn->near_ptr, Pn->near_col,
                       Pn->near_w, sigb, Pn->near_t, Pn->selfc_t, Pn->perm, out_tree, ke,
                       (float*)out, rs);
  ED_LAUNCH_CHECK("l2p_out");
  ED_MARK(8);
  (void)NC;
  return ED_OK;
}

int fmm_launch(ed_plan* Pn, const void* sigma, int sig64, int sig_tree, const void* selfi,
               void* out, int out64, int out_tree, hipStream_t st, int nrhs) {
#define ED_RUN(PP)                                                                             \
  return nrhs > 1 ? fmm_run<PP, true>(Pn, sigma, sig64, sig_tree, selfi, out, out64, out_tree, \
                                      nrhs, st)                                                \
                  : fmm_run<PP, false>(Pn, sigma, sig64, sig_tree, selfi, out, out64, out_tree, \
                                       1, st);
  switch (Pn->p) {
    case 4: ED_RUN(4)
    case 6: ED_RUN(6)
    case 8: ED_RUN(8)
    case 10: ED_RUN(10)
    case 12: ED_RUN(12)
    case 14: ED_RUN(14)
    case 16: ED_RUN(16)
    default: return fail(ED_ERR_ARG, "expansion order must be one of 4, 6, 8, 10, 12, 14, 16");
  }
#undef ED_RUN
}

int m2l_shape(int p, int prec, int* ncp, int* kp, int* cols, int* enabled) {
  switch (p) {
#define ED_SHAPE(PP)                                                          \
  case PP:                                                                    \
    *ncp = M2LShape<PP>::NCP;                                                 \
    if (prec == 0) {                                                          \
      *kp = M2LShape<PP>::KP; *cols = M2LShape<PP>::COLS;                     \
      *enabled = M2LShape<PP>::ENABLED ? 1 : 0;                               \
    } else {                                                                  \
      *kp = M2LShapeH<PP>::KP; *cols = M2LShapeH<PP>::COLS;                   \
      *enabled = prec > 0 ? 1 : 0;                                            \
    }                                                                         \
    return ED_OK;
    ED_SHAPE(4) ED_SHAPE(6) ED_SHAPE(8) ED_SHAPE(10) ED_SHAPE(12)
#undef ED_SHAPE
    case 14:
    case 16:  // M2L tables beyond the LDS: VALU form only
      *ncp = (p + 1) * (p + 1);
      *kp = *ncp;
      *cols = 0;
      *enabled = 0;
      return ED_OK;
    default: return fail(ED_ERR_ARG, "unsupported order");
  }
}

}  // namespace ed

// Per-stage device time of the FMM matvec (FP64 in/out, original order),
// averaged over `reps` launches, between the marks the launcher records
// (0 start, 1 after gather, 2 after the upward pass, 3 after the M2L GEMM,
// 4 after the M2L sums, 5 after the downward pass, 7 after P2P, 8 end):
// stage_ms[0..7] = gather, upward (P2M + M2M), M2L, M2L sums, downward
// (L2L), 0, P2P (inline here), l2p_out (L2P + near + self + output).
static int fmm_profile(ed_plan* P, const double* sigma, const double* selfi, double* out,
                       int nrhs, int32_t reps, double* stage_ms, int32_t n_stages, hipStream_t st) {
  using namespace ed;
  constexpr int NE = 9;  // mark slots
  constexpr int from[8] = {0, 1, 2, 3, 4, -1, 5, 7};
  constexpr int to[8] = {1, 2, 3, 4, 5, -1, 7, 8};
  hipEvent_t ev[NE];
  for (int k = 0; k < NE; ++k)
    if (hipEventCreate(&ev[k]) != hipSuccess) return fail(ED_ERR_DEVICE, "ed_fmm_profile: event");
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int rc = ED_OK;
  P->prof_events = ev;
  for (int r = 0; r < reps && rc == ED_OK; ++r) {
    rc = fmm_launch(P, sigma, 1, 0, selfi, out, 1, 0, st, nrhs);
    if (rc != ED_OK) break;
    if (hipEventSynchronize(ev[NE - 1]) != hipSuccess) { rc = fail(ED_ERR_DEVICE, "ed_fmm_profile: sync"); break; }
    for (int k = 0; k < 8; ++k) {
      if (from[k] < 0) continue;
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ev[from[k]], ev[to[k]]);
      acc[k] += ms;
    }
  }
  P->prof_events = nullptr;
  for (int k = 0; k < NE; ++k) (void)hipEventDestroy(ev[k]);
  for (int k = 0; k < n_stages && k < 8; ++k) stage_ms[k] = acc[k] / reps;
  return rc;
}

extern "C" int ed_fmm_profile(ed_plan* P, const double* sigma, const double* selfi, double* out,
                              int32_t reps, double* stage_ms, int32_t n_stages, ed_stream stream) {
  using namespace ed;
  if (!P || !sigma || !out || !stage_ms || reps < 1) return fail(ED_ERR_ARG, "ed_fmm_profile: bad argument");
  return fmm_profile(P, sigma, selfi, out, 1, reps, stage_ms, n_stages, (hipStream_t)stream);
}

// The same for the multi-RHS matvec of nrhs (1..8) right-hand sides.
extern "C" int ed_fmm_profile_multi(ed_plan* P, const double* sigma, const double* selfi,
                                    double* out, int32_t nrhs, int32_t reps, double* stage_ms,
                                    int32_t n_stages, ed_stream stream) {
  using namespace ed;
  if (!P || !sigma || !out || !stage_ms || reps < 1 || nrhs < 1 || nrhs > 8)
    return fail(ED_ERR_ARG, "ed_fmm_profile_multi: bad argument");
  if (nrhs > 1) {
    const int rc = multi_reserve(P, nrhs);
    if (rc != ED_OK) return rc;
  }
  return fmm_profile(P, sigma, selfi, out, nrhs, reps, stage_ms, n_stages, (hipStream_t)stream);
}
