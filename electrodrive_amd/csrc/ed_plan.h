// Device-resident FMM plan (one per geometry).  All buffers are allocated at
// ed_plan_create; ed_fmm_matvec_* allocates nothing.
#pragma once

#include <hip/hip_runtime.h>

#include "ed_internal.h"

namespace ed {
struct DenseItem;
struct DenseSymDevice;
}

struct ed_plan {
  int device = 0;
  int p = 8, nc = 81, ws = 1, leaf_size = 64, max_level = 16;
  int m2l_precision = 0;
  int64_t n = 0, n_nodes = 0, n_leaves = 0;
  ed::HostTree host;

  // points in tree order: x, y, z relative to the root centre (FP32), w = q;
  // pts[n] is a zero-charge sentinel far away (the P2P's padding source)
  float4* pts = nullptr;
  float* area_t = nullptr;   // areas, tree order
  float* sig_t = nullptr;    // sigma, tree order (written per matvec)
  float* selfc_t = nullptr;  // per-area self term, tree order (per matvec)
  float* near_t = nullptr;   // near-field correction, tree order (per one-RHS matvec)
  int32_t* perm = nullptr;   // tree position -> original index

  // nodes
  float4* node_ctr = nullptr;   // centre (relative), w = 1/s (s = box side)
  int2* node_rng = nullptr;     // [begin, end)
  int32_t* node_parent = nullptr;
  int32_t* node_octant = nullptr;  // octant within parent
  int32_t* first_child = nullptr;
  int32_t* n_child = nullptr;
  float* M = nullptr;  // n_nodes * nc
  float* L = nullptr;  // n_nodes * nc
  float* ym2m = nullptr;   // n_nodes * nc: each node's own part of its local
                           // expansion (M2L sums, before the ancestors' part)

  // leaves (node ids) and their U lists (CSR over leaf ordinal)
  int32_t* pt_node = nullptr;   // n, leaf node of each tree-ordered panel
  // Multi-right-hand-side P2P work items (ed::P2PItem), their source
  // segments (per leaf: the leaf itself, then its other U-list leaves) and
  // per tree-ordered panel (first partial row, number of rows); the rows of
  // 64 partial sums per item and RHS sit in multi.part
  void* p2p_items = nullptr;
  int2* p2p_seg = nullptr;
  int2* p2p_ref = nullptr;
  int64_t n_p2p_items = 0;
  // One-right-hand-side P2P: every unordered pair of distinct U-list leaves
  // is evaluated once for both directions (ed::P2PSymItem: r_ij and its
  // rsqrt once, the target side summed per lane, the source side by a
  // cross-lane column sum).  The smaller leaf of a pair is the target side,
  // and the targets of all the smaller neighbours of a source leaf are packed
  // 64 to a chunk (a chunk is a few lane ranges, ed::P2PSymSeg, each the
  // panels of one target leaf).  A leaf with itself runs one-directional
  // items (ed::P2PItem over
  // p2s_seg).  Partial sums sit in rows per leaf: leaf L (panels b .. b +
  // n_L - 1) owns K_L rows of n_L slots from p2s_leaf[L].x, row k slot i - b
  // for panel i; every row is written whole, by the items of one (pair,
  // source piece), one chunk or one piece of the own stream, and k_l2p_out
  // adds a panel's rows in order.  (The multi-right-hand-side matvec keeps
  // the one-directional schedule above: with R charges per source the column
  // sums would cost more than the shared rsqrt saves.)
  void* p2s_items = nullptr;
  int2* p2s_seg = nullptr;
  void* p2s_sym = nullptr;
  void* p2s_tseg = nullptr;  // ed::P2PSymSeg: target segments of the chunks
  int64_t n_p2s_items = 0, n_p2s_sym = 0;
  float* p2s_part = nullptr;
  int2* p2s_leaf = nullptr;  // per node: (first slot, rows K); leaves only
  int64_t n_p2s_slots = 0, n_p2s_sym_interactions = 0;
  int32_t* leaves = nullptr;
  int32_t* u_ptr = nullptr;
  int32_t* u_src = nullptr;  // node ids
  int64_t n_u = 0, n_p2p_interactions = 0, n_p2p_src = 0;

  // upward/downward passes: per level l, nodes of level l that have children
  std::vector<int64_t> parents_begin;  // host offsets into parents_list per level
  int32_t* parents_list = nullptr;

  // M2L: CSR over target nodes (v_ptr over all nodes), source ids and table ids
  int32_t* v_ptr = nullptr;
  int32_t* v_src = nullptr;
  int32_t* v_tab = nullptr;
  int64_t n_v = 0;
  int n_off = 0;                 // number of distinct offsets (tables)
  std::vector<int32_t> slot_to_tab;

  // M2L in GEMM form: pairs grouped by offset table, tiles of kM2LTile pairs
  int32_t* m2l_pair_src = nullptr;   // n_v, grouped by table
  uint8_t* m2l_pair_sgn = nullptr;   // n_v: reflection of the pair's offset onto its table
  int32_t* m2l_pair_tgt = nullptr;
  int32_t* m2l_tile_tab = nullptr;   // per tile: table id
  int32_t* m2l_tile_beg = nullptr;   // per tile: first pair
  int32_t* m2l_tile_cnt = nullptr;   // per tile: pair count
  int64_t n_m2l_tiles = 0;
  int32_t* m2l_tgt_ptr = nullptr;    // CSR target -> rows of Ypair (n_nodes+1)
  int32_t* m2l_tgt_row = nullptr;    // Ypair row of each (target, pair)
  float* ypair = nullptr;            // n_v * ncp (ncp = nc padded)
  int ncp = 96;

  // operator tables (column-major: T[tab][j][k] = out k <- in j)
  float* t_m2l = nullptr;       // n_off * nc * nc, fp32 (VALU path)
  float* t_m2l_mfma = nullptr;  // n_off * ncp * kp, row-major padded, scaled (MFMA A)
  uint16_t* t_m2l_h = nullptr;  // n_off * ncp * kp16, bf16/fp16 (scaled)
  // split-FP16 tables of a 16-bit plan (the same tiles and table ids): the
  // solver's residual updates run the M2L in split FP16 through them
  const uint16_t* t_m2l_h3 = nullptr;
  // M2L form of the next matvecs when >= 0 (3: split FP16 with t_m2l_h3 on a
  // 16-bit plan); -1: the plan's own form
  int m2l_prec_override = -1;
  float* m2l_ascale = nullptr;  // per input coefficient: sqrt((n+m)!(n-m)!)
  float* m2l_binv = nullptr;    // per output coefficient: sqrt((n+m)!(n-m)!)
  const uint8_t* m2l_flip = nullptr;  // per coefficient: reflection sign bits (shared)
  unsigned* m2l_mmax = nullptr;    // 64 slots (128 B apart) of max |M * ascale| as float bits
  float* t_m2m = nullptr;       // 8 * nc * nc
  const float4* t_m2m7 = nullptr;  // octant-7 M2M operator alone, padded to float4s
  float* t_l2l = nullptr;       // 8 * nc * nc

  // near-field correction CSR in tree order
  int32_t* near_ptr = nullptr;  // n + 1
  int32_t* near_col = nullptr;
  float* near_w = nullptr;
  int64_t n_near = 0;


  // FP64 copies in tree order for the accurate (dense) residual operator
  double* cent64_t = nullptr;  // n x 3
  double* area64_t = nullptr;  // n
  double* near_w64 = nullptr;  // near CSR weights in FP64

  // coarse space of the solver's preconditioner (tree order, k columns)
  int coarse_k = 0;
  double* coarse_z = nullptr;     // k x n, orthonormal
  double* coarse_az = nullptr;    // k x n, A Z with the FMM operator
  double* coarse_einv = nullptr;  // k x k, (Z^T A Z)^-1
  double* coarse_cy = nullptr;
  double* coarse_part = nullptr;
  double* coarse_tmp = nullptr;   // n, coarse solve staging
  // symmetric dense FP64 operator of the refinement (built on first use)
  ed::DenseSymDevice* dense_sym = nullptr;
  // sparse approximate inverse smoother (column-major ELL, 32 slots per row)
  int32_t* sai_col = nullptr;
  double* sai_val = nullptr;
  int32_t* sai_cnt = nullptr;

  // stage events while ed_fmm_profile runs (8 markers), else NULL
  hipEvent_t* prof_events = nullptr;
  // P2P runs on a side stream beside the far-field passes (fork by an event
  // on the caller's stream, join by an event under graph capture, so that
  // capture records both branches, else by the signal below);
  // the stream is shared by all plans on the device and never destroyed
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // where the P2P is the longer branch (p2p_blocks_per_cu == 0) and outside
  // graph capture, the join is a stream memory operation on a signal instead
  // of an event: the side stream writes sync_join = k (the matvec's sequence
  // number) after the P2P and the caller's stream waits for >= k.  The write
  // and the wait cost about 5 us on the host against 8-15 us for an event
  // record plus cross-stream wait.  (The fork stays an event: a side stream
  // waiting on a signal value that the caller's stream writes later resumed
  // tens of microseconds late.)  Every write comes from the device's one side
  // stream, in order, so the value only grows.  NULL: events only.
  uint64_t* sync_join = nullptr;
  uint64_t sync_seq = 0;
  int n_cu = 256;  // compute units of the plan's device (caps the one-RHS P2P grid)
  // one-RHS P2P grid cap in blocks per CU (0: one block per item block); set
  // at plan creation from the estimated lengths of the two matvec branches
  int p2p_blocks_per_cu = 0;

  // GMRES workspace (allocated lazily by the solver, owned by the plan)
  void* gmres_ws = nullptr;

  // buffers of the multi-right-hand-side matvec (allocated on first use for
  // up to `cap` right-hand sides)
  struct Multi {
    int cap = 0;
    float* qT = nullptr;       // n x R charges, point-major (tree order)
    float* sig = nullptr;      // cap x n sigma (tree order)
    float* M = nullptr;        // cap x n_nodes x nc
    float* L = nullptr;        // cap x n_nodes x nc
    float* ym2m = nullptr;     // cap x n_nodes x nc
    float* ypair = nullptr;    // cap x pair_stride
    int64_t pair_stride = 0;   // n_v x ncp
    float* part = nullptr;     // cap x 64 x n_p2p_items
    unsigned* mmax = nullptr;  // cap x 64 x 32 multipole-maximum slots
  } multi;
};

namespace ed {
// P2P work item: up to 64 consecutive targets of one leaf (tree positions
// tgt0 .. tgt0 + ntc - 1) against `cnt` consecutive points of that leaf's
// source stream, starting at offset `off` of segment `seg`; the item writes
// its partial sums to p2p_part[out + t].  slices = min(64 / ntc, 8).
struct P2PItem {
  int32_t tgt0, ntc, seg, off, cnt, out, slices, pad;
};
constexpr int kP2PItemSteps = 256;
// Symmetric P2P item: the nl <= 64 targets of chunk segments seg0 ..
// seg0 + nseg - 1 against the sources s0 .. s0 + cnt - 1 (one leaf, piece q
// of it) in both directions; the sources' sums go to part[outS + j - s0].
struct P2PSymItem {
  int32_t seg0, nseg, q, nl, s0, cnt, outS, pad;
};
// Lanes lane0 .. (the next segment's lane0) - 1 of a chunk hold the panels
// panel0 + (l - lane0) of one target leaf, whose sums for source piece q go
// to part[slot0 + (l - lane0) + q * stride] (stride: the leaf's size).
struct P2PSymSeg {
  int32_t lane0, panel0, slot0, stride;
};
constexpr int kP2PSymSources = 256;
// Grid cap of the one-RHS P2P in blocks per compute unit, used when the
// far-field chain is the longer of the two branches of a matvec.
constexpr int kP2PBlocksPerCU = 5;  // sources per symmetric item
// Work item of the split dense operator: up to 512 target rows and one
// contiguous source range of their geometry.
struct DenseItem {
  int64_t row0;
  int32_t nrows;
  int32_t split;
  int64_t src0, src1;
};
void dense_schedule(const std::vector<int64_t>& off, std::vector<DenseItem>& items,
                    std::vector<int32_t>& row_nsplit, int* max_split);
// Symmetric FP64 dense operator (one right-hand side): tiles of 512 rows a x
// 512 sources b >= a of one geometry, one partial slot per tile for the rows
// and one for the columns, and per row tile the list of slots to add
// (slot * 2 + 1 for a column slot).
constexpr int kDenseSymTile = 512;
struct DenseSymItem {
  int64_t r0, s0;
  int32_t nr, ns, diag, pad;
};
struct DenseSymSchedule {
  std::vector<DenseSymItem> items;
  std::vector<int64_t> trow0;
  std::vector<int32_t> tnr, tptr, tent;
};
struct DenseSymDevice {
  DenseSymItem* items = nullptr;
  int64_t* trow0 = nullptr;
  int32_t *tnr = nullptr, *tptr = nullptr, *tent = nullptr;
  double* part = nullptr;  // 2 x n_items x kDenseSymTile: row slots, then column slots
  int64_t n_items = 0, n_tiles = 0;
};
void dense_sym_schedule(const std::vector<int64_t>& off, DenseSymSchedule& S);
int dense_sym_upload(const DenseSymSchedule& S, DenseSymDevice& D);
void dense_sym_free(DenseSymDevice& D);
int dense_sym_run(const DenseSymDevice& D, const double* C, const double* A, const double* sigma,
                  const double* selfi, double* out, hipStream_t st);
template <typename T>
int dense_run(const DenseItem* items, int64_t n_items, const int32_t* row_nsplit, int max_split,
              const T* C, const T* A, const T* sigma, const T* selfi, T* out, int64_t n,
              int64_t nrhs, T* partial, hipStream_t st);
int plan_upload(ed_plan* P);
int multi_reserve(ed_plan* P, int nrhs);
void multi_free(ed_plan* P);
void plan_free(ed_plan* P);
// FMM matvec with tree-order q/sigma/self already in place; writes the result
// either in original order (out_orig) or tree order (out_tree), FP32 or FP64.
int fmm_launch(ed_plan* P, const void* sigma, int sigma_is_f64, int sigma_tree_order,
               const void* self_integrals, void* out, int out_is_f64,
               int out_tree_order, hipStream_t stream, int nrhs = 1);
void gmres_free(ed_plan* P);
void gmres_invalidate_graph(ed_plan* P);
// z = M_sai u (+ Z cy when k > 0), tree order
void sai_apply_launch(ed_plan* P, const double* u, int k, const double* Z, const double* cy,
                      double* z, hipStream_t st);
void sai_free(ed_plan* P);
}  // namespace ed
