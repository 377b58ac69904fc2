/*
 * ednative.h — C ABI of the MI355X (gfx950) BEM/FMM solve hot path.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory on the
 * current device; `ed_stream` is a hipStream_t (NULL = the default stream).
 * Every function returns ED_OK (0) or an ED_ERR_* status; ed_last_error()
 * gives a thread-local message for the last failure.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the Electrodrive repository root).
 */
#ifndef EDNATIVE_H
#define EDNATIVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ED_OK 0
#define ED_ERR_ARG 1        /* shape / value misuse  -> Python ValueError   */
#define ED_ERR_DEVICE 2     /* no device / HIP error -> Python RuntimeError */
#define ED_ERR_ALLOC 3      /* device allocation failed                      */
#define ED_ERR_LAUNCH 4     /* kernel launch failed                          */
#define ED_ERR_NONFINITE 5  /* non-finite values were produced / clamped     */
#define ED_ERR_INTERNAL 6

typedef void* ed_stream;
typedef struct ed_plan ed_plan;

const char* ed_last_error(void);
int32_t ed_abi_version(void);

/* ------------------------------------------------------------------------ */
/* FMM plan: replaces LaplaceFmm3D.__init__ / make_laplace_fmm_backend       */
/* (fmm3d/bem_fmm.py:127-320, :888-940): builds the octree, interaction      */
/* lists (fmm3d/tree.py:968-1050, fmm3d/interaction_lists.py:531-577) and    */
/* translation operators once per geometry.  Geometry is host FP64, as the   */
/* reference requires CPU geometry (bem_fmm.py:169-181).                     */
/* ------------------------------------------------------------------------ */
typedef struct ed_plan_opts {
  int32_t leaf_size;     /* max panels per leaf (reference default 64)        */
  int32_t order;         /* expansion order p: 4, 6, ..., 16 (reference
                            default 8); 14 and 16 run the VALU M2L          */
  int32_t ws;            /* well-separation radius in boxes: 1 or 2          */
  int32_t max_level;     /* deepest octree level, <= 20                      */
  int32_t m2l_precision; /* 0 = fp32 MFMA, 1 = bf16 MFMA, 2 = fp16 MFMA,
                            3 = split fp16 (hi + lo, three fp16 MFMAs),
                            -1 = VALU                                       */
  int32_t reserved[3];
} ed_plan_opts;

void ed_plan_opts_default(ed_plan_opts* opts);

int ed_plan_create(const double* centroids_host, const double* areas_host,
                   int64_t n, const ed_plan_opts* opts, ed_plan** out);
int ed_plan_destroy(ed_plan* plan);

/* Batched plan over several independent geometries (no reference
 * counterpart: the reference batches right-hand sides only,
 * core/bem_kernel.py:715-810).  Panels of geometry g are
 * [geom_offsets[g], geom_offsets[g+1]) of the input arrays.  Each geometry
 * gets its own octree and lists; the trees are merged level-major so each FMM
 * stage is one launch over all geometries.  Matvecs and solves are
 * block-diagonal: geometries never interact. */
int ed_plan_create_batched(const double* centroids_host, const double* areas_host,
                           int64_t n, const int64_t* geom_offsets, int64_t n_geoms,
                           const ed_plan_opts* opts, ed_plan** out);
/* Number of geometries and (if offsets_host non-NULL) their panel offsets. */
int ed_plan_geometries(const ed_plan* plan, int64_t* n_geoms_out, int64_t* offsets_host,
                       int64_t capacity);

/* Counters (TreeStatistics / InteractionStats analogue, fmm3d/tree.py:120-162,
 * fmm3d/interaction_lists.py:59-88).  out[i] for i < n_out:
 * 0 n_points 1 n_nodes 2 n_leaves 3 depth 4 n_p2p_leaf_pairs
 * 5 n_p2p_interactions 6 n_m2l_pairs 7 n_m2l_offsets 8 order 9 ws
 * 10 n_near_pairs 11 n_levels_with_m2l 12 n_p2p_source_points (sum over P2P
 * leaf pairs of the source leaf size) 13 n_p2p_symmetric_interactions
 * (ordered interactions the one-RHS P2P evaluates in symmetric pairs)
 * 14 n_p2p_items (one-RHS P2P work items) 15 n_p2p_partials (one-RHS P2P
 * partial sums per matvec) */
int ed_plan_stats(const ed_plan* plan, int64_t* out, int32_t n_out);

/* Tree-order permutation (index_map, fmm3d/tree.py:428-485): perm[i] is the
 * original index of tree position i.  Host output of length n. */
int ed_plan_permutation(const ed_plan* plan, int64_t* perm_host);

/* FMM matvec: replaces LaplaceFmm3D.matvec (fmm3d/bem_fmm.py:760-885):
 * out_i = K_E * sum_{j != i} sigma_j A_j / |c_i - c_j|  (FMM-approximated)
 *         + self_integrals_i * sigma_i * A_i            (if non-NULL)
 *         + sum_j delta_ij sigma_j                      (if near attached).
 * sigma / self_integrals / out are device arrays in original panel order;
 * non-finite outputs are clamped to 0 (bem_fmm.py:741-749).  The f64 variant
 * takes and returns FP64 with FP32 arithmetic inside.  No allocation.
 * The P2P runs on a per-device side stream that forks from `stream` and
 * joins it again before the output; the plan's buffers are shared by its
 * calls, so calls on one plan from several streams must be ordered by the
 * caller (as for one stream). */
int ed_fmm_matvec_f32(ed_plan* plan, const float* sigma, const float* self_integrals,
                      float* out, ed_stream stream);
int ed_fmm_matvec_f64(ed_plan* plan, const double* sigma, const double* self_integrals,
                      double* out, ed_stream stream);
/* The same matvec for nrhs (1..8) right-hand sides at once: sigma and out are
 * (nrhs, n) row-major FP64 device arrays in original panel order,
 * self_integrals (n) is shared.  Extends LaplaceFmm3D.matvec
 * (fmm3d/bem_fmm.py:760-885), which takes one vector: the P2P evaluates each
 * r_ij once for all right-hand sides and every far-field launch covers all of
 * them.  Each right-hand side's result equals ed_fmm_matvec_f64 on it alone
 * bit for bit.  Buffers for the extra right-hand sides are allocated on the
 * first call with a larger nrhs and kept with the plan. */
int ed_fmm_matvec_multi_f64(ed_plan* plan, const double* sigma, const double* self_integrals,
                            double* out, int32_t nrhs, ed_stream stream);

/* Per-stage device time of the FMM matvec (no reference counterpart: the
 * reference logs MultipoleOpStats counters, fmm3d/multipole_operators.py:
 * 984-1041, not times).  Runs the f64 matvec `reps` times with HIP events
 * between stages on `stream` (P2P inline, not on the side stream);
 * stage_ms[0..7] = gather, upward (P2M + M2M), M2L, M2L sums, downward
 * (L2L), 0, P2P, output (L2P + P2P partials + near + self), averaged; every
 * stage time includes the cost of its event pair (a few microseconds). */
int ed_fmm_profile(ed_plan* plan, const double* sigma, const double* self_integrals,
                   double* out, int32_t reps, double* stage_ms, int32_t n_stages,
                   ed_stream stream);
/* The same stage times for the multi-RHS matvec (ed_fmm_matvec_multi_f64) of
 * nrhs (1..8) right-hand sides: sigma and out (nrhs, n) row-major FP64. */
int ed_fmm_profile_multi(ed_plan* plan, const double* sigma, const double* self_integrals,
                         double* out, int32_t nrhs, int32_t reps, double* stage_ms,
                         int32_t n_stages, ed_stream stream);

/* Attach the near-singular correction (core/bem.py:1838-1957): pairs (P,2)
 * int64 (i, j) and FP64 weights delta_ij, host arrays in original order.  They
 * are converted to a tree-ordered CSR applied inside ed_fmm_matvec_*.  P = 0
 * detaches. */
int ed_plan_attach_near(ed_plan* plan, const int64_t* pairs_host,
                        const double* weights_host, int64_t n_pairs);
/* Same CSR from device arrays (no host round trip; no reference counterpart):
 * pairs (n_pairs, 2) int64 i-major with valid indices (what ed_near_pairs_f64
 * produces), weights (n_pairs) FP64, both device memory in original order.
 * Unsorted rows or out-of-range indices -> ED_ERR_ARG and an empty
 * correction. */
int ed_plan_attach_near_device(ed_plan* plan, const int64_t* pairs, const double* weights,
                               int64_t n_pairs, ed_stream stream);

/* ------------------------------------------------------------------------ */
/* Dense single-layer matvec: replaces _bem_matvec_core_torch and            */
/* bem_matvec_batched (core/bem_kernel.py:157-230, :715-810).                */
/* centroids (n,3), areas (n), sigma (nrhs,n), out (nrhs,n) device arrays;   */
/* self_integrals may be NULL (diagonal 0).                                  */
/* ------------------------------------------------------------------------ */
int ed_dense_matvec_f64(const double* centroids, const double* areas,
                        const double* sigma, const double* self_integrals,
                        double* out, int64_t n, int64_t nrhs, ed_stream stream);
int ed_dense_matvec_f32(const float* centroids, const float* areas,
                        const float* sigma, const float* self_integrals,
                        float* out, int64_t n, int64_t nrhs, ed_stream stream);
/* No reference counterpart: ed_dense_matvec_f64 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same inputs and outputs (one geometry,
 * nrhs >= 1), the split form's kernels; it allocates nothing, copies nothing
 * and does not synchronise the stream.  Its partial sums use a fixed 16 MiB
 * device array that each GPU the library runs on holds until the process
 * ends: n is limited to 131072 panels (split x n <= 2^21 doubles; larger n ->
 * ED_ERR_ARG), right-hand sides run in passes that fit.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller. */
int ed_dense_matvec_f64_noalloc(const double* centroids, const double* areas,
                                const double* sigma, const double* self_integrals,
                                double* out, int64_t n, int64_t nrhs, ed_stream stream);

/* ------------------------------------------------------------------------ */
/* Quadrature                                                                */
/* ------------------------------------------------------------------------ */
/* Induced potential V (m) and field E (m, 3) at arbitrary targets (m, 3) from
 * panel charges: V_i = sum_j K_E sigma_j A_j rsqrt(max(r^2, 1e-24)),
 * E_i = sum_j K_E sigma_j A_j (x_i - c_j) rsqrt(max(r^2, 1e-24))^3, non-finite
 * results -> 0 (bem_potential_targets / bem_E_field_targets,
 * core/bem_kernel.py:233-335, :818-1008).  V or E may be NULL (not
 * computed).  Device arrays, original order. */
int ed_targets_f64(const double* targets, int64_t m, const double* centroids,
                   const double* areas, const double* sigma, int64_t n, double* V, double* E,
                   ed_stream stream);
int ed_targets_f32(const float* targets, int64_t m, const float* centroids, const float* areas,
                   const float* sigma, int64_t n, float* V, float* E, ed_stream stream);
/* Near-singular correction of target potentials (core/bem.py:940-1047): for
 * panels with |c_j - x_i| <= factor R_j, A_j > 0, sigma_j != 0,
 * V_i += sigma_j (I_near - K_E A_j / max(|c_j - x_i|, 1e-12)), I_near the
 * panel integral refined `depth` (0..6) times toward x_i with the
 * order-`order` rule.  V (m) is updated in place; device FP64. */
int ed_near_targets_f64(const double* targets, int64_t m, const double* centroids,
                        const double* areas, const double* panel_vertices, const double* sigma,
                        int64_t n, double factor, int32_t depth, int32_t order, double* V,
                        ed_stream stream);

/* Self integrals I_i = sum_k K_E w_k / max(|c_i - y_k|, 1e-12) over the panel
 * refined `depth` times toward its centroid with the order-`order` rule
 * (core/bem.py:1777-1816, core/bem_quadrature.py:217-308).  panel_vertices
 * (n,3,3), centroids (n,3), out (n) device FP64. */
int ed_self_integrals_f64(const double* panel_vertices, const double* centroids,
                          int64_t n, int32_t depth, int32_t order, double* out,
                          ed_stream stream);

/* Near pairs (core/bem.py:475-555): (i, j), i != j,
 * 0 < |c_i - c_j| <= factor (R_i + R_j), R = sqrt(A/pi), i-major, j ascending.
 * *n_pairs_out (a host int64) receives the pair count; the pairs are written
 * to pairs_out when it is non-NULL and capacity >= that count, else nothing
 * is written (call again with a large enough buffer).  Device inputs, device
 * outputs. */
int ed_near_pairs_f64(const double* centroids, const double* areas, int64_t n,
                      double factor, int64_t* pairs_out, int64_t capacity,
                      int64_t* n_pairs_out, ed_stream stream);

/* Near weights (core/bem.py:558-667): delta_ij = sum_k K_E w_k/max(|c_i-y_k|,
 * 1e-12) - K_E A_j / max(|c_i - c_j|, 1e-12), panel j uniformly refined
 * `depth` times.  Device arrays. */
int ed_near_weights_f64(const double* centroids, const double* areas,
                        const double* panel_vertices, const int64_t* pairs,
                        int64_t n_pairs, int32_t order, int32_t depth,
                        double* weights_out, ed_stream stream);

/* Near correction apply: replaces the CUDA extension's near_quadrature_matvec
 * (core/bem_near_cuda.cpp:61-233, core/bem_near_cuda_kernel.cu:110-232) and
 * _apply_near_quadrature_matvec_torch (core/bem.py:675-734):
 * out = v_far + sum_j sigma_j * delta_ij.  When weights is NULL, delta_ij is
 * computed on the fly with the unrefined order-`quad_order` rule, as the CUDA
 * kernel does.  Deterministic (pairs must be i-major; no atomics). */
int ed_near_apply_f64(const double* v_far, const double* sigma,
                      const double* centroids, const double* areas,
                      const double* panel_vertices, const int64_t* pairs,
                      const double* weights, int64_t n_pairs, int64_t n,
                      int32_t quad_order, double* out, ed_stream stream);

/* ------------------------------------------------------------------------ */
/* Morton keys: replaces morton_keys_from_points (fmm3d/geometry.py:736-799).*/
/* points (n,3) device; bbox_center/bbox_half (3) host or NULL (tight bbox). */
/* keys (n) device int64.  Bit-exact with the reference.                      */
/* ------------------------------------------------------------------------ */
int ed_morton_keys_f64(const double* points, int64_t n, int32_t level,
                       const double* bbox_center, const double* bbox_half,
                       int64_t* keys, ed_stream stream);
int ed_morton_keys_f32(const float* points, int64_t n, int32_t level,
                       const float* bbox_center, const float* bbox_half,
                       int64_t* keys, ed_stream stream);

/* ------------------------------------------------------------------------ */
/* GMRES in FP64 on the device: replaces gmres_restart                       */
/* (core/bem_solver.py:125-920) when A is the plan's operator.               */
/* ------------------------------------------------------------------------ */
typedef struct ed_gmres_opts {
  int32_t restart;        /* Krylov dimension per cycle                      */
  int32_t maxiter;        /* total Arnoldi steps                             */
  double tol;             /* relative tolerance                              */
  int32_t precond;        /* 0 none, 1 Jacobi on diag (1/diag, bad -> 1),    */
                          /* 2 sparse approximate inverse (ed_plan_build_sai) */
  int32_t refine;         /* 1: outer FP64 iterative refinement with the     */
                          /*    dense FP64 operator as the residual operator */
  int32_t max_refine;     /* outer refinement steps                          */
  double inner_tol;       /* relative tolerance of each inner solve          */
  int32_t use_graph;      /* capture one Arnoldi step in a hipGraph (default */
                          /*    0: direct launches measured 3-8 % faster)  */
  int32_t check_every;    /* host poll interval in Arnoldi steps             */
  int32_t residual_update; /* 1: refinement residuals b - A_fmm x with the  */
                          /*    split-FP16 FMM while they are far above its  */
                          /*    error, one dense FP64 residual once small    */
                          /*    enough that the FMM product of the next      */
                          /*    correction is accurate to well below tol,    */
                          /*    then updates r <- r - A_fmm d (error bound   */
                          /*    tracked); an initial guess with |r| >= |b|   */
                          /*    is replaced by 0.  0: every refinement       */
                          /*    residual after the first is a dense one      */
} ed_gmres_opts;

typedef struct ed_gmres_info {
  int32_t code;           /* 0 converged, 1 maxiter, 2 stagnation, 3/5/6 non-finite */
  int32_t iters;          /* Arnoldi steps (matvecs) in total                */
  int32_t cycles;
  int32_t refine_steps;
  int32_t first_iter_below_tol; /* first step whose estimate met tol, or -1 */
  double resid;           /* final true residual norm ||b - A x||_2          */
  double resid0;          /* initial true residual norm                      */
  double rel_resid;       /* resid / ||b||_2                                 */
  double tol_abs;
  int32_t n_accurate;     /* residuals taken with the dense FP64 operator    */
  int32_t n_updated;      /* residuals updated with the FMM product          */
  double resid_bound;     /* bound on |updated - true residual| (0 if none)  */
  double update_err;      /* measured relative error of the update operator  */
} ed_gmres_info;

void ed_gmres_opts_default(ed_gmres_opts* opts);

/* b (n), x (n, in: initial guess, out: solution), diag (n, Jacobi diagonal,
 * may be NULL), self_integrals (n, per-area diagonal of the operator) — all
 * device FP64 in original order.  On a batched plan without a coarse space
 * every geometry gets its own Krylov space (same launches) and converges at
 * ||r_g|| <= tol ||b_g||; `resid` / `rel_resid` are over the whole vector. */
int ed_gmres_solve(ed_plan* plan, const double* b, double* x, const double* diag,
                   const double* self_integrals, const ed_gmres_opts* opts,
                   ed_gmres_info* info, ed_stream stream);

/* ed_gmres_solve for nrhs <= 8 right-hand sides at once: b, x are nrhs x n
 * (row r = right-hand side r, original order, device FP64); diag and
 * self_integrals as above, shared.  Replaces one bem_solve call per
 * right-hand side (core/bem.py:1369-2700 solves one charge set per call).
 * Every (right-hand side, geometry) block runs its own Krylov space and stops
 * at ||r|| <= tol ||b|| of that block; all blocks share each launch (one
 * multi-right-hand-side FMM matvec per Arnoldi step, as
 * ed_fmm_matvec_multi_f64).  `info` covers the whole solve: iters counts the
 * shared Arnoldi steps, resid / rel_resid are over all nrhs x n entries. */
int ed_gmres_solve_multi(ed_plan* plan, const double* b, double* x, int32_t nrhs,
                         const double* diag, const double* self_integrals,
                         const ed_gmres_opts* opts, ed_gmres_info* info, ed_stream stream);

/* Per-block norms of the last ed_gmres_solve / ed_gmres_solve_multi on this
 * plan (no reference counterpart; the reference reports one residual per
 * solve, core/bem_solver.py:703-730): resid[k] = ||b_k - A x_k|| as the solver
 * last measured it (the accurate FP64 operator when it refined, tree-order
 * sums), bnorm[k] = ||b_k||, host arrays of `cap` entries.  Blocks are
 * (right-hand side, geometry) pairs, right-hand side major
 * (*per_geometry = 1), or whole right-hand sides when a coarse space couples
 * the geometries (*per_geometry = 0); *n_blocks is their number. */
int ed_gmres_block_resid(ed_plan* plan, double* resid, double* bnorm, int32_t cap,
                         int32_t* n_blocks, int32_t* per_geometry);

/* Per-poll trace of the last ed_gmres_solve / ed_gmres_solve_multi on this
 * plan, the native counterpart of the SOLVER_TRACE rows gmres_restart writes
 * (core/bem_solver.py:19-28): 8 doubles a row = host time (s since the
 * epoch), outer refinement step, restart cycle, Arnoldi steps so far, true
 * residual norm (NaN if not measured at that point), preconditioned residual
 * norm (at an Arnoldi poll: the solver's estimate), Krylov dimension, phase
 * (0 initial residual of an inner solve, 1 Arnoldi poll, 2 cycle end,
 * 3 refinement residual with the accurate operator, 4 residual update
 * r <- r - A_fmm d: residual column = the updated residual, preconditioned
 * column = the bound on its distance from the true residual, 5 refinement
 * residual b - A_fmm x taken while it is far above the FMM's error).  *n_rows receives the
 * row count; min(cap, *n_rows) rows are copied when rows is non-NULL. */
int ed_gmres_trace(const ed_plan* plan, double* rows, int32_t cap, int32_t* n_rows);

/* Sparse approximate inverse smoother for ed_gmres_solve (precond = 2; no
 * reference counterpart, whose setup_preconditioner offers Jacobi and mass
 * scaling, core/bem_solver.py:36-117).  Row i of M minimises
 * |m^T A[J(i), I(i)] - e_i^T| with J(i) the panels within
 * factor_pattern (R_i + R_j) of panel i (at most 32, nearest first) and I(i)
 * those within factor_rows (R_i + R_j) (at most 128); A is the plan's
 * operator (near correction included when attached) with self_integrals
 * (n, device FP64, original order) on the diagonal.  Rebuild after
 * ed_plan_attach_near. */
int ed_plan_build_sai(ed_plan* plan, const double* self_integrals, double factor_pattern,
                      double factor_rows, ed_stream stream);
/* z = M u with the built smoother alone (original order, device FP64). */
int ed_plan_sai_apply(ed_plan* plan, const double* u, double* z, ed_stream stream);

/* Coarse space of ed_gmres_solve's preconditioner (no reference counterpart):
 * M^-1 u = S (u - A Z c) + Z c, c = (Z^T A Z)^-1 Z^T u (S the Jacobi or the
 * sparse approximate inverse smoother), with Z (n x k,
 * column-major, original order, device FP64) orthonormalised here and A the
 * plan's FMM operator with `self_integrals`.  k <= 32; k = 0 removes it.
 * Exact on span(Z), Jacobi elsewhere.  Used with the learned basis. */
int ed_plan_set_coarse(ed_plan* plan, const double* Z, int32_t k,
                       const double* self_integrals, ed_stream stream);

/* Coarse solve on the plan's coarse space (no reference counterpart):
 * out = Z (Z^T A Z)^-1 Z^T u, original order, device FP64; the initial guess
 * of the learned-preconditioned solve (u = b).  Needs ed_plan_set_coarse. */
int ed_plan_coarse_solve(ed_plan* plan, const double* u, double* out, ed_stream stream);

/* FP64 level-1 kernels for GMRES over an arbitrary operator callable: they
 * replace the torch.dot / norm / axpy steps of gmres_restart's Arnoldi loop
 * (core/bem_solver.py:536-541, :771-773).  Deterministic reductions.
 * ed_vec_dot_f64 needs >= 241 doubles of device scratch and synchronises. */
int ed_vec_dot_f64(const double* x, const double* y, int64_t n, double* scratch,
                   double* result_host, ed_stream stream);
int ed_vec_axpy_f64(double a, const double* x, double* y, int64_t n, ed_stream stream);
int ed_vec_scale_f64(double a, const double* x, double* y, int64_t n, ed_stream stream);
/* y = d .* x (Jacobi apply, core/bem_solver.py:80-83) */
int ed_vec_mul_f64(const double* d, const double* x, double* y, int64_t n, ed_stream stream);

/* ------------------------------------------------------------------------ */
/* Learned preconditioner network (no reference counterpart: the reference  */
/* has no learned preconditioner).  FP32, row-major, device pointers.        */
/* ------------------------------------------------------------------------ */
/* C = act(op(A) op(B) + bias) (+ beta * C); op = transpose when trans_*;
 * act 0 = identity, 1 = SiLU; pre_act (same layout as C) receives the
 * pre-activation when non-NULL; bias (N) may be NULL. */
int ed_gemm_f32(int32_t trans_a, int32_t trans_b, int64_t M, int64_t N, int64_t K,
                const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                int64_t ldc, const float* bias, int32_t act, float* pre_act, float beta,
                ed_stream stream);
int ed_act_backward_f32(const float* pre_act, const float* dy, float* dx, int64_t n,
                        int32_t act, ed_stream stream);
/* out[j] = sum_i X[i][j], fixed order; scratch >= 64 * cols floats */
int ed_colsum_f32(const float* X, int64_t rows, int64_t cols, float* scratch, float* out,
                  ed_stream stream);
int ed_adam_f32(float* param, const float* grad, float* m, float* v, int64_t n, float lr,
                float beta1, float beta2, float eps, int32_t step, float weight_decay,
                ed_stream stream);
/* grad = 2 w (y - t); *loss_host = sum w (y - t)^2 (synchronises); part >= 240 */
int ed_wmse_f32(const float* y, const float* t, const float* w, int64_t n, float* grad,
                float* part, float* loss_host, ed_stream stream);
/* 12 per-panel features (geometry, RHS potential, Jacobi guess, Morton-leaf
 * means) in original panel order; inputs device FP64 (n,3)/(n). */
int ed_panel_features(ed_plan* plan, const double* centroids, const double* normals,
                      const double* areas, const double* b, const double* diag, float* F,
                      ed_stream stream);
int32_t ed_panel_feature_count(void);
/* Projection loss for training a learned coarse basis (no reference
 * counterpart): per solve s (rows s*m .. s*m+m-1), Z = [Y | X],
 * c = argmin |t - Zc|^2 + ridge tr(Z^T Z)/k |c|^2, loss[s] = |t - Zc|^2/|t|^2,
 * dY = -2 (t - Zc) c_Y^T / |t|^2.  ky >= 1, ky + kx <= 32; device FP32. */
int ed_proj_loss_f32(const float* Y, int32_t ky, const float* X, int32_t kx, const float* T,
                     int64_t n_sol, int32_t m, float ridge, float* dY, float* loss,
                     ed_stream stream);

/* ------------------------------------------------------------------------ */
/* Reference-mode octree and interaction lists (host; bit-exact indexing).   */
/* The FMM runs on the lattice octree of ed_plan_create; these reproduce the */
/* reference's own adaptive tree and pair lists for callers of its tree and  */
/* list API.  Points are host arrays, FP64 (is_f32 = 0) or FP32.             */
/* ------------------------------------------------------------------------ */
typedef struct ed_ref_tree ed_ref_tree;
typedef struct ed_ref_lists ed_ref_lists;

/* Replaces build_fmm_tree / _build_adaptive_octree (fmm3d/tree.py:801-1050):
 * tight per-node boxes, split at the box centre with a strict '>', octant
 * x + 2y + 4z, stable partition, node ids in depth-first pre-order, leaf iff
 * n <= leaf_size, level >= max_depth or max half extent < 1e-7.  points (n,3)
 * must be finite; leaf_size > 0; 1 <= max_depth <= 64 (tree.py:1018-1031). */
int ed_ref_tree_create(const void* points, int32_t is_f32, int64_t n, int32_t leaf_size,
                       int32_t max_depth, ed_ref_tree** out);
/* out[0..3] = n_points, n_nodes, n_leaves, deepest level (-1 when empty). */
int ed_ref_tree_sizes(const ed_ref_tree* tree, int64_t* out);
/* The FmmTree SoA (fmm3d/tree.py:286-339); any pointer may be NULL:
 * index_map (n) tree position -> original index; centers, half_extents
 * (n_nodes,3) and radii (n_nodes) in the points' dtype; levels (n_nodes)
 * int32; parents (n_nodes, -1 for the root); children (n_nodes,8), -1 for a
 * missing octant; ranges (n_nodes,2) [start, end) in tree order; is_leaf
 * (n_nodes) bytes. */
int ed_ref_tree_export(const ed_ref_tree* tree, int64_t* index_map, void* centers,
                       void* half_extents, void* radii, int32_t* levels, int64_t* parents,
                       int64_t* children, int64_t* ranges, uint8_t* is_leaf);
int ed_ref_tree_destroy(ed_ref_tree* tree);

/* Replaces build_interaction_lists / _build_interaction_lists
 * (fmm3d/interaction_lists.py:374-577): dual-tree traversal with the MAC
 * max(2 r_s, 2 r_t) / dist < theta (:273-297) and the refinement rule of
 * :306-337.  target NULL (or == source) is the one-tree case (a node with
 * itself is never admissible).  0 < theta < 1.  max_pairs < 0: no cap;
 * otherwise the traversal stops pushing once that many pairs were emitted and
 * reports truncated = 1 (the reference then raises).  sort_pairs != 0 sorts
 * pairs by (tgt, src), else by (src, tgt). */
int ed_ref_lists_create(const ed_ref_tree* source, const ed_ref_tree* target, double theta,
                        int64_t max_pairs, int32_t sort_pairs, ed_ref_lists** out);
/* out[0..7] = n_p2p_pairs, n_m2l_pairs, mac_accept, mac_reject, leaf_leaf,
 * stack_pops, max_stack_size, truncated (InteractionStats,
 * interaction_lists.py:59-88). */
int ed_ref_lists_sizes(const ed_ref_lists* lists, int64_t* out);
/* Pairs as (src, tgt) rows: p2p (n_p2p,2), m2l (n_m2l,2); NULL skips. */
int ed_ref_lists_export(const ed_ref_lists* lists, int64_t* p2p_pairs, int64_t* m2l_pairs);
int ed_ref_lists_destroy(ed_ref_lists* lists);

/* ------------------------------------------------------------------------ */
/* Host-only helpers (no device needed): expose the plan builder's octree,   */
/* lists and operator tables so they can be checked on a CPU-only machine.   */
/* ------------------------------------------------------------------------ */
/* Builds the lattice tree on the host.  Returns counts in out_counts[0..3] =
 * n_nodes, n_leaves, n_v_pairs, n_u_pairs; when the array pointers are
 * non-NULL also fills: perm (n), node_level (n_nodes), node_begin, node_end,
 * node_parent (n_nodes), v_pairs (n_v_pairs,2) as (target, source),
 * u_pairs (n_u_pairs,2) as (target leaf, source leaf). */
int ed_host_tree_build(const double* centroids, int64_t n, int32_t leaf_size,
                       int32_t max_level, int32_t ws, int64_t* out_counts,
                       int64_t* perm, int32_t* node_level, int64_t* node_begin,
                       int64_t* node_end, int64_t* node_parent,
                       int64_t* v_pairs, int64_t* u_pairs);

/* Operator tables in FP64: kind 0 = M2L for integer offset (ox,oy,oz)
 * (unscaled, caller multiplies by 1/s), kind 1 = M2M child octant, kind 2 =
 * L2L child octant (octant = x | y<<1 | z<<2).  out is (p+1)^2 x (p+1)^2
 * row-major. */
int ed_host_operator(int32_t order, int32_t kind, int32_t ox, int32_t oy,
                     int32_t oz, int32_t octant, double* out);

/* Morton keys on the host (same arithmetic as the device kernel). */
int ed_host_morton_keys_f64(const double* points, int64_t n, int32_t level,
                            const double* bbox_center, const double* bbox_half,
                            int64_t* keys);

#ifdef __cplusplus
}
#endif

#endif /* EDNATIVE_H */
