"""Owning wrapper of the native ``ed_plan`` (tree, lists, operators on HBM)."""
import ctypes

import numpy as np
import torch

from .. import _native as nat
from .config import M2L_PRECISIONS

STAT_NAMES = ("n_points", "n_nodes", "n_leaves", "depth", "n_p2p_leaf_pairs",
              "n_p2p_interactions", "n_m2l_pairs", "n_m2l_offsets", "order", "ws",
              "n_near_pairs", "n_levels_with_m2l", "n_p2p_source_points",
              "n_p2p_symmetric_interactions", "n_p2p_items", "n_p2p_partials")


class FmmPlan:
    """Native FMM plan for one panel geometry (centroids, areas on the host).

    ``matvec`` evaluates K_E * sum_{j != i} sigma_j A_j / r_ij (FMM), plus the
    per-area self term and the attached near correction, on device tensors in
    original panel order.  No allocation happens inside ``matvec`` when ``out``
    is given.
    """

    def __init__(self, centroids, areas, *, leaf_size=64, order=8, ws=1, max_level=16,
                 m2l_precision="fp16x3", geom_offsets=None):
        nat.require_gpu()
        L = nat.lib()
        C = np.ascontiguousarray(np.asarray(centroids, dtype=np.float64).reshape(-1, 3))
        A = np.ascontiguousarray(np.asarray(areas, dtype=np.float64).reshape(-1))
        if C.shape[0] != A.shape[0]:
            raise ValueError("centroids and areas must have the same length")
        self.n = int(C.shape[0])
        opts = nat.PlanOpts()
        L.ed_plan_opts_default(ctypes.byref(opts))
        opts.leaf_size = int(leaf_size)
        opts.order = int(order)
        opts.ws = int(ws)
        opts.max_level = int(max_level)
        opts.m2l_precision = int(M2L_PRECISIONS[m2l_precision]
                                 if isinstance(m2l_precision, str) else m2l_precision)
        h = ctypes.c_void_p()
        if geom_offsets is None:
            nat.check(L.ed_plan_create(C.ctypes.data_as(ctypes.c_void_p),
                                       A.ctypes.data_as(ctypes.c_void_p), self.n,
                                       ctypes.byref(opts), ctypes.byref(h)), "ed_plan_create")
        else:
            go = np.ascontiguousarray(np.asarray(geom_offsets, dtype=np.int64))
            nat.check(L.ed_plan_create_batched(C.ctypes.data_as(ctypes.c_void_p),
                                               A.ctypes.data_as(ctypes.c_void_p), self.n,
                                               go.ctypes.data_as(ctypes.c_void_p), go.size - 1,
                                               ctypes.byref(opts), ctypes.byref(h)),
                      "ed_plan_create_batched")
        self._h = h
        self.order, self.ws, self.leaf_size = int(order), int(ws), int(leaf_size)
        self.coarse_k = 0
        self.m2l_precision = m2l_precision

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("plan has been destroyed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            nat.lib().ed_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        out = (ctypes.c_int64 * len(STAT_NAMES))()
        nat.check(nat.lib().ed_plan_stats(self.handle, out, len(STAT_NAMES)), "ed_plan_stats")
        return {k: int(v) for k, v in zip(STAT_NAMES, out)}

    def geometries(self):
        """Panel offsets of the geometries of this plan (length n_geoms + 1)."""
        g = ctypes.c_int64(0)
        nat.check(nat.lib().ed_plan_geometries(self.handle, ctypes.byref(g), None, 0),
                  "ed_plan_geometries")
        off = np.zeros(g.value + 1, dtype=np.int64)
        nat.check(nat.lib().ed_plan_geometries(self.handle, ctypes.byref(g),
                                               off.ctypes.data_as(ctypes.c_void_p), off.size),
                  "ed_plan_geometries")
        return off

    def permutation(self):
        perm = np.empty(self.n, dtype=np.int64)
        nat.check(nat.lib().ed_plan_permutation(self.handle, perm.ctypes.data_as(ctypes.c_void_p)),
                  "ed_plan_permutation")
        return perm

    def attach_near(self, pairs, weights):
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int64).reshape(-1, 2))
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(-1))
        if pairs.shape[0] != w.shape[0]:
            raise ValueError("pairs and weights length mismatch")
        nat.check(nat.lib().ed_plan_attach_near(self.handle, pairs.ctypes.data_as(ctypes.c_void_p),
                                                w.ctypes.data_as(ctypes.c_void_p), pairs.shape[0]),
                  "ed_plan_attach_near")

    def attach_near_device(self, pairs, weights):
        """Near CSR from device tensors: pairs (P, 2) int64 i-major, weights (P,)
        float64 (the output of ``build_near_pairs`` / ``near_correction_weights``)."""
        if (not pairs.is_cuda or pairs.dtype != torch.int64 or pairs.dim() != 2
                or pairs.shape[1] != 2):
            raise ValueError("pairs must be an int64 device tensor of shape (P, 2)")
        if not weights.is_cuda or weights.dtype != torch.float64 or weights.shape != (pairs.shape[0],):
            raise ValueError("weights must be a float64 device tensor of length P")
        pairs, weights = pairs.contiguous(), weights.contiguous()
        nat.check(nat.lib().ed_plan_attach_near_device(self.handle, nat.ptr(pairs), nat.ptr(weights),
                                                       pairs.shape[0], nat.stream_ptr(pairs.device)),
                  "ed_plan_attach_near_device")

    def matvec(self, sigma, self_integrals=None, out=None, stream=None):
        if not sigma.is_cuda or sigma.dim() != 1 or sigma.shape[0] != self.n:
            raise ValueError("sigma must be a 1-D device tensor of length n")
        if sigma.dtype not in (torch.float32, torch.float64):
            raise ValueError("sigma must be float32 or float64")
        sigma = sigma.contiguous()
        if self_integrals is not None:
            if (not self_integrals.is_cuda or self_integrals.dtype != sigma.dtype
                    or self_integrals.shape != sigma.shape):
                raise ValueError("self_integrals must match sigma (device, dtype, shape)")
            self_integrals = self_integrals.contiguous()
        if out is None:
            out = torch.empty_like(sigma)
        st = stream if stream is not None else nat.stream_ptr(sigma.device)
        fn = (nat.lib().ed_fmm_matvec_f64 if sigma.dtype == torch.float64
              else nat.lib().ed_fmm_matvec_f32)
        nat.check(fn(self.handle, nat.ptr(sigma), nat.ptr(self_integrals), nat.ptr(out), st),
                  "ed_fmm_matvec")
        return out

    def matvec_multi(self, sigma, self_integrals=None, out=None, stream=None):
        """One matvec for R = sigma.shape[0] (1..8) right-hand sides: sigma
        (R, n) FP64 device tensor; returns (R, n).  Each row equals
        ``matvec`` of that row alone."""
        if (not sigma.is_cuda or sigma.dim() != 2 or sigma.shape[1] != self.n
                or sigma.dtype != torch.float64):
            raise ValueError("sigma must be a (R, n) float64 device tensor")
        R = int(sigma.shape[0])
        if not 1 <= R <= 8:
            raise ValueError("1..8 right-hand sides per call")
        sigma = sigma.contiguous()
        if self_integrals is not None:
            if (not self_integrals.is_cuda or self_integrals.dtype != torch.float64
                    or self_integrals.shape != (self.n,)):
                raise ValueError("self_integrals must be a float64 device tensor of length n")
            self_integrals = self_integrals.contiguous()
        if out is None:
            out = torch.empty_like(sigma)
        st = stream if stream is not None else nat.stream_ptr(sigma.device)
        nat.check(nat.lib().ed_fmm_matvec_multi_f64(self.handle, nat.ptr(sigma),
                                                    nat.ptr(self_integrals), nat.ptr(out), R,
                                                    st), "ed_fmm_matvec_multi_f64")
        return out

    # stage slots of ed_fmm_profile ("" = a slot the launcher leaves empty)
    STAGES = ("gather", "upward", "m2l", "m2l_reduce", "downward", "", "p2p", "l2p_out")

    def profile(self, sigma, self_integrals=None, reps=10):
        """Average device milliseconds per matvec stage (HIP events): gather,
        upward (P2M + M2M), m2l (GEMM), m2l_reduce (per-node sums), downward
        (L2L chains), p2p (inline), l2p_out."""
        if not sigma.is_cuda or sigma.dtype != torch.float64:
            raise ValueError("profile takes a float64 device sigma")
        out = torch.empty_like(sigma)
        ms = (ctypes.c_double * len(self.STAGES))()
        nat.check(nat.lib().ed_fmm_profile(self.handle, nat.ptr(sigma), nat.ptr(self_integrals),
                                           nat.ptr(out), int(reps), ms, len(self.STAGES),
                                           nat.stream_ptr(sigma.device)), "ed_fmm_profile")
        return {k: v for k, v in zip(self.STAGES, list(ms)) if k}

    def profile_multi(self, sigma, self_integrals=None, reps=10):
        """``profile`` of the multi-RHS matvec: sigma (R, n) FP64, R in 1..8."""
        if (not sigma.is_cuda or sigma.dim() != 2 or sigma.shape[1] != self.n
                or sigma.dtype != torch.float64):
            raise ValueError("sigma must be a (R, n) float64 device tensor")
        R = int(sigma.shape[0])
        if not 1 <= R <= 8:
            raise ValueError("1..8 right-hand sides per call")
        sigma = sigma.contiguous()
        out = torch.empty_like(sigma)
        ms = (ctypes.c_double * len(self.STAGES))()
        nat.check(nat.lib().ed_fmm_profile_multi(self.handle, nat.ptr(sigma),
                                                 nat.ptr(self_integrals), nat.ptr(out), R,
                                                 int(reps), ms, len(self.STAGES),
                                                 nat.stream_ptr(sigma.device)),
                  "ed_fmm_profile_multi")
        return {k: v for k, v in zip(self.STAGES, list(ms)) if k}

    def set_coarse(self, Z, self_integrals=None):
        """Coarse space (n, k) float64 device tensor (k <= 32) for the
        solver's preconditioner; None removes it."""
        if Z is None:
            nat.check(nat.lib().ed_plan_set_coarse(self.handle, None, 0, None, None),
                      "ed_plan_set_coarse")
            self.coarse_k = 0
            return
        if not Z.is_cuda or Z.dtype != torch.float64 or Z.dim() != 2 or Z.shape[0] != self.n:
            raise ValueError("Z must be a float64 device tensor of shape (n, k)")
        Zc = Z.t().contiguous()  # column-major
        nat.check(nat.lib().ed_plan_set_coarse(self.handle, nat.ptr(Zc), int(Z.shape[1]),
                                               nat.ptr(self_integrals),
                                               nat.stream_ptr(Z.device)), "ed_plan_set_coarse")
        self.coarse_k = int(Z.shape[1])

    def build_sai(self, self_integrals, factor_pattern=2.0, factor_rows=4.0):
        """Sparse approximate inverse smoother (solver ``precond=2``)."""
        if not self_integrals.is_cuda or self_integrals.dtype != torch.float64:
            raise ValueError("self_integrals must be a float64 device tensor")
        nat.check(nat.lib().ed_plan_build_sai(self.handle, nat.ptr(self_integrals.contiguous()),
                                              float(factor_pattern), float(factor_rows),
                                              nat.stream_ptr(self_integrals.device)),
                  "ed_plan_build_sai")

    def sai_apply(self, u):
        """z = M_sai u (the smoother alone)."""
        if not u.is_cuda or u.dtype != torch.float64 or u.shape != (self.n,):
            raise ValueError("u must be a float64 device tensor of length n")
        u = u.contiguous()
        z = torch.empty_like(u)
        nat.check(nat.lib().ed_plan_sai_apply(self.handle, nat.ptr(u), nat.ptr(z),
                                              nat.stream_ptr(u.device)), "ed_plan_sai_apply")
        return z

    def coarse_solve(self, u):
        """Z (Z^T A Z)^-1 Z^T u on the coarse space set by ``set_coarse``."""
        if not u.is_cuda or u.dtype != torch.float64 or u.shape != (self.n,):
            raise ValueError("u must be a float64 device tensor of length n")
        u = u.contiguous()
        out = torch.empty_like(u)
        nat.check(nat.lib().ed_plan_coarse_solve(self.handle, nat.ptr(u), nat.ptr(out),
                                                 nat.stream_ptr(u.device)), "ed_plan_coarse_solve")
        return out

    def block_residuals(self):
        """Per-block residual and right-hand-side norms of the last ``solve``
        as the native solver measured them (the FP64 operator when it
        refined): (resid, bnorm, per_geometry) with blocks (right-hand side,
        geometry), right-hand side major, when per_geometry, else whole
        right-hand sides (a coarse space couples the geometries)."""
        L = nat.lib()
        nb, pg = ctypes.c_int32(0), ctypes.c_int32(0)
        nat.check(L.ed_gmres_block_resid(self.handle, None, None, 0, ctypes.byref(nb),
                                         ctypes.byref(pg)), "ed_gmres_block_resid")
        r = np.zeros(nb.value)
        bn = np.zeros(nb.value)
        nat.check(L.ed_gmres_block_resid(self.handle, r.ctypes.data, bn.ctypes.data, nb.value,
                                         ctypes.byref(nb), ctypes.byref(pg)),
                  "ed_gmres_block_resid")
        return r, bn, bool(pg.value)

    TRACE_PHASES = ("initial", "arnoldi", "cycle_end", "refine", "update", "approx")

    def trace(self):
        """Rows of the last native solve's per-poll trace (ed_gmres_trace) as
        dicts with the SOLVER_TRACE columns (core/bem_solver.py:19-28) plus
        refine_step: ts, cycle, iter, resid_true_l2 (NaN where not measured),
        resid_precond_l2 (the solver's estimate at an Arnoldi poll),
        arnoldi_dim and phase (initial / arnoldi / cycle_end / refine: a
        refinement residual with the dense FP64 operator / update: a residual
        updated with the FMM product of the last correction, resid_precond_l2
        then holding the bound on its distance from the true residual /
        approx: a refinement residual b - A_fmm x far above the FMM's
        error)."""
        L = nat.lib()
        nr = ctypes.c_int32(0)
        nat.check(L.ed_gmres_trace(self.handle, None, 0, ctypes.byref(nr)), "ed_gmres_trace")
        rows = np.zeros((max(nr.value, 1), 8))
        nat.check(L.ed_gmres_trace(self.handle, rows.ctypes.data_as(ctypes.c_void_p), nr.value,
                                   ctypes.byref(nr)), "ed_gmres_trace")
        out = []
        for r in rows[:nr.value]:
            out.append(dict(ts=float(r[0]), refine_step=int(r[1]), cycle=int(r[2]), iter=int(r[3]),
                            resid_true_l2=float(r[4]), resid_precond_l2=float(r[5]),
                            arnoldi_dim=int(r[6]), phase=self.TRACE_PHASES[int(r[7])]))
        return out

    def solve(self, b, x0=None, diag=None, self_integrals=None, *, restart=50, maxiter=2000,
              tol=1e-8, precond=True, refine=True, max_refine=8, inner_tol=1e-4,
              use_graph=False, check_every=4, residual_update=True):
        """Native FP64 GMRES (+ iterative refinement); returns (x, info dict).
        ``b`` of shape (n,) or (R, n) with R <= 8: the R right-hand sides are
        solved together (one multi-RHS FMM matvec per Arnoldi step), each
        (right-hand side, geometry) block to tol relative to its own |b|."""
        def f64(t, rows=False):
            if t is None:
                return None
            ok = t.is_cuda and t.dtype == torch.float64 and (
                t.shape == (self.n,) or (rows and t.dim() == 2 and t.shape[1] == self.n
                                         and 1 <= t.shape[0] <= 8))
            if not ok:
                raise ValueError("solver vectors must be float64 device tensors of length n "
                                 "(right-hand sides: (n,) or (R, n), R <= 8)")
            return t.contiguous()
        b = f64(b, rows=True)
        x = torch.zeros_like(b) if x0 is None else f64(x0, rows=True).clone()
        if x.shape != b.shape:
            raise ValueError("x0 must have the shape of b")
        diag, self_integrals = f64(diag), f64(self_integrals)
        L = nat.lib()
        o = nat.GmresOpts()
        L.ed_gmres_opts_default(ctypes.byref(o))
        o.restart, o.maxiter, o.tol = int(restart), int(maxiter), float(tol)
        pc = {"none": 0, "jacobi": 1, "sai": 2}.get(precond, None) if isinstance(precond, str) \
            else int(bool(precond))
        if pc is None:
            raise ValueError(f"unknown preconditioner {precond!r} (none, jacobi, sai)")
        o.precond, o.refine, o.max_refine = pc, int(bool(refine)), int(max_refine)
        o.inner_tol, o.use_graph, o.check_every = float(inner_tol), int(bool(use_graph)), int(check_every)
        o.residual_update = int(bool(residual_update))
        info = nat.GmresInfo()
        if b.dim() == 2:
            nat.check(L.ed_gmres_solve_multi(self.handle, nat.ptr(b), nat.ptr(x), int(b.shape[0]),
                                             nat.ptr(diag), nat.ptr(self_integrals),
                                             ctypes.byref(o), ctypes.byref(info),
                                             nat.stream_ptr(b.device)), "ed_gmres_solve_multi")
        else:
            nat.check(L.ed_gmres_solve(self.handle, nat.ptr(b), nat.ptr(x), nat.ptr(diag),
                                       nat.ptr(self_integrals), ctypes.byref(o),
                                       ctypes.byref(info), nat.stream_ptr(b.device)),
                      "ed_gmres_solve")
        d = {k: getattr(info, k) for k, _ in nat.GmresInfo._fields_}
        d["success"] = d["code"] == 0
        return x, d
