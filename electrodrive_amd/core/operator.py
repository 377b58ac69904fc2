"""Single-layer collocation operator of one panel geometry, assembled on the GPU.

``BemOperator`` performs the set-up of core/bem.py:1715-1960 natively:
self integrals (refined rule, core/bem.py:1777-1816, or the equal-area disk
when ``refined_self=False``), near pairs and near-correction weights
(core/bem.py:1838-1957), the FMM plan, and the right-hand side
b = bc - V_free (core/bem.py:2056).  Two operators share the same definition

    (A sigma)_i = sum_{j != i} K_E A_j sigma_j / r_ij + self_corr_i A_i sigma_i
                  + sum_j delta_ij sigma_j          (near correction, optional)

``approx`` evaluates it with the FP32 FMM (the GMRES operator); ``accurate``
with the FP64 dense kernel plus the FP64 near correction (the residual operator
of the outer iterative refinement).
"""
import math
import threading

import numpy as np
import torch

from .. import _native as nat
from ..fmm3d.plan import FmmPlan
from . import bem_quadrature as bq
from .bem_kernel import dense_matvec
from .kernels import K_E


# Inner (FMM-operator) solve depth per M2L precision.  FP32: 1e-5, so that the
# first correction reaches the FMM's own accuracy and two FP64 refinement
# residuals (dense, the dominant solve cost) suffice for 1e-8 instead of three.
# 16-bit forms: about 7x (BF16, matvec error 1.4e-3) and 3x (FP16, 1.6e-4)
# their operator error; at the 300k torus these minimise the time to 1e-8
# (BF16 2e-2 -> 1e-2: 6 -> 5 dense residuals; FP16 3e-3 -> 5e-4: 4 -> 3 in
# most solves), deeper asks stall the inner GMRES at the operator's rounding.
INNER_TOL = {"fp32": 1e-5, "fp16x3": 1e-5, "bf16": 1e-2, "fp16": 5e-4}

# target-panel pair count from which potential_at(method="auto") takes the
# FMM with a one-shot tree: at 1M targets x 82k panels (8e10 pairs) the FP64
# direct kernel takes 47 ms and the FMM 113 ms, most of it the host tree
FMM_TARGET_PAIRS = 1 << 38


class FmmOperator:
    """Callable A(sigma) backed by a native plan; exposes the native solver to
    ``gmres_restart``."""

    def __init__(self, plan: FmmPlan, self_corr):
        self.plan = plan
        self.self_corr = self_corr

    def __call__(self, sigma):
        s = sigma.to(self.self_corr.device, torch.float64).contiguous()
        return self.plan.matvec(s, self.self_corr).to(sigma.device, sigma.dtype)

    def _ed_native_solve(self, b, x0=None, diag=None, restart=50, tol=1e-8, maxiter=500):
        dev = self.self_corr.device
        bb = b.to(dev, torch.float64).contiguous()
        x0d = None if x0 is None else x0.to(dev, torch.float64).contiguous()
        dd = None if diag is None else diag.to(dev, torch.float64).contiguous()
        x, info = self.plan.solve(bb, x0d, dd, self.self_corr, restart=restart, maxiter=maxiter,
                                  tol=tol, precond=dd is not None, refine=False)
        out = dict(iters=info["iters"], resid=info["resid"], success=info["success"],
                   code=info["code"], resid_true0=info["resid0"], tol_abs=info["tol_abs"],
                   first_iter_below_tol=info["first_iter_below_tol"], cycles=info["cycles"])
        if out["success"]:
            out["converged_on"] = "true"
        return x, out


class BemOperator:
    """``mesh`` is one TriMesh or a list of TriMesh (a batch of independent
    geometries: one plan, block-diagonal operator, geometries never interact)."""

    def __init__(self, mesh, *, order=8, ws=1, leaf_size=64, max_level=16,
                 m2l_precision="fp16x3", refined_self=True, self_depth=3, near=True,
                 near_depth=3, near_order=2, distance_factor=1.5, device=None):
        nat.require_gpu()
        self.device = torch.device(device or "cuda")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        meshes = list(mesh) if isinstance(mesh, (list, tuple)) else [mesh]
        self.meshes = meshes
        self.mesh = meshes[0] if len(meshes) == 1 else None
        sizes = [m.n_panels for m in meshes]
        self.geom_offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.n = int(self.geom_offsets[-1])
        cents = np.concatenate([m.centroids for m in meshes])
        areas = np.concatenate([m.areas for m in meshes])
        voff = np.cumsum([0] + [m.vertices.shape[0] for m in meshes])
        verts = np.concatenate([m.vertices for m in meshes])
        tris = np.concatenate([m.triangles + voff[k] for k, m in enumerate(meshes)])
        dev = self.device
        # The plan (host tree, interaction lists, schedules) needs only the
        # host geometry: it is built on a worker thread (the C call releases
        # the GIL) while this thread runs the device quadrature below.
        plan_box = {}

        def make_plan():
            try:
                with torch.cuda.device(dev):
                    plan_box["plan"] = FmmPlan(
                        cents, areas, leaf_size=leaf_size, order=order, ws=ws,
                        max_level=max_level, m2l_precision=m2l_precision,
                        geom_offsets=None if len(meshes) == 1 else self.geom_offsets)
            except BaseException as exc:  # re-raised on the calling thread
                plan_box["error"] = exc

        worker = threading.Thread(target=make_plan, name="fmm-plan", daemon=True)
        worker.start()
        with torch.cuda.device(dev):
            self.C = torch.as_tensor(cents, dtype=torch.float64).to(dev).contiguous()
            self.A = torch.as_tensor(areas, dtype=torch.float64).to(dev).contiguous()
            # panel vertex coordinates (n, 3, 3), gathered on the device
            vd = torch.as_tensor(verts, dtype=torch.float64).to(dev)
            self.PV = vd[torch.as_tensor(tris, dtype=torch.int64).to(dev)].contiguous()
            if refined_self:
                self.I_self, self.self_corr = bq.self_integrals(self.C, self.PV, self.A,
                                                                depth=self_depth, order=near_order)
            else:
                vals = 2.0 * math.pi * K_E * torch.sqrt(self.A.clamp_min(0.0) / math.pi)
                self.I_self = vals
                self.self_corr = vals / self.A.clamp_min(torch.finfo(torch.float64).tiny)
            self.near_pairs = None
            self.near_w = None
            if near and self.n > 1:
                parts = []
                for g in range(len(meshes)):
                    b, e = int(self.geom_offsets[g]), int(self.geom_offsets[g + 1])
                    if e - b > 1:
                        parts.append(bq.build_near_pairs(self.C[b:e], self.A[b:e],
                                                         distance_factor) + b)
                self.near_pairs = (torch.cat(parts) if parts
                                   else torch.zeros((0, 2), dtype=torch.int64, device=dev))
                self.near_w = bq.near_correction_weights(self.C, self.A, self.PV, self.near_pairs,
                                                         quad_order=near_order, max_depth=near_depth)
            worker.join()
            if "error" in plan_box:
                raise plan_box["error"]
            self.plan = plan_box["plan"]
            if self.near_pairs is not None and self.near_pairs.shape[0] > 0:
                self.plan.attach_near_device(self.near_pairs, self.near_w)
        self.fmm = FmmOperator(self.plan, self.self_corr)
        # Jacobi diagonal of bem_solve: the raw self integral (core/bem.py:2059)
        self.diag = self.I_self.clamp_min(1e-18)

    # -- right-hand sides ------------------------------------------------------
    def rhs_point_charges(self, charges, bc=0.0, per_geometry=False):
        """b = bc - V_free with V_free = sum_k K_E q_k / max(|c - r_k|, 1e-12)
        (core/bem.py:291-320, :2056).  charges: iterable of (q, (x, y, z)).
        With ``per_geometry`` (batches), ``charges`` holds one charge list per
        geometry and each list acts on its own geometry only."""
        V = torch.zeros(self.n, dtype=torch.float64, device=self.device)
        if per_geometry and len(charges) != len(self.meshes):
            raise ValueError("per_geometry charges need one list per geometry")
        groups = charges if per_geometry else [charges] * len(self.meshes)
        for g, chs in enumerate(groups):
            b, e = int(self.geom_offsets[g]), int(self.geom_offsets[g + 1])
            C = self.C[b:e]
            for q, pos in chs:
                p = torch.tensor(pos, dtype=torch.float64, device=self.device)
                r = torch.linalg.norm(C - p[None, :], dim=1).clamp_min(1e-12)
                V[b:e] = V[b:e] + (K_E * float(q)) / r
        bcv = torch.as_tensor(bc, dtype=torch.float64, device=self.device)
        return (bcv - V).contiguous()

    # -- operators ---------------------------------------------------------------
    def approx(self, sigma):
        return self.plan.matvec(sigma.to(self.device, torch.float64).contiguous(), self.self_corr)

    def accurate(self, sigma):
        s = sigma.to(self.device, torch.float64).contiguous()
        v = torch.empty_like(s)
        for g in range(len(self.meshes)):
            b, e = int(self.geom_offsets[g]), int(self.geom_offsets[g + 1])
            v[b:e] = dense_matvec(s[b:e], self.C[b:e], self.A[b:e], self.self_corr[b:e])
        if self.near_pairs is not None and self.near_pairs.shape[0] > 0:
            v = bq.near_quadrature_matvec(v, s, self.C, self.A, None, self.near_pairs,
                                          weights=self.near_w)
        return v

    def enable_sai(self, factor_pattern=2.0, factor_rows=4.0):
        """Builds the sparse approximate inverse smoother (``precond="sai"``)."""
        self.plan.build_sai(self.self_corr, factor_pattern, factor_rows)
        self._sai = True

    def solve(self, b, x0=None, *, tol=1e-8, restart=50, maxiter=2000, refine=True,
              inner_tol=None, max_refine=None, precond=True, use_graph=False, check_every=4,
              residual_update=True):
        """FP64 GMRES on the FMM operator with FP64 iterative refinement against
        the accurate operator (``refine=True``); returns (sigma, info).
        ``precond``: True / "jacobi" (core/bem.py:2141-2158), "sai" (sparse
        approximate inverse, built on first use) or False / "none".
        ``b`` of shape (R, n), R <= 8, solves R right-hand sides together (one
        bem_solve per charge set in the reference); sigma has b's shape.
        ``residual_update``: once a dense residual is small enough, later
        refinement residuals are updated with the split-FP16 FMM product of
        each correction (r <- r - A_fmm d) instead of recomputed with the
        O(N^2) dense operator; the solver bounds the distance of the updated
        residual from the true one (info ``resid_bound``) and counts both
        kinds (``n_accurate``, ``n_updated``)."""
        if x0 is None:
            x0 = (b / self.diag).contiguous()  # core/bem.py:2060
        # The inner solves run on the FMM operator, whose rounding (16-bit M2L:
        # ~1e-3 bf16, ~2e-4 fp16 relative) is not linear in sigma: each inner
        # solve is asked for no more than that floor allows, and the FP64
        # refinement takes the residual the rest of the way.
        prec = self.plan.m2l_precision
        if inner_tol is None:
            inner_tol = INNER_TOL.get(prec, 1e-4)
        if max_refine is None:
            max_refine = 16 if prec in ("bf16", "fp16") else 8
        if precond == "sai" and not getattr(self, "_sai", False):
            self.enable_sai()
        use_diag = precond not in (False, None, "none")
        G = len(self.meshes)
        multi = b.dim() == 2
        kw = dict(restart=restart, maxiter=maxiter, precond=precond, refine=refine,
                  max_refine=max_refine, inner_tol=inner_tol, use_graph=use_graph,
                  check_every=check_every, residual_update=residual_update)
        if G == 1 and not multi:
            return self.plan.solve(b, x0, self.diag if use_diag else None, self.self_corr,
                                   tol=tol, **kw)
        # batches: the operator is block diagonal, so each (right-hand side,
        # geometry) block is scaled to a unit right-hand side.  Without a
        # coarse space the native solver runs one Krylov space per block and
        # stops each at tol |b_block|; a coarse space couples the geometries of
        # one right-hand side into one system, solved to tol / sqrt(G) so that
        # every block still meets tol relative to its own |b_g|
        B = b if multi else b[None]
        gid = torch.as_tensor(np.repeat(np.arange(G), np.diff(self.geom_offsets)),
                              device=B.device)
        off = [int(v) for v in self.geom_offsets]

        def block_norms(V):
            # one reduction per geometry (fixed order; an index_add_ would sum
            # with atomics and change the scaling in its last bits)
            return torch.stack([torch.linalg.norm(V[:, off[g]:off[g + 1]], dim=1)
                                for g in range(G)], 1)
        nb = block_norms(B)
        scale = torch.where(nb > 0.0, 1.0 / nb, torch.ones_like(nb))[:, gid]
        Bs, Xs = B * scale, (x0 if multi else x0[None]) * scale
        if not multi:
            Bs, Xs = Bs[0], Xs[0]
        x, info = self.plan.solve(Bs.contiguous(), Xs.contiguous(),
                                  self.diag if use_diag else None, self.self_corr,
                                  tol=tol if self.plan.coarse_k == 0 else tol / math.sqrt(G),
                                  **kw)
        X = (x if multi else x[None]) / scale
        if refine:
            nbh = nb.cpu().numpy()
            rn, bn, per_geometry = self.plan.block_residuals()
            if per_geometry:
                # the solver's own FP64 residual per block (scaled units):
                # relative residuals are scale-free, absolute ones are
                # rn_block * |b_block|
                rn, bn = rn.reshape(nbh.shape), bn.reshape(nbh.shape)
                rel = np.where(nbh > 0.0, rn / np.where(bn > 0.0, bn, 1.0), 0.0)
                res = np.where(nbh > 0.0, rn * nbh, rn)
            else:
                Rr = torch.stack([B[r] - self.accurate(X[r]) for r in range(B.shape[0])])
                res = block_norms(Rr).cpu().numpy()
                rel = np.where(nbh > 0.0, res / np.where(nbh > 0.0, nbh, 1.0), 0.0)
            info["max_geometry_rel_resid"] = float(rel.max())
            info["resid"] = float(np.sqrt((res ** 2).sum()))
            info["rel_resid"] = info["resid"] / max(float(np.sqrt((nbh ** 2).sum())), 1e-300)
            if multi:
                info["rhs_rel_resid"] = (np.sqrt((res ** 2).sum(1))
                                         / np.maximum(np.sqrt((nbh ** 2).sum(1)), 1e-300)).tolist()
        return (X if multi else X[0]), info

    # -- post-solve evaluation (core/bem.py:1441-1522) -------------------------
    def free_space(self, points, charges, field=False):
        """Potential (or field) of point charges at points: K_E q / max(r, 1e-12)
        (core/bem.py:291-320) or K_E q (x - r_k) / max(r, 1e-12)^3 (:340-371)."""
        P = torch.as_tensor(points, dtype=torch.float64).to(self.device).reshape(-1, 3)
        out = torch.zeros((P.shape[0], 3) if field else P.shape[0], dtype=torch.float64,
                          device=self.device)
        for q, pos in charges:
            p = torch.tensor(pos, dtype=torch.float64, device=self.device)
            R = P - p[None, :]
            r = torch.linalg.norm(R, dim=1).clamp_min(1e-12)
            out = out + ((K_E * float(q)) * R / (r ** 3)[:, None] if field
                         else (K_E * float(q)) / r)
        return out

    def potential_at(self, points, sigma, charges=(), near=True, near_depth=None,
                     distance_factor=1.5, quad_order=2, method="auto"):
        """V_total = V_free + V_induced at arbitrary points (core/bem.py:
        1441-1522), the induced part by the native target kernel (``method=
        "direct"``: FP64, O(M N)) or by the FMM (``"fmm"``: the panels and the
        targets in one lattice tree, targets as charge-free points, O(M + N),
        FP32 with the plan's expansion order), plus (``near``) the
        near-singular correction of core/bem.py:940-1047.  ``"auto"`` takes
        the FMM from M N >= 2^38 on; for many densities at the same targets
        ``fmm_targets(points)`` keeps the tree."""
        from .bem_kernel import bem_potential_targets
        P = torch.as_tensor(points, dtype=torch.float64).to(self.device).reshape(-1, 3)
        s = sigma.to(self.device, torch.float64)
        if method == "auto":
            method = "fmm" if P.shape[0] * self.n >= FMM_TARGET_PAIRS else "direct"
        if method == "fmm":
            V = self._fmm_potential_targets(P, s)
        elif method == "direct":
            V = bem_potential_targets(P, self.C, self.A, s)
        else:
            raise ValueError(f"method must be 'auto', 'direct' or 'fmm', got {method!r}")
        if near:
            V = bq.near_quadrature_potentials(V, P, self.C, self.A, s, self.PV,
                                              distance_factor=distance_factor,
                                              quad_order=quad_order, max_depth=near_depth)
        return V + self.free_space(P, charges)

    def fmm_targets(self, points):
        """An FMM evaluator of the induced potential at fixed target points
        for any number of densities (``FmmTargets.potential(sigma)``): the
        panels and the targets in one lattice tree, built once."""
        return FmmTargets(self, points)

    def _fmm_potential_targets(self, P, s):
        ev = FmmTargets(self, P)
        try:
            return ev.potential(s)
        finally:
            ev.close()

    def field_at(self, points, sigma, charges=()):
        """E_total = E_free + E_induced at arbitrary points (core/bem.py:1497-1520)."""
        from .bem_kernel import bem_E_field_targets
        P = torch.as_tensor(points, dtype=torch.float64).to(self.device).reshape(-1, 3)
        E = bem_E_field_targets(P, self.C, self.A, sigma.to(self.device, torch.float64))
        return E + self.free_space(P, charges, field=True)

    def total_charge(self, sigma):
        return float((sigma.to(torch.float64) * self.A).sum())


class FmmTargets:
    """sum_j K_E A_j sigma_j / |x - c_j| at fixed targets by FMM matvecs over
    the panels and the targets together: targets are zero-area (charge-free)
    points of one lattice tree, so they receive the far field by L2P from
    their leaves and the near field by P2P like panels while contributing
    nothing.  Built once per target set (host tree O((N + M) log(N + M))),
    then O(N + M) per density; FP32 with the operator's order and leaf size
    and split-FP16 M2L (relative error ~2e-6 against the FP64 direct sum)."""

    def __init__(self, op, points):
        P = torch.as_tensor(points, dtype=torch.float64).reshape(-1, 3)
        if not bool(torch.isfinite(P).all()):
            raise ValueError("target points must be finite")
        self.op, self.m, self.plan = op, int(P.shape[0]), None
        if self.m == 0:
            return
        pts = np.concatenate([op.C.cpu().numpy(), P.cpu().numpy()])
        area = np.concatenate([op.A.cpu().numpy(), np.zeros(self.m)])
        self.plan = FmmPlan(pts, area, leaf_size=op.plan.leaf_size, order=op.plan.order,
                            ws=op.plan.ws, m2l_precision="fp16x3")

    def potential(self, sigma):
        op = self.op
        if self.m == 0:
            return torch.zeros(0, dtype=torch.float64, device=op.device)
        if self.plan is None:
            raise RuntimeError("FmmTargets is closed")
        s = sigma.to(op.device, torch.float64).reshape(-1)
        if s.shape[0] != op.n:
            raise ValueError(f"sigma must have {op.n} entries, got {s.shape[0]}")
        sig = torch.cat([s, torch.zeros(self.m, dtype=torch.float64, device=op.device)])
        return self.plan.matvec(sig)[op.n:].clone()

    def close(self):
        if self.plan is not None:
            torch.cuda.current_stream(self.op.device).synchronize()  # before the plan goes
            self.plan.close()
            self.plan = None


_ = np
