"""``bem_solve`` and ``BEMSolution`` (core/bem.py:1369-1542, :1550-2700): the
single-layer Dirichlet BEM with explicit point charges, end to end on the
device — mesh a parsed spec, build the operator, solve, measure the
boundary-condition residual, refine h until the target residual, a plateau or
the pass limit, and return the reference's result dict with a solution that
evaluates V and E anywhere.

What is the reference's and what is this build's:

- refinement policy, pass bookkeeping and result keys follow core/bem.py:
  h_{k+1} = refine_factor h_k from ``initial_h``; a pass records h, dof,
  bc_resid_linf = max |V_free + A sigma - bc| at the centroids and
  energy_A = 1/2 sum V sigma A; the best pass is the smallest residual (ties
  within max(1e-3 target, 1e-12) go to the finer mesh); stop at
  bc_resid_linf <= target_bc_inf_norm (with min_panels / min_refine_passes),
  or after plateau_max passes without a relative improvement of
  plateau_rel_improvement;
- the operator is the reference operator (dense single layer with the
  per-area self term; refined self integrals and near-pair corrections only
  when use_near_quadrature / use_near_quadrature_matvec ask for them, at the
  reference's depths: 3 with a sphere or two planes, else 4 (self) and 3
  (near));
- the solve is this build's: FP64 GMRES on the FMM operator (order / leaf from
  the autotuner, or the configured ones) with FP64 iterative refinement
  against the FP64 dense operator, to gmres_tol times ||b - A x0|| (the
  reference's tolerance convention, core/bem_solver.py:404-407), with the SAI
  smoother unless ``precond`` says otherwise;
- failure recovery follows core/bem.py:2170-2303: a preconditioned solve
  that fails (or has a non-finite residual) is retried once without the
  preconditioner, then handed to the damped Jacobi fall-back solve
  (core/bem.py:1156-1251: x += omega D^-1 (b - A x) from x = 0 until
  ||r|| <= tol max(||r_0||, 1)) when allow_jacobi_fallback is set; if that
  fails too the result is the reference's error dict.

Out of scope (reference control plane): run manifests, control.json polling,
the BEM intercept recorder, the near_alpha neighbour blending (off by
default), tile-size autotuning.  ``differentiable=True`` solves the
reference's differentiable operator (dense + self term, no near correction)
with ``solve_diffbem`` (adjoint gradients, core/diffbem.py) and returns a
solution whose sigma and evaluations carry gradients.
"""
from __future__ import annotations

import math
from typing import Any, Dict, List, Optional, Tuple

import numpy as np
import torch

from .. import _native as nat
from ..fmm3d.autotune import autotune_for_problem
from ..fmm3d.config import FmmConfig
from ..orchestration.parser import CanonicalSpec
from .bem_kernel import (_bem_E_field_targets_core_torch, _bem_matvec_core_torch,
                         _bem_potential_targets_core_torch, bem_E_field_targets,
                         bem_potential_targets)
from .bem_mesh import TriMesh, generate_mesh
from .bem_quadrature import near_quadrature_potentials
from .diffbem import solve_diffbem
from .kernels import K_E
from .operator import BemOperator


# -- point-charge fields (core/bem.py:291-371) ---------------------------------
def _point_charges(spec) -> List[Tuple[float, Tuple[float, float, float]]]:
    out = []
    for ch in getattr(spec, "charges", []) or []:
        if ch.get("type") == "point" and ch.get("pos") is not None:
            out.append((float(ch.get("q", 0.0)), tuple(float(v) for v in ch["pos"])))
    return out


def _free_space_at_points(spec, P: torch.Tensor, dtype, device) -> torch.Tensor:
    """V_free(x) = sum_k K_E q_k / max(|x - r_k|, 1e-12)."""
    V = torch.zeros(P.shape[0], device=device, dtype=dtype)
    for q, pos in _point_charges(spec):
        r = torch.linalg.norm(P - torch.tensor(pos, device=device, dtype=dtype)[None, :], dim=1)
        V = V + (K_E * q) / r.clamp_min(1e-12)
    return V


def _free_space_E_field_at_points(spec, P: torch.Tensor, dtype, device) -> torch.Tensor:
    """E_free(x) = sum_k K_E q_k (x - r_k) / max(|x - r_k|, 1e-12)^3."""
    E = torch.zeros(P.shape[0], 3, device=device, dtype=dtype)
    for q, pos in _point_charges(spec):
        R = P - torch.tensor(pos, device=device, dtype=dtype)[None, :]
        r = torch.linalg.norm(R, dim=1).clamp_min(1e-12)
        E = E + (K_E * q) * R / (r ** 3)[:, None]
    return E


def _bc_vector(spec, mesh: TriMesh, device, dtype) -> torch.Tensor:
    """Per-panel potential of the owning conductor (``potential``, default 0;
    conductors keyed by ``id`` or index; unknown ids 0) (core/bem.py:373-403)."""
    id_to_v = {}
    for i, c in enumerate(getattr(spec, "conductors", []) or []):
        try:
            v = float(c.get("potential", 0.0))
        except (TypeError, ValueError):
            v = 0.0
        id_to_v[int(c.get("id", i))] = v
    vals = np.array([id_to_v.get(int(cid), 0.0) for cid in mesh.conductor_ids], dtype=np.float64)
    return torch.as_tensor(vals, device=device, dtype=dtype)


def _offset_from_boundary(spec, P: torch.Tensor, areas: torch.Tensor, dtype, device,
                          sign: float = 1.0) -> torch.Tensor:
    """Nudge evaluation points lying on an ideal plane or sphere by 1e-9 times
    the median panel radius, along the normal (core/bem.py:406-472)."""
    if P.numel() == 0 or spec is None:
        return P
    P2 = P.clone()
    if areas.numel() > 0:
        req = float(torch.median(torch.sqrt(areas.clamp_min(1e-30) / math.pi)))
    else:
        req = 1e-3
    delta = sign * 1e-9 * req
    eps = torch.finfo(dtype).eps
    for c in spec.conductors:
        kind = c.get("type")
        if kind == "plane":
            z0 = float(c.get("z", 0.0))
            mask = torch.abs(P2[:, 2] - z0) < 10.0 * eps
            if bool(mask.any()):
                P2[mask, 2] = z0 + delta
        elif kind == "sphere":
            try:
                ctr = torch.tensor([float(x) for x in c.get("center", [0.0, 0.0, 0.0])],
                                   device=device, dtype=dtype)
                a = float(c.get("radius", 1.0))
            except (TypeError, ValueError):
                continue
            r = torch.linalg.norm(P2 - ctr[None, :], dim=1)
            mask = torch.abs(r - a) < 10.0 * eps * max(1.0, a)
            if bool(mask.any()):
                v = P2[mask] - ctr[None, :]
                v = v / torch.linalg.norm(v, dim=1, keepdim=True).clamp_min(1e-24)
                P2[mask] = ctr[None, :] + (a + delta) * v
    return P2


def compute_bem_capacitive_energy(V_total, sigma, areas) -> float:
    """U = 1/2 sum V_i sigma_i A_i in FP64 (core/bem.py:1254-1274)."""
    try:
        return float(0.5 * torch.sum(V_total.to(torch.float64) * sigma.to(torch.float64)
                                     * areas.to(torch.float64)))
    except RuntimeError:
        return float("nan")


class BEMSolution:
    """Evaluates the solved surface charge anywhere (core/bem.py:1369-1542):
    ``eval(p)`` -> potential, ``eval_V_E_batched(P)`` -> (V[N], E[N, 3]);
    total field = point charges + induced, with the reference's boundary
    nudge, and the near-singular target correction when ``near_quadrature``
    is on and panel vertices are given."""

    def __init__(self, spec: CanonicalSpec, centroids, areas, sigma, device, dtype,
                 tile_size: int, normals=None, differentiable: bool = False,
                 panel_vertices=None, near_quadrature: bool = False, near_quad_order: int = 2,
                 near_quad_dist_factor: float = 1.5, near_quad_max_depth: Optional[int] = None):
        self._spec = spec
        self._C = centroids
        self._A = areas
        self._N = normals
        self._S = sigma
        self._device = device
        self._dtype = dtype
        self._tile = tile_size
        self._differentiable = bool(differentiable)
        self._panel_vertices = panel_vertices
        self._near_quad_enabled = bool(near_quadrature and panel_vertices is not None
                                       and not self._differentiable)
        self._near_quad_order = int(near_quad_order)
        self._near_quad_dist_factor = float(near_quad_dist_factor)
        self._near_quad_max_depth = near_quad_max_depth
        self.meta: Dict[str, Any] = {}

    def _eval_at_points(self, P: torch.Tensor, compute_V: bool, compute_E: bool):
        P = torch.as_tensor(P).to(device=self._device, dtype=self._dtype).reshape(-1, 3)
        Ps = _offset_from_boundary(self._spec, P, self._A, self._dtype, self._device)
        V = E = None
        if compute_V:
            if self._differentiable:
                V_ind = _bem_potential_targets_core_torch(Ps, self._C, self._A, self._S)
            else:
                V_ind = bem_potential_targets(Ps, self._C, self._A, self._S)
            if self._near_quad_enabled and self._panel_vertices.numel() > 0:
                V_ind = near_quadrature_potentials(
                    V_ind, Ps, self._C, self._A, self._S, self._panel_vertices,
                    distance_factor=self._near_quad_dist_factor, quad_order=self._near_quad_order,
                    max_depth=self._near_quad_max_depth)
            V = _free_space_at_points(self._spec, Ps, self._dtype, self._device) + V_ind
        if compute_E:
            field = (_bem_E_field_targets_core_torch if self._differentiable
                     else bem_E_field_targets)
            E = (_free_space_E_field_at_points(self._spec, Ps, self._dtype, self._device)
                 + field(Ps, self._C, self._A, self._S))
        return V, E

    def eval(self, p) -> float:
        V, _ = self._eval_at_points(torch.tensor([[float(p[0]), float(p[1]), float(p[2])]]),
                                    True, False)
        return float(V[0]) if V is not None and V.numel() > 0 else 0.0

    def eval_V_E_batched(self, P):
        V, E = self._eval_at_points(P, True, True)
        return V, E


def _jacobi_fallback_solve(A, b, diag, logger, *, maxiter: int = 1024, tol: float = 1e-6,
                           omega: float = 0.8):
    """Damped Jacobi for A x = b (core/bem.py:1156-1251): x_0 = 0,
    x += omega D^-1 r, r = b - A x, until ||r|| <= tol max(||r_0||, 1);
    returns (x, info) with iters, resid, success, solver = "jacobi_fallback"."""
    def log(level, msg, **kw):
        if logger is not None:
            getattr(logger, level)(msg, **kw)

    x = torch.zeros_like(b)
    dinv = 1.0 / diag.clamp_min(1e-18)
    with torch.no_grad():
        r = b - A(x)
        rn = float(torch.linalg.norm(r))
        if not math.isfinite(rn):
            log("error", "Jacobi fallback: initial residual is non-finite.", r0=rn)
            return x, {"iters": 0, "resid": rn, "success": False, "solver": "jacobi_fallback"}
        target = tol * max(rn, 1.0)
        log("warning", "Jacobi fallback: starting iterative solve.", r0=rn, tol=tol,
            target=target, maxiter=maxiter)
        it = 0
        for k in range(maxiter):
            it = k + 1
            x = x + omega * dinv * r
            r = b - A(x)
            rn = float(torch.linalg.norm(r))
            if not math.isfinite(rn):
                log("error", "Jacobi fallback: residual became non-finite.", iter=it, resid=rn)
                return x, {"iters": it, "resid": rn, "success": False,
                           "solver": "jacobi_fallback"}
            if k == 0 or (k + 1) % 25 == 0:
                log("info", "Jacobi fallback iter.", iter=it, resid=rn)
            if rn <= target:
                log("info", "Jacobi fallback converged.", iters=it, resid=rn, target=target)
                return x, {"iters": it, "resid": rn, "success": True,
                           "solver": "jacobi_fallback"}
        log("warning", "Jacobi fallback reached maxiter without convergence.", iters=it,
            resid=rn, target=target)
        return x, {"iters": it, "resid": rn, "success": False, "solver": "jacobi_fallback"}


def _failed(info) -> Tuple[bool, float, int]:
    try:
        resid = float(info.get("resid", math.inf))
    except (TypeError, ValueError):
        resid = math.inf
    return (not bool(info.get("success", True))) or (not math.isfinite(resid)), resid, \
        int(info.get("code", 0) or 0)


def _solve_with_recovery(solve, jacobi, precond, allow_jacobi, logger):
    """The reference's recovery chain around one linear solve
    (core/bem.py:2170-2303).  ``solve(precond)`` -> (sigma, info);
    ``jacobi()`` -> (sigma, info).  Returns (sigma, info, None) or
    (None, None, (message, extra keys)) for the error result."""
    def log(level, msg, **kw):
        if logger is not None:
            getattr(logger, level)(msg, **kw)

    sigma, info = solve(precond)
    bad, resid, code = _failed(info)
    if precond != "none" and bad:
        log("warning", "GMRES reported failure or non-finite residual with preconditioner; "
            "retrying without preconditioner.", gmres_resid=resid, gmres_code=code)
        sigma, info = solve("none")
        info = dict(info)
        info["retried_without_precond"] = True
        bad, resid, code = _failed(info)
    if not bad:
        return sigma, info, None
    if not allow_jacobi:
        log("error", "BEM linear solve failed; GMRES reported failure and Jacobi fallback is "
            "disabled.", gmres_resid=resid, gmres_code=code)
        return None, None, (f"BEM linear solve failed: GMRES(code={code}, resid={resid})",
                            dict(gmres_resid=resid, gmres_code=code))
    log("warning", "GMRES reported failure; attempting Jacobi fallback.", gmres_resid=resid,
        gmres_code=code)
    sigma_j, info_j = jacobi()
    try:
        resid_j = float(info_j.get("resid", math.inf))
    except (TypeError, ValueError):
        resid_j = math.inf
    if bool(info_j.get("success", False)) and math.isfinite(resid_j):
        log("info", "Jacobi fallback succeeded.", resid=resid_j, iters=info_j.get("iters"))
        return sigma_j, dict(info_j), None
    log("error", "Jacobi fallback failed after GMRES failure.", gmres_resid=resid,
        gmres_code=code, jacobi_resid=resid_j, jacobi_iters=info_j.get("iters"))
    return None, None, (f"BEM linear solve failed: GMRES(code={code}, resid={resid}) "
                        f"and Jacobi(resid={resid_j})",
                        dict(gmres_resid=resid, gmres_code=code, jacobi_resid=resid_j,
                             jacobi_iters=info_j.get("iters")))


def _near_depths(spec) -> Tuple[int, int]:
    """(self / evaluation depth, near-matvec depth) (core/bem.py:1762-1774)."""
    kinds = [c.get("type") for c in spec.conductors]
    if any(k == "sphere" for k in kinds) or kinds.count("plane") >= 2:
        return 3, 3
    return 4, 3


def _operator_shape(cfg, n: int, mesh: TriMesh, logger) -> Tuple[int, int]:
    order = getattr(cfg, "fmm_order", None)
    leaf = getattr(cfg, "fmm_leaf_size", None)
    if order is not None and leaf is not None:
        return int(order), int(leaf)
    if n < int(getattr(cfg, "fmm_min_panels", 4096)):
        return int(order or 10), int(leaf or 64)
    base = FmmConfig(expansion_order=10, leaf_size=64, mac_theta=0.8,
                     m2l_precision=getattr(cfg, "m2l_precision", "fp16x3"))
    res = autotune_for_problem(base, n, dry_run=False, target_rel_error=1e-5,
                               centroids=mesh.centroids, areas=mesh.areas, use_cache=False)
    if logger:
        logger.info("FMM autotune.", order=res.config.native_order, leaf=res.config.leaf_size,
                    **{k: v for k, v in res.metrics.items() if k in ("ms", "rel_err")})
    return int(order or res.config.native_order), int(leaf or res.config.leaf_size)


def bem_solve(spec: CanonicalSpec, cfg, logger, differentiable: bool = False) -> Dict[str, Any]:
    """The reference's h-refined BEM solve; returns the result dict
    (boundary_samples, sample_points, solution, surface_charge_density,
    mesh_stats, gmres_stats, refinement_history) or {"error": ...}."""
    nat.require_gpu()
    device = torch.device("cuda", torch.cuda.current_device())
    dtype = torch.float64

    def log(level, msg, **kw):
        if logger is not None:
            getattr(logger, level)(msg, **kw)

    current_h = float(getattr(cfg, "initial_h", 0.3))
    refine_factor = float(getattr(cfg, "refine_factor", 0.5))
    target_bc = float(getattr(cfg, "target_bc_inf_norm", 1e-7))
    tiny_bc_band = max(target_bc * 1e-3, 1e-12)
    max_passes = int(getattr(cfg, "max_refine_passes", 3))
    restart = max(1, min(int(getattr(cfg, "gmres_restart", 50)), 128))
    gtol = float(getattr(cfg, "gmres_tol", 1e-8))
    maxiter = int(getattr(cfg, "gmres_maxiter", 500))
    use_precond = bool(getattr(cfg, "use_precond", True))
    precond = getattr(cfg, "precond", "sai") if use_precond else "none"
    plateau_max = getattr(cfg, "plateau_max", None)
    plateau_max = None if plateau_max is None else int(plateau_max)
    plateau_rel = float(getattr(cfg, "plateau_rel_improvement", 0.05))
    near_eval = bool(getattr(cfg, "use_near_quadrature", False))
    near_mv = bool(getattr(cfg, "use_near_quadrature_matvec", False))
    near_order = int(getattr(cfg, "near_quadrature_order", 2))
    near_f = float(getattr(cfg, "near_quadrature_distance_factor", 1.5))
    self_depth, near_depth = _near_depths(spec)
    charges = _point_charges(spec)

    history: List[Dict[str, Any]] = []
    best: Optional[Dict[str, Any]] = None
    best_bc = float("inf")
    plateau = 0
    stop_reason = "unset"

    def error_result(msg: str, **extra) -> Dict[str, Any]:
        payload: Dict[str, Any] = {"error": msg}
        if history:
            payload["refinement_history"] = history
        if best is not None:
            payload["best_bc_resid_linf"] = best.get("bc_resid_linf")
            payload["best_dof"] = best.get("dof")
        payload.update(extra)
        return payload

    log("info", "BEM solver start.", device=str(device), max_refine_passes=max_passes)
    for rp in range(max_passes):
        log("info", f"Refine pass {rp + 1}: target h={current_h:.4f}", refine_pass=rp + 1,
            target_h=current_h)
        try:
            mesh = generate_mesh(spec, target_h=current_h, logger=logger)
        except NotImplementedError as e:
            log("error", "Mesh generation failed.", error=str(e))
            break
        N = mesh.n_panels
        if N == 0:
            log("error", "Generated mesh has zero panels; aborting.")
            break
        order, leaf = _operator_shape(cfg, N, mesh, logger)
        op = BemOperator(mesh, order=order, ws=1, leaf_size=leaf,
                         m2l_precision=getattr(cfg, "m2l_precision", "fp16x3"),
                         refined_self=near_eval or near_mv, self_depth=self_depth,
                         near=near_mv, near_depth=near_depth, near_order=near_order,
                         distance_factor=near_f, device=device)
        bc = _bc_vector(spec, mesh, device, dtype)
        V_free = _free_space_at_points(spec, op.C, dtype, device)
        if not bool(torch.isfinite(bc).all()):
            return error_result("BEM BC vector contains NaN/Inf.")
        if not bool(torch.isfinite(V_free).all()):
            return error_result("Free-space potential on centroids contains NaN/Inf.")
        b = (bc - V_free).contiguous()
        x0 = (b / op.diag).contiguous()
        if not bool(torch.isfinite(x0).all()):
            x0 = torch.zeros_like(b)
        # the reference's tolerance is relative to ||b - A x0||; the native
        # solver's is relative to ||b||
        if differentiable:
            # the reference's differentiable operator: the dense matvec with
            # the self term, no near-field correction (core/bem.py:2015-2022)
            C_d, A_d, sc_d = op.C, op.A, op.self_corr

            def matvec_diff(sg, C_d=C_d, A_d=A_d, sc_d=sc_d):
                return _bem_matvec_core_torch(C_d, A_d, sg, self_integrals=sc_d)

            d_out = solve_diffbem(spec, cfg, logger, C=op.C, N=op.C, A=op.A, rhs=b,
                                  matvec=matvec_diff, x0=x0)
            sigma, info = d_out["sigma"], dict(d_out["stats"])
            mv_final = matvec_diff
        else:
            r0 = float(torch.linalg.norm(b - op.approx(x0)))
            bn = float(torch.linalg.norm(b))
            tol_b = gtol * (r0 / bn if bn > 0.0 else 1.0)
            sigma, info, err = _solve_with_recovery(
                lambda pc: op.solve(b, x0, tol=tol_b, restart=restart, maxiter=maxiter,
                                    precond=pc),
                lambda: _jacobi_fallback_solve(op.approx, b, op.diag, logger,
                                               maxiter=int(getattr(cfg, "jacobi_maxiter", 256)),
                                               tol=float(getattr(cfg, "jacobi_tol", 1e-5)),
                                               omega=float(getattr(cfg, "jacobi_omega", 0.8))),
                precond, bool(getattr(cfg, "allow_jacobi_fallback", True)), logger)
            if err is not None:
                return error_result(err[0], **err[1])
            mv_final = op.accurate
        info = dict(info)
        info["tol"] = gtol
        info["used_preconditioner"] = precond != "none" and not info.get(
            "retried_without_precond", False) and info.get("solver") != "jacobi_fallback"
        info["fmm_order"], info["fmm_leaf_size"] = order, leaf
        log("info", "BEM linear solve done.", iters=info.get("iters"), resid=info.get("resid"),
            success=info.get("success"))
        if not info.get("success", False) or not math.isfinite(float(info.get("resid", math.inf))):
            return error_result(f"BEM linear solve failed: GMRES(code={info.get('code')}, "
                                f"resid={info.get('resid')})", gmres_code=info.get("code"),
                                gmres_resid=info.get("resid"))
        if not bool(torch.isfinite(sigma).all()):
            return error_result("BEM sigma contains NaN/Inf after linear solve.")
        V_tot = torch.nan_to_num(V_free + mv_final(sigma), nan=0.0, posinf=0.0, neginf=0.0)
        bc_resid = float(torch.max(torch.abs(V_tot.detach() - bc)))
        energy = compute_bem_capacitive_energy(V_tot.detach(), sigma.detach(), op.A)
        log("info", "Pass results.", bc_resid_linf=bc_resid, energy_A=f"{energy:.6e}")
        payload = {"pass": rp + 1, "h": current_h, "dof": N, "bc_resid_linf": bc_resid,
                   "energy_A": energy, "gmres_info": info, "tile_size": 0,
                   "artifacts": (spec, op.C, op.A, torch.as_tensor(mesh.normals, device=device),
                                 sigma, V_tot, mesh)}
        history.append(payload)
        if bc_resid < best_bc:
            improvement = best_bc - bc_resid if math.isfinite(best_bc) else float("inf")
            best, best_bc = payload, bc_resid
            plateau = 0 if improvement > max(1e-16, plateau_rel * max(best_bc, 1e-16)) \
                else plateau + 1
        elif best is not None and abs(bc_resid - best_bc) <= tiny_bc_band \
                and N > int(best.get("dof", 0)):
            best, best_bc, plateau = payload, bc_resid, 0
        else:
            plateau += 1
        have_min = (N >= int(getattr(cfg, "min_panels", 0))
                    and rp + 1 >= int(getattr(cfg, "min_refine_passes", 1)))
        if bc_resid <= target_bc and have_min:
            stop_reason = "target_bc"
            break
        if plateau_max is not None and plateau >= plateau_max and have_min:
            stop_reason = "plateau"
            break
        current_h *= refine_factor
    if stop_reason == "unset":
        stop_reason = "max_passes_reached"
    if best is None:
        return error_result("BEM solve failed (no passes)")

    spec_b, C_b, A_b, N_b_rm, sigma_b, Vtot_b, mesh_b = best["artifacts"]
    n_b = best["dof"]
    eval_f = near_f
    kinds = {c.get("type") for c in spec.conductors}
    if "sphere" in kinds:
        eval_f = max(eval_f, 6.0)
    elif "plane" in kinds:
        eval_f = max(eval_f, 3.0)
    pv = torch.as_tensor(mesh_b.panel_vertices, device=device, dtype=dtype) if near_eval else None
    solution = BEMSolution(spec_b, C_b, A_b, sigma_b, device, dtype, 0, normals=N_b_rm,
                           differentiable=differentiable, panel_vertices=pv,
                           near_quadrature=near_eval and not differentiable,
                           near_quad_order=near_order, near_quad_dist_factor=eval_f,
                           near_quad_max_depth=self_depth)
    solution.meta["energy_A"] = best["energy_A"]
    solution.meta["stop_reason"] = stop_reason
    n_samples = min(1024, max(1, n_b))
    step = max(1, n_b // n_samples)
    idx = torch.arange(0, n_b, step, device=device)[:n_samples]
    mesh_stats: Dict[str, Any] = {"n_panels": n_b, "total_area": float(A_b.sum()),
                                  "bc_residual_linf": float(best["bc_resid_linf"]),
                                  "h_final": float(best["h"]), "tile_size_final": 0,
                                  "gpu_mem_peak_mb": float(torch.cuda.max_memory_allocated(device))
                                  / 2 ** 20}
    if "plane" in {c.get("type") for c in spec_b.conductors} and mesh_stats["total_area"] > 0:
        mesh_stats["patch_L"] = math.sqrt(mesh_stats["total_area"])
    return {"boundary_samples": Vtot_b[idx].detach().cpu().numpy().tolist(),
            "sample_points": C_b[idx].cpu().numpy().tolist(),
            "solution": solution,
            "surface_charge_density": sigma_b.detach().cpu().numpy(),
            "mesh_stats": mesh_stats,
            "gmres_stats": dict(best.get("gmres_info", {})),
            "refinement_history": history,
            "stop_reason": stop_reason}
