"""Learned coarse space and initial guess for the BEM solve.

A 4-layer MLP (width 256, SiLU; native HIP kernels, see ``mlp.py``) maps 12
per-panel features -- normalised centroid, normal, log relative area, RHS
potential b, b^3, the Jacobi guess b/diag and their Morton-leaf means
(``ed_panel_features``) -- to ``N_OUT`` charge-density channels.  The channels
of every RHS component (one per point charge, one for a constant boundary
potential), together with fixed smooth columns (monomials of the normalised
centroid up to ``degree`` and the Jacobi guess), form the coarse basis Z of
the native solver's preconditioner (``ed_plan_set_coarse``):

    M^-1 u = D^-1 (u - A Z c) + Z c,   c = (Z^T A Z)^-1 Z^T u,

and the initial guess is the coarse solution x0 = Z (Z^T A Z)^-1 Z^T b
(``ed_plan_coarse_solve``): the predicted panel charge density.

Training (``train.py``) minimises the projection loss |sigma - Z c*|^2 /
|sigma|^2 of solver-produced solutions onto span(Z) (``ed_proj_loss_f32``), so
the network learns channels that, with the fixed columns, span the solution.

Plain Jacobi GMRES stays the fallback and the judge: every solve reports its
true residual against the FP64 operator, and ``solve`` retries without the
learned parts when the preconditioned solve does not converge.
"""
import os

import numpy as np
import torch

from .. import _native as nat
from ..core.kernels import K_E
from .mlp import NativeMLP

N_FEATURES = 12
N_OUT = 8
MAX_COARSE = 32
DEFAULT_DIMS = (N_FEATURES, 256, 256, 256, N_OUT)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                               "precond_mlp.npz")


def panel_features(op, b):
    """(n, 12) float32 device features of a BemOperator and RHS b."""
    n = op.n
    if op.mesh is not None:
        normals = op.mesh.normals
    else:
        normals = np.concatenate([m.normals for m in op.meshes])
    Nt = torch.as_tensor(normals, dtype=torch.float64).to(op.device).contiguous()
    F = torch.empty((n, N_FEATURES), dtype=torch.float32, device=op.device)
    nat.check(nat.lib().ed_panel_features(op.plan.handle, nat.ptr(op.C), nat.ptr(Nt),
                                          nat.ptr(op.A), nat.ptr(b.contiguous()),
                                          nat.ptr(op.diag), nat.ptr(F),
                                          nat.stream_ptr(op.device)), "ed_panel_features")
    return F


def monomial_basis(op, degree):
    """(n, (d+1)(d+2)(d+3)/6) monomials of the centroid, normalised per
    geometry to [-1, 1]^3 (single-geometry operators)."""
    c = op.C
    c = c - 0.5 * (c.max(0).values + c.min(0).values)
    c = c / c.abs().max().clamp_min(1e-300)
    cols = []
    for a in range(degree + 1):
        for bb in range(degree + 1 - a):
            for d in range(degree + 1 - a - bb):
                cols.append(c[:, 0] ** a * c[:, 1] ** bb * c[:, 2] ** d)
    return torch.stack(cols, 1)


def fixed_columns(op, b, degree):
    """Fixed coarse columns: monomials up to ``degree`` and the Jacobi guess
    b/diag scaled to unit max."""
    jac = b / op.diag
    jac = jac / jac.abs().max().clamp_min(1e-300)
    return torch.cat([monomial_basis(op, degree), jac[:, None]], 1)


def point_charge_potential(op, q, pos):
    """V_k = K_E q / max(|c - r_k|, 1e-12) at the collocation points
    (core/bem.py:291-320)."""
    p = torch.tensor(pos, dtype=torch.float64, device=op.device)
    r = torch.linalg.norm(op.C - p[None, :], dim=1).clamp_min(1e-12)
    return (K_E * float(q)) / r


def rhs_components(op, charges=(), bc=0.0):
    """RHS b = bc - sum_k V_k split into its components [bc * 1, -V_1, ...]."""
    comps = []
    if bc != 0.0:
        comps.append(torch.full((op.n,), float(bc), dtype=torch.float64, device=op.device))
    for q, pos in charges:
        comps.append(-point_charge_potential(op, q, pos))
    return comps


class LearnedPreconditioner:
    def __init__(self, net=None, degree=2, weights=None, seed=0):
        if net is None:
            net = NativeMLP(DEFAULT_DIMS, seed=seed)
            path = DEFAULT_WEIGHTS if weights is None else weights
            if os.path.exists(path):
                d = np.load(path, allow_pickle=False)
                net = NativeMLP(tuple(int(v) for v in d["dims"]), seed=seed)
                net.load_state_dict({"dims": d["dims"].tolist(), "params": d["params"]})
                if "degree" in d:
                    degree = int(d["degree"])
            elif weights is not None:
                raise FileNotFoundError(path)
        self.net = net
        self.degree = degree

    @property
    def n_out(self):
        return self.net.dims[-1]

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        sd = self.net.state_dict()
        np.savez(path, dims=np.asarray(sd["dims"], dtype=np.int64), params=sd["params"],
                 degree=np.int64(self.degree), **{k: np.asarray(v) for k, v in meta.items()})

    def channels(self, op, b_comp):
        """(n, n_out) float64 network channels for one RHS component."""
        return self.net.forward(panel_features(op, b_comp)).to(torch.float64)

    def coarse_basis(self, op, b, components=None):
        """Z = [channels of each component, fixed columns], at most 32 columns
        (components beyond the budget are merged into one)."""
        if op.mesh is None:
            raise ValueError("the learned coarse space is built for single-geometry operators")
        comps = [b] if not components else list(components)
        fixed = fixed_columns(op, b, self.degree)
        budget = (MAX_COARSE - fixed.shape[1]) // self.n_out
        if budget < 1:
            raise ValueError("degree too high for the coarse budget")
        if len(comps) > budget:
            comps = comps[:budget - 1] + [sum(comps[budget - 1:])]
        Z = torch.cat([self.channels(op, c) for c in comps] + [fixed], 1)
        return Z.contiguous()

    def predict(self, op, b, components=None):
        """Predicted panel charge density: the coarse solution Z (Z^T A Z)^-1 Z^T b
        on the learned basis (leaves the coarse space set on the plan)."""
        op.plan.set_coarse(self.coarse_basis(op, b, components), op.self_corr)
        return op.plan.coarse_solve(b.contiguous())

    def solve(self, op, b=None, *, charges=(), bc=0.0, tol=1e-8, use_coarse=True, use_x0=True,
              **kw):
        """Learned-preconditioned solve; falls back to plain Jacobi GMRES when
        it does not converge.  Pass either ``b`` or the RHS description
        (``charges`` as (q, (x, y, z)) and ``bc``) -- the latter gives one
        learned component per charge.  Returns (x, info) with info['learned']."""
        comps = None
        if b is None:
            comps = rhs_components(op, charges, bc)
            b = op.rhs_point_charges(list(charges), bc=bc)
        try:
            x0 = self.predict(op, b, comps)
            if not use_coarse:
                op.plan.set_coarse(None)
            x, info = op.solve(b, x0=x0 if use_x0 else None, tol=tol, **kw)
        finally:
            op.plan.set_coarse(None)
        info["learned"] = True
        if info["code"] != 0:
            x, info = op.solve(b, tol=tol, **kw)
            info["learned"] = False
        return x, info


def baseline_solve(op, b, degree=2, tol=1e-8, **kw):
    """The same two-level method without the network: coarse space of the
    fixed columns only, x0 its coarse solution."""
    op.plan.set_coarse(fixed_columns(op, b, degree).contiguous(), op.self_corr)
    try:
        x0 = op.plan.coarse_solve(b.contiguous())
        x, info = op.solve(b, x0=x0, tol=tol, **kw)
    finally:
        op.plan.set_coarse(None)
    return x, info
