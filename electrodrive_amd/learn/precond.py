"""Learned GMRES initial guess and coarse space for the BEM solve.

A 4-layer MLP (width 256, SiLU; native HIP kernels, see ``mlp.py``) maps 12
per-panel features — normalised centroid, normal, log relative area, RHS
potential b, b^3, the Jacobi guess b/diag and their Morton-leaf means
(``ed_panel_features``) — to the panel charge density, normalised by the
geometry's largest Jacobi guess.  It is used twice:

* initial guess x0 = scale * net(features);
* coarse space of the native solver's preconditioner (``ed_plan_set_coarse``):
  Z = [x0, smooth monomials of the normalised centroid up to ``degree``],
  M^-1 u = D^-1 (u - A Z c) + Z c, c = (Z^T A Z)^-1 Z^T u.

Plain Jacobi GMRES stays the fallback and the judge: the solve always reports
its true residual against the FP64 operator, and ``solve`` retries without the
learned parts when the preconditioned solve does not converge.
"""
import ctypes
import os

import numpy as np
import torch

from .. import _native as nat
from .mlp import NativeMLP

N_FEATURES = 12
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


def jacobi_scale(op, b):
    """Per-panel normaliser: max |b/diag| of the panel's geometry."""
    jac = (b / op.diag).abs()
    s = torch.empty_like(jac)
    for g in range(len(op.meshes)):
        lo, hi = int(op.geom_offsets[g]), int(op.geom_offsets[g + 1])
        s[lo:hi] = jac[lo:hi].max().clamp_min(1e-300)
    return s


def monomial_basis(op, degree):
    cols = []
    for g in range(len(op.meshes)):
        lo, hi = int(op.geom_offsets[g]), int(op.geom_offsets[g + 1])
        c = op.C[lo:hi]
        c = c - 0.5 * (c.max(0).values + c.min(0).values)
        c = c / c.abs().max().clamp_min(1e-300)
        block = []
        for a in range(degree + 1):
            for bb in range(degree + 1 - a):
                for d in range(degree + 1 - a - bb):
                    block.append(c[:, 0] ** a * c[:, 1] ** bb * c[:, 2] ** d)
        cols.append(torch.stack(block, 1))
    if len(cols) == 1:
        return cols[0]
    # block-diagonal basis: each geometry gets its own copy of the functions
    k = cols[0].shape[1]
    Z = torch.zeros((op.n, k * len(cols)), dtype=torch.float64, device=op.device)
    for g, blk in enumerate(cols):
        lo, hi = int(op.geom_offsets[g]), int(op.geom_offsets[g + 1])
        Z[lo:hi, g * k:(g + 1) * k] = blk
    return Z


class LearnedPreconditioner:
    def __init__(self, net=None, degree=2, weights=None):
        self.net = net if net is not None else NativeMLP((N_FEATURES, 256, 256, 256, 1))
        path = weights if weights is not None else (DEFAULT_WEIGHTS if net is None else None)
        if path is not None and os.path.exists(path):
            d = np.load(path, allow_pickle=False)
            self.net.load_state_dict({"dims": d["dims"].tolist(), "params": d["params"]})
        self.degree = degree

    def save(self, path):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        sd = self.net.state_dict()
        np.savez(path, dims=np.asarray(sd["dims"], dtype=np.int64), params=sd["params"])

    def predict(self, op, b):
        F = panel_features(op, b)
        y = self.net.forward(F).view(-1).to(torch.float64)
        return (y * jacobi_scale(op, b)).contiguous()

    def coarse_basis(self, op, b, x0=None):
        x0 = self.predict(op, b) if x0 is None else x0
        Z = [x0[:, None]]
        if self.degree >= 0:
            Z.append(monomial_basis(op, self.degree))
        Z = torch.cat(Z, 1)
        if Z.shape[1] > 32:
            Z = Z[:, :32]
        return Z.contiguous()

    def solve(self, op, b, tol=1e-8, use_coarse=True, use_x0=True, **kw):
        """Learned-preconditioned solve; falls back to plain Jacobi GMRES when
        it does not converge.  Returns (x, info) with info['learned'] set."""
        x0 = self.predict(op, b)
        if use_coarse:
            op.plan.set_coarse(self.coarse_basis(op, b, x0), op.self_corr)
        try:
            x, info = op.solve(b, x0=x0 if use_x0 else None, tol=tol, **kw)
        finally:
            if use_coarse:
                op.plan.set_coarse(None)
        info["learned"] = True
        if info["code"] != 0:
            x, info = op.solve(b, tol=tol, **kw)
            info["learned"] = False
        return x, info


_ = ctypes
