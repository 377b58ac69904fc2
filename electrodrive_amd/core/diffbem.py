"""Differentiable BEM solve (core/diffbem.py): sigma = A^-1 b with gradients
to the right-hand side and to every tensor the matvec closes over.

The reference wraps its torch matvec in a xitorch LinearOperator.  Here the
forward solve is the package's GMRES on the (native) matvec, and the gradient
comes from the implicit-function theorem with one adjoint solve per backward
pass: sigma solves r(sigma, theta, b) = b - A_theta sigma = 0, so

    dL/db = lambda,   dL/dtheta = -lambda^T d(A_theta sigma)/dtheta,
    A^T lambda = dL/dsigma.

``solve_diffbem`` returns sigma as ``sigma* + A^-1 (r - r.detach())`` in
autograd terms: the value is the converged sigma*, and the backward of the
custom node turns dL/dsigma into lambda (GMRES on v -> A^T v, the
vector-Jacobian product of the matvec, which the native kernels of
``bem_matvec_diff`` supply), after which autograd carries lambda through
r = b - A_theta sigma* to b and theta.  The adjoint right-hand side is scaled
to unit norm for the solve, so the absolute floor of the GMRES tolerance does
not cut it short.
"""
from __future__ import annotations

import math
from typing import Any, Callable, Dict, Optional

import torch

from .bem_solver import gmres_restart


def _adjoint_apply(matvec: Callable, v: torch.Tensor) -> torch.Tensor:
    """A^T v as the vector-Jacobian product of the (linear) matvec."""
    with torch.enable_grad():
        x = torch.zeros_like(v, requires_grad=True)
        y = matvec(x)
        (gx,) = torch.autograd.grad(y, x, grad_outputs=v, allow_unused=True)
    return torch.zeros_like(v) if gx is None else gx.detach()


class _ImplicitSolve(torch.autograd.Function):
    """Value: sigma*.  Backward: dL/dsigma -> lambda = A^-T dL/dsigma, handed
    to the residual r (whose graph reaches b and the matvec's tensors)."""

    @staticmethod
    def forward(ctx, r, sigma_star, matvec, opts):
        ctx.matvec, ctx.opts = matvec, opts
        return sigma_star.clone()

    @staticmethod
    def backward(ctx, g):
        g = g.detach().contiguous()
        gn = float(torch.linalg.norm(g))
        if gn == 0.0 or not math.isfinite(gn):
            return torch.zeros_like(g), None, None, None
        mv, o = ctx.matvec, ctx.opts
        lam, info = gmres_restart(lambda v: _adjoint_apply(mv, v), g / gn,
                                  restart=o["restart"], tol=o["tol"], maxiter=o["maxiter"])
        o["adjoint_info"] = info
        return lam * gn, None, None, None


def solve_diffbem(spec: Optional[Any], cfg: Any, logger: Optional[Any] = None, *,
                  C: torch.Tensor, N: torch.Tensor, A: torch.Tensor, rhs: torch.Tensor,
                  matvec: Callable[[torch.Tensor], torch.Tensor],
                  x0: Optional[torch.Tensor] = None) -> Dict[str, Any]:
    """Autograd-capable solve of matvec(sigma) = rhs (core/diffbem.py):
    returns {"sigma": Tensor [N], "stats": {"solver", "resid", "iters",
    "success", ...}}.  ``C``, ``N``, ``A`` are the panel tensors the matvec
    uses (gradients reach them through the matvec); ``cfg`` supplies
    gmres_tol, gmres_maxiter and gmres_restart."""
    del spec, N
    n = int(C.shape[0])
    b = rhs.reshape(-1)
    if b.shape[0] != n:
        raise RuntimeError(f"rhs has {b.shape[0]} entries for {n} panels")
    opts = {"tol": float(getattr(cfg, "gmres_tol", 1e-6)),
            "maxiter": int(getattr(cfg, "gmres_maxiter", 200)),
            "restart": max(1, min(int(getattr(cfg, "gmres_restart", 50)), 128))}
    with torch.no_grad():
        sigma_star, info = gmres_restart(lambda s: matvec(s).detach(), b.detach(),
                                         restart=opts["restart"], tol=opts["tol"],
                                         maxiter=opts["maxiter"],
                                         x0=None if x0 is None else x0.reshape(-1).detach())
    sigma_star = sigma_star.to(b.dtype)
    # residual with sigma* held fixed: its graph reaches rhs and the tensors
    # inside the matvec, not sigma*
    r = b - matvec(sigma_star.detach())
    sigma = _ImplicitSolve.apply(r, sigma_star.detach(), matvec, opts)
    stats = {"solver": "gmres_adjoint", "resid": float(info.get("resid", float("nan"))),
             "iters": int(info.get("iters", 0)), "success": bool(info.get("success", False)),
             "code": int(info.get("code", -1)), "tol_abs": info.get("tol_abs")}
    if logger is not None:
        logger.info("DiffBEM solve done.", **stats)
    return {"sigma": sigma, "stats": stats}
