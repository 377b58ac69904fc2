"""Restarted left-preconditioned GMRES (mirror of core/bem_solver.py).

``gmres_restart`` keeps the reference signature, info dict and codes
(core/bem_solver.py:125-216).  Two execution paths:

* ``A`` is an :class:`~electrodrive_amd.core.operator.FmmOperator` (or any
  object exposing ``_ed_native_solve``) and the preconditioner is Jacobi/none:
  the whole solve runs in the native FP64 device solver (``ed_gmres_solve``,
  one host poll every few Arnoldi steps, hipGraph-captured steps).
* any other callable: the reference control flow (MGS, Givens on host floats,
  full restart cycles, true-residual test at cycle end, retry without the
  preconditioner for N <= 512) with every vector operation done by the native
  FP64 kernels ``ed_vec_*`` on device memory.

Work precision is FP64 in both paths (the reference promotes FP32 input to
FP64 only for N <= 64, core/bem_solver.py:228-233); x is returned in b's dtype.
"""
import math
from typing import Any, Callable, Dict, Optional, Tuple, Union

import torch

from .. import _native as nat

SOLVER_TRACE_HEADERS = ["ts", "pass_idx", "cycle", "iter", "resid_true_l2", "resid_precond_l2",
                        "arnoldi_dim", "phase"]


class _Vec:
    """Native FP64 level-1 operations on device tensors."""

    def __init__(self, n, dev):
        self.n = n
        self.dev = dev
        self.scratch = torch.empty(256, dtype=torch.float64, device=dev)
        self.st = nat.stream_ptr(dev)
        self.L = nat.lib()

    def dot(self, x, y):
        out = nat.c_dbl(0.0)
        nat.check(self.L.ed_vec_dot_f64(nat.ptr(x), nat.ptr(y), self.n, nat.ptr(self.scratch),
                                        nat.ctypes.byref(out), self.st), "ed_vec_dot_f64")
        return out.value

    def axpy(self, a, x, y):
        nat.check(self.L.ed_vec_axpy_f64(float(a), nat.ptr(x), nat.ptr(y), self.n, self.st),
                  "ed_vec_axpy_f64")

    def scale(self, a, x, y):
        nat.check(self.L.ed_vec_scale_f64(float(a), nat.ptr(x), nat.ptr(y), self.n, self.st),
                  "ed_vec_scale_f64")

    def mul(self, d, x, y):
        nat.check(self.L.ed_vec_mul_f64(nat.ptr(d), nat.ptr(x), nat.ptr(y), self.n, self.st),
                  "ed_vec_mul_f64")

    def nrm2(self, x):
        return math.sqrt(max(self.dot(x, x), 0.0))


class _DiagPrecond:
    """Jacobi / mass preconditioner with the native elementwise kernel."""

    def __init__(self, inv):
        self.inv = inv

    def __call__(self, x):
        y = torch.empty_like(x, dtype=torch.float64)
        xx = x.to(torch.float64).contiguous()
        v = _Vec(xx.numel(), xx.device)
        v.mul(self.inv, xx, y)
        return y.to(x.dtype)


def setup_preconditioner(method: Optional[str], A_diag=None, areas=None, logger=None):
    """core/bem_solver.py:36-117."""
    if method is None or method == "none":
        return None
    if method in ("diag", "jacobi"):
        if A_diag is None:
            if areas is None:
                return None
            method = "mass"
        else:
            d = A_diag.to(torch.float64).clone()
            bad = (~torch.isfinite(d)) | (d.abs() < 1e-18)
            d[bad] = 1.0
            pc = _DiagPrecond((1.0 / d).contiguous())
            pc.kind, pc.diag = "jacobi", A_diag
            return pc
    if method == "mass":
        if areas is None:
            return None
        a = areas.to(torch.float64).clone()
        bad = (~torch.isfinite(a)) | (a.abs() < 1e-18)
        if not bool((~bad).any()):
            return None
        a[bad] = 1.0
        pc = _DiagPrecond((1.0 / a).contiguous())
        pc.kind, pc.diag = "mass", areas
        return pc
    return None


def _info(iters, resid, success, code, r0, tol_abs, used, **kw):
    d = dict(iters=int(iters), resid=float(resid), success=bool(success), code=int(code),
             resid_true0=float(r0), tol_abs=float(tol_abs), used_preconditioner=bool(used))
    d.update(kw)
    return d


@torch.no_grad()
def gmres_restart(A: Callable, b, M: Optional[Callable] = None, restart: int = 80,
                  tol: float = 1e-8, maxiter: int = 500, logger=None, x0=None, telemetry=None,
                  pass_index: int = 0, live_control_poll_fn=None, log_every: int = 1, *,
                  callback=None, callback_context=None, precond=None, areas=None, A_diag=None,
                  validate_operators: bool = False,
                  _retry_without_precond: bool = False) -> Tuple[Any, Dict[str, Any]]:
    if restart <= 0:
        raise ValueError("restart must be positive")
    if maxiter <= 0:
        raise ValueError("maxiter must be positive")
    nat.require_gpu()
    orig_dtype, orig_device = b.dtype, b.device
    N = int(b.numel())
    dev = b.device if b.is_cuda else torch.device("cuda", torch.cuda.current_device())

    if M is None and precond is not None:
        M_local = (setup_preconditioner(precond, A_diag=A_diag, areas=areas)
                   if isinstance(precond, str) else precond)
    else:
        M_local = M

    # ---- native fast path --------------------------------------------------
    native = getattr(A, "_ed_native_solve", None)
    if native is not None and (M_local is None or getattr(M_local, "kind", None) == "jacobi"):
        diag = None if M_local is None else M_local.diag
        x, info = native(b, x0=x0, diag=diag, restart=restart, tol=tol, maxiter=maxiter)
        info["used_preconditioner"] = M_local is not None
        if telemetry is not None:
            # one SOLVER_TRACE row per host poll of the native solver
            for row in A.plan.trace():
                try:
                    telemetry.append_row(dict(ts=row["ts"], pass_idx=int(pass_index),
                                              cycle=row["cycle"], iter=row["iter"],
                                              resid_true_l2=row["resid_true_l2"],
                                              resid_precond_l2=row["resid_precond_l2"],
                                              arnoldi_dim=row["arnoldi_dim"],
                                              phase=row["phase"]))
                except Exception:
                    pass
        return x.to(orig_device, orig_dtype), info

    # ---- generic path: reference control flow on native vector kernels ------
    V = _Vec(N, dev)

    def A_apply(v):
        out = A(v.to(orig_dtype).to(orig_device))
        return out.to(dev, torch.float64).contiguous()

    def M_apply(v):
        if M_local is None:
            return v
        return M_local(v.to(orig_dtype).to(orig_device)).to(dev, torch.float64).contiguous()

    if M_local is not None:
        try:
            probe = M_apply(torch.ones(N, dtype=torch.float64, device=dev))
            if probe.shape != b.shape or not bool(torch.isfinite(probe).all()):
                M_local = None
        except Exception:
            M_local = None
    used = M_local is not None

    def retry():
        return gmres_restart(A, b, None, restart, tol, maxiter, logger, x0, telemetry,
                             pass_index, live_control_poll_fn, log_every, callback=callback,
                             callback_context=callback_context, precond=None, areas=None,
                             validate_operators=False, _retry_without_precond=True)

    x = (torch.zeros(N, dtype=torch.float64, device=dev) if x0 is None
         else x0.to(dev, torch.float64).clone().contiguous())
    bw = b.to(dev, torch.float64).contiguous()
    r_true = bw.clone()
    V.axpy(-1.0, A_apply(x), r_true)
    r_true_norm = V.nrm2(r_true)
    if not math.isfinite(r_true_norm):
        return x.to(orig_device, orig_dtype), _info(0, r_true_norm, False, 3, r_true_norm,
                                                    float("nan"), used)
    r = M_apply(r_true)
    beta = V.nrm2(r)
    if not math.isfinite(beta):
        if M_local is not None and not _retry_without_precond:
            return retry()
        return x.to(orig_device, orig_dtype), _info(0, r_true_norm, False, 7, r_true_norm,
                                                    float("nan"), used)
    r_norm = beta
    init_true = r_true_norm if r_true_norm > 1e-16 else 1.0
    tol_abs = max(tol * init_true, tol * 1e-12, 1e-12)
    if beta == 0.0:
        if r_true_norm <= tol_abs:
            return x.to(orig_device, orig_dtype), _info(0, 0.0, True, 0, r_true_norm, tol_abs,
                                                        used, converged_on="true")
        return x.to(orig_device, orig_dtype), _info(0, r_true_norm, False, 6, r_true_norm,
                                                    tol_abs, used)
    total = 0
    max_cycles = max(1, (maxiter + restart - 1) // restart)
    for cycle in range(max_cycles):
        Vb = torch.empty((restart + 1, N), dtype=torch.float64, device=dev)
        V.scale(1.0 / beta, r, Vb[0])
        H = [[0.0] * restart for _ in range(restart + 1)]
        g = [0.0] * (restart + 1)
        g[0] = beta
        giv = []
        j = 0
        while j < restart and total < maxiter:
            if live_control_poll_fn is not None:
                state = live_control_poll_fn() or {}
                if state.get("terminate"):
                    return x.to(orig_device, orig_dtype), _info(total, float("nan"), False, 2,
                                                                 r_true_norm, tol_abs, used)
            w = M_apply(A_apply(Vb[j]))
            total += 1
            for i in range(j + 1):
                H[i][j] = V.dot(Vb[i], w)
                V.axpy(-H[i][j], Vb[i], w)
            H[j + 1][j] = V.nrm2(w)
            h_next = H[j + 1][j]
            if (not math.isfinite(h_next)) or h_next <= 0.0:
                if M_local is not None and not _retry_without_precond and N <= 512:
                    return retry()
                break
            V.scale(1.0 / h_next, w, Vb[j + 1])
            for ig, (cs, sn) in enumerate(giv):
                t = cs * H[ig][j] + sn * H[ig + 1][j]
                H[ig + 1][j] = -sn * H[ig][j] + cs * H[ig + 1][j]
                H[ig][j] = t
            den = math.sqrt(H[j][j] ** 2 + H[j + 1][j] ** 2)
            if den < 1e-14 or not math.isfinite(den):
                cs, sn = 1.0, 0.0
            else:
                cs, sn = H[j][j] / den, H[j + 1][j] / den
            giv.append((cs, sn))
            H[j][j] = cs * H[j][j] + sn * H[j + 1][j]
            H[j + 1][j] = 0.0
            g_j1 = -sn * g[j]
            g[j] = cs * g[j]
            g[j + 1] = g_j1
            r_norm = abs(g[j + 1])
            if not math.isfinite(r_norm):
                if M_local is not None and not _retry_without_precond and N <= 512:
                    return retry()
                return x.to(orig_device, orig_dtype), _info(total, r_norm, False, 5, r_true_norm,
                                                            tol_abs, used)
            if callback is not None:
                try:
                    ctx = dict(callback_context or {})
                    ctx.setdefault("pass_index", int(pass_index))
                    ctx.setdefault("arnoldi_dim", int(j + 1))
                    callback(int(total), float(r_norm), ctx)
                except Exception:
                    pass
            j += 1
        m = j
        if m == 0:
            if M_local is not None and not _retry_without_precond and N <= 512:
                return retry()
            return x.to(orig_device, orig_dtype), _info(total, r_norm, False, 2, r_true_norm,
                                                        tol_abs, used)
        y = [0.0] * m
        for i in range(m - 1, -1, -1):
            s = g[i]
            for k in range(i + 1, m):
                s -= H[i][k] * y[k]
            y[i] = s / H[i][i] if H[i][i] != 0.0 else 0.0
        for k in range(m):
            V.axpy(y[k], Vb[k], x)
        r_true = bw.clone()
        V.axpy(-1.0, A_apply(x), r_true)
        r_true_norm = V.nrm2(r_true)
        if not math.isfinite(r_true_norm):
            if M_local is not None and not _retry_without_precond and N <= 512:
                return retry()
            return x.to(orig_device, orig_dtype), _info(total, r_true_norm, False, 6, r_true_norm,
                                                        tol_abs, used)
        r = M_apply(r_true)
        beta = V.nrm2(r)
        r_norm = beta
        if telemetry is not None:
            try:
                import time
                telemetry.append_row(dict(ts=time.time(), pass_idx=int(pass_index),
                                          cycle=int(cycle), iter=int(total),
                                          resid_true_l2=r_true_norm, resid_precond_l2=r_norm,
                                          arnoldi_dim=int(m), phase="cycle_end"))
            except Exception:
                pass
        if r_true_norm <= tol_abs:
            return x.to(orig_device, orig_dtype), _info(total, r_true_norm, True, 0, r_true_norm,
                                                        tol_abs, used, converged_on="true")
        if total >= maxiter:
            break
    if M_local is not None and not _retry_without_precond and N <= 512:
        return retry()
    return x.to(orig_device, orig_dtype), _info(total, r_norm, False, 1, r_true_norm, tol_abs,
                                                used)


__all__ = ["SOLVER_TRACE_HEADERS", "setup_preconditioner", "gmres_restart"]
_ = (Union, Dict)
