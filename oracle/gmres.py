"""Restarted left-preconditioned GMRES restated from core/bem_solver.py (numpy).

Test infrastructure only (see oracle/__init__.py).

- ``setup_preconditioner``  core/bem_solver.py:36-117 (Jacobi 1/diag with
  non-finite or |d|<1e-18 entries -> 1; mass 1/A).
- ``gmres_restart``  core/bem_solver.py:125-920: MGS Arnoldi (:536-541),
  Givens rotations (:590-614), y = R^-1 g (:762-769), x += V y (:771-773),
  convergence only on the TRUE residual at cycle end (:854-865),
  tol_abs = max(tol*||b - A x0||, tol*1e-12, 1e-12) (:404-407), return codes
  0..7 (:203-216), retry without the preconditioner for N <= 512 (:472).
  Work dtype is float64 (the reference promotes to float64 only for N <= 64,
  :228-233; the oracle always works in float64).

Parity: pinned by SPD systems with known solutions (tests/test_gmres.py of the
reference, :12-44) and against numpy.linalg.solve on small dense systems.
"""
import math

import numpy as np


def setup_preconditioner(method, A_diag=None, areas=None):
    if method is None or method == "none":
        return None
    if method in ("diag", "jacobi"):
        if A_diag is None:
            if areas is None:
                return None
            method = "mass"
        else:
            d = np.array(A_diag, dtype=float, copy=True)
            bad = (~np.isfinite(d)) | (np.abs(d) < 1e-18)
            d[bad] = 1.0
            inv = 1.0 / d
            return lambda x: inv * x
    if method == "mass":
        if areas is None:
            return None
        a = np.array(areas, dtype=float, copy=True)
        bad = (~np.isfinite(a)) | (np.abs(a) < 1e-18)
        if not np.any(~bad):
            return None
        a[bad] = 1.0
        inv = 1.0 / a
        return lambda x: inv * x
    return None


def gmres_restart(A, b, M=None, restart=80, tol=1e-8, maxiter=500, x0=None,
                  precond=None, A_diag=None, areas=None,
                  _retry_without_precond=False):
    if restart <= 0:
        raise ValueError("restart must be positive")
    if maxiter <= 0:
        raise ValueError("maxiter must be positive")
    b = np.asarray(b, dtype=float)
    N = b.size
    M_local = M
    if M_local is None and precond is not None:
        M_local = (setup_preconditioner(precond, A_diag, areas)
                   if isinstance(precond, str) else precond)
    if M_local is not None:
        probe = M_local(np.ones_like(b))
        if probe.shape != b.shape or not np.all(np.isfinite(probe)):
            M_local = None
    Mx = (lambda v: v) if M_local is None else M_local
    used = M_local is not None

    def retry():
        return gmres_restart(A, b, None, restart, tol, maxiter, x0, None, None,
                             None, True)

    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=float, copy=True)
    r_true = b - A(x)
    r_true_norm = float(np.linalg.norm(r_true))
    if not math.isfinite(r_true_norm):
        return x, dict(iters=0, resid=r_true_norm, success=False, code=3,
                       resid_true0=r_true_norm, tol_abs=float("nan"),
                       used_preconditioner=used)
    r = Mx(r_true)
    beta = float(np.linalg.norm(r))
    if not math.isfinite(beta):
        if M_local is not None and not _retry_without_precond:
            return retry()
        return x, dict(iters=0, resid=r_true_norm, success=False, code=7,
                       resid_true0=r_true_norm, tol_abs=float("nan"),
                       used_preconditioner=used)
    r_norm = beta
    init_true = r_true_norm if r_true_norm > 1e-16 else 1.0
    tol_abs = max(tol * init_true, tol * 1e-12, 1e-12)
    if beta == 0.0:
        if r_true_norm <= tol_abs:
            return x, dict(iters=0, resid=0.0, success=True, code=0,
                           resid_true0=r_true_norm, tol_abs=tol_abs,
                           used_preconditioner=used, converged_on="true")
        return x, dict(iters=0, resid=r_true_norm, success=False, code=6,
                       resid_true0=r_true_norm, tol_abs=tol_abs,
                       used_preconditioner=used)
    total = 0
    max_cycles = max(1, (maxiter + restart - 1) // restart)
    for _cycle in range(max_cycles):
        V = [r / beta]
        H = np.zeros((restart + 1, restart))
        g = np.zeros(restart + 1)
        g[0] = beta
        giv = []
        j = 0
        while j < restart and total < maxiter:
            w = Mx(A(V[j]))
            total += 1
            for i in range(j + 1):
                H[i, j] = float(np.dot(V[i], w))
                w = w - H[i, j] * V[i]
            H[j + 1, j] = float(np.linalg.norm(w))
            h_next = H[j + 1, j]
            if (not math.isfinite(h_next)) or h_next <= 0.0:
                if M_local is not None and not _retry_without_precond and N <= 512:
                    return retry()
                break
            V.append(w / H[j + 1, j])
            for ig, (cs, sn) in enumerate(giv):
                t = cs * H[ig, j] + sn * H[ig + 1, j]
                H[ig + 1, j] = -sn * H[ig, j] + cs * H[ig + 1, j]
                H[ig, j] = t
            den = math.sqrt(H[j, j] ** 2 + H[j + 1, j] ** 2)
            if den < 1e-14 or not math.isfinite(den):
                cs, sn = 1.0, 0.0
            else:
                cs, sn = H[j, j] / den, H[j + 1, j] / den
            giv.append((cs, sn))
            H[j, j] = cs * H[j, j] + sn * H[j + 1, j]
            H[j + 1, j] = 0.0
            g_j1 = -sn * g[j]
            g[j] = cs * g[j]
            g[j + 1] = g_j1
            r_norm = abs(float(g[j + 1]))
            if not math.isfinite(r_norm):
                if M_local is not None and not _retry_without_precond and N <= 512:
                    return retry()
                return x, dict(iters=total, resid=r_norm, success=False, code=5,
                               resid_true0=r_true_norm, tol_abs=tol_abs,
                               used_preconditioner=used)
            j += 1
        m = j
        if m == 0:
            if M_local is not None and not _retry_without_precond and N <= 512:
                return retry()
            return x, dict(iters=total, resid=r_norm, success=False, code=2,
                           resid_true0=r_true_norm, tol_abs=tol_abs,
                           used_preconditioner=used)
        R = H[:m, :m]
        y = np.zeros(m)
        for i in range(m - 1, -1, -1):
            y[i] = (g[i] - R[i, i + 1:m] @ y[i + 1:m]) / R[i, i]
        x = x + np.stack(V[:m], axis=1) @ y
        r_true = b - A(x)
        r_true_norm = float(np.linalg.norm(r_true))
        if not math.isfinite(r_true_norm):
            if M_local is not None and not _retry_without_precond and N <= 512:
                return retry()
            return x, dict(iters=total, resid=r_true_norm, success=False, code=6,
                           resid_true0=r_true_norm, tol_abs=tol_abs,
                           used_preconditioner=used)
        r = Mx(r_true)
        beta = float(np.linalg.norm(r))
        r_norm = beta
        if r_true_norm <= tol_abs:
            return x, dict(iters=total, resid=r_true_norm, success=True, code=0,
                           resid_true0=r_true_norm, tol_abs=tol_abs,
                           used_preconditioner=used, converged_on="true")
        if total >= maxiter:
            break
    if M_local is not None and not _retry_without_precond and N <= 512:
        return retry()
    return x, dict(iters=total, resid=r_norm, success=False, code=1,
                   resid_true0=r_true_norm, tol_abs=tol_abs,
                   used_preconditioner=used)
