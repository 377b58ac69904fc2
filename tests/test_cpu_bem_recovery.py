"""bem_solve's failure recovery (core/bem.py:1156-1251, :2170-2303), host logic
on the CPU: the damped Jacobi fall-back solve on small dense systems against
numpy, and every branch of the retry chain with stub solvers (the result keys
and messages of the reference's error dict)."""
import math

import numpy as np
import torch

from electrodrive_amd.core.bem import _jacobi_fallback_solve, _solve_with_recovery


class _Log:
    def __init__(self):
        self.msgs = []

    def __getattr__(self, level):
        return lambda msg, **kw: self.msgs.append((level, msg))


def test_jacobi_fallback_converges_diverges_and_rejects_nonfinite():
    rng = np.random.default_rng(0)
    n = 40
    M = rng.uniform(-0.2, 0.2, (n, n)) + np.diag(rng.uniform(8.0, 10.0, n))  # diagonally dominant
    b = rng.normal(size=n)
    A = torch.tensor(M)
    d = torch.tensor(np.diag(M).copy())
    x, info = _jacobi_fallback_solve(lambda v: A @ v, torch.tensor(b), d, _Log(), maxiter=500,
                                     tol=1e-10, omega=0.8)
    assert info["success"] and info["solver"] == "jacobi_fallback" and info["iters"] > 1
    r0 = np.linalg.norm(b)
    assert info["resid"] <= 1e-10 * max(r0, 1.0)
    assert np.linalg.norm(x.numpy() - np.linalg.solve(M, b)) < 1e-9 * np.linalg.norm(b)
    # omega lambda_max(D^-1 A) > 2: the iteration grows and never meets the target
    Mb = np.ones((n, n)) + np.eye(n)
    x, info = _jacobi_fallback_solve(lambda v: torch.tensor(Mb) @ v, torch.tensor(b),
                                     torch.full((n,), 2.0), _Log(), maxiter=30, tol=1e-6)
    assert not info["success"] and info["iters"] >= 1
    # a non-finite right-hand side stops before the first update
    bb = torch.tensor(b)
    bb[3] = float("nan")
    x, info = _jacobi_fallback_solve(lambda v: A @ v, bb, d, _Log())
    assert info["iters"] == 0 and not info["success"] and torch.count_nonzero(x) == 0


def _stub(outcomes):
    calls = []

    def solve(pc):
        calls.append(pc)
        return torch.zeros(3), dict(outcomes[pc])
    return solve, calls


def test_recovery_chain_branches():
    ok = {"success": True, "resid": 1e-9, "code": 0, "iters": 5}
    fail = {"success": False, "resid": 0.3, "code": 1, "iters": 2}
    nonfinite = {"success": True, "resid": float("inf"), "code": 0}
    jac_ok = lambda: (torch.ones(3), {"success": True, "resid": 1e-7, "iters": 9,  # noqa: E731
                                      "solver": "jacobi_fallback"})
    jac_bad = lambda: (torch.ones(3), {"success": False, "resid": 4.0, "iters": 256,  # noqa: E731
                                       "solver": "jacobi_fallback"})
    # 1. the preconditioned solve succeeds: no retry
    solve, calls = _stub({"sai": ok})
    s, info, err = _solve_with_recovery(solve, jac_bad, "sai", True, _Log())
    assert err is None and calls == ["sai"] and "retried_without_precond" not in info
    # 2. it fails; the retry without preconditioner succeeds
    solve, calls = _stub({"sai": fail, "none": ok})
    s, info, err = _solve_with_recovery(solve, jac_bad, "sai", True, _Log())
    assert err is None and calls == ["sai", "none"] and info["retried_without_precond"]
    # 3. a non-finite residual counts as a failure
    solve, calls = _stub({"sai": nonfinite, "none": ok})
    assert _solve_with_recovery(solve, jac_bad, "sai", True, _Log())[2] is None
    assert calls == ["sai", "none"]
    # 4. both fail, Jacobi disabled: the reference's GMRES error
    solve, calls = _stub({"sai": fail, "none": fail})
    s, info, err = _solve_with_recovery(solve, jac_ok, "sai", False, _Log())
    assert s is None and err[0] == "BEM linear solve failed: GMRES(code=1, resid=0.3)"
    assert err[1] == {"gmres_resid": 0.3, "gmres_code": 1}
    # 5. both fail, Jacobi succeeds
    log = _Log()
    s, info, err = _solve_with_recovery(solve, jac_ok, "sai", True, log)
    assert err is None and info["solver"] == "jacobi_fallback" and torch.equal(s, torch.ones(3))
    assert any("Jacobi fallback succeeded" in m for _, m in log.msgs)
    # 6. both fail, Jacobi fails: error with the Jacobi keys
    s, info, err = _solve_with_recovery(solve, jac_bad, "sai", True, _Log())
    assert err[0] == "BEM linear solve failed: GMRES(code=1, resid=0.3) and Jacobi(resid=4.0)"
    assert err[1] == {"gmres_resid": 0.3, "gmres_code": 1, "jacobi_resid": 4.0,
                      "jacobi_iters": 256}
    # 7. no preconditioner to drop: straight to Jacobi, no retry
    solve, calls = _stub({"none": fail})
    s, info, err = _solve_with_recovery(solve, jac_ok, "none", True, _Log())
    assert calls == ["none"] and info["solver"] == "jacobi_fallback"
    _ = math
