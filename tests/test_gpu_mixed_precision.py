"""GPU parity of the mixed-precision solve: BF16 / FP16 M2L on MFMA (FP32
accumulation) inside FP64 GMRES, with FP64 iterative refinement back to a
1e-8 relative residual against the accurate operator.

Workload: torus R = 1, r = 0.3 (the reference mesher's vertex and triangle
order, core/bem_mesh.py:180-212) at ~24k panels, one point charge drawn from
U([-2, 2]^3) outside the surface (rejected within 0.05 of it), refined self
integrals and near correction (depth 3).

Bounds:
- solver code 0 and ||b - A_acc x|| / ||b|| <= 1e-8, A_acc the FP64 dense +
  FP64 near operator (itself equal to the oracle's dense matvec to 1e-12 in
  tests/test_gpu_fmm.py);
- the solution within 2e-6 (relative L2) of the split-FP16 (FP32-accurate)
  M2L solve, refined to the same residual;
- the 16-bit FMM matvec against the oracle's FP64 dense matvec
  (oracle/csrc/dense_ref.c): BF16 <= 5e-3, FP16 <= 1e-3 relative L2;
- residual updates (later refinement residuals r <- r - A_fmm d with the
  split-FP16 FMM): fewer dense residuals than the all-dense refinement, which
  also reaches 1e-8, and the solver's final updated residual within its
  bound and 1e-9 |b| (a tenth of the tolerance) of the true one.
"""
import numpy as np
import pytest
import torch

from oracle import cdense

pytestmark = pytest.mark.gpu

ORDER, LEAF = 10, 128


def _charge(rng):
    while True:
        p = rng.uniform(-2.0, 2.0, 3)
        if np.hypot(np.hypot(p[0], p[1]) - 1.0, p[2]) > 0.3 + 0.05:
            return [(1.0, tuple(p))]


@pytest.fixture(scope="module")
def torus_case():
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator
    mesh = bem_mesh.torus_for_panels(1.0, 0.3, 24000)
    rng = np.random.default_rng(4)
    ch = _charge(rng)
    op = BemOperator(mesh, order=ORDER, leaf_size=LEAF, m2l_precision="fp16x3", near=True,
                     self_depth=3, near_depth=3)
    op.enable_sai()
    b = op.rhs_point_charges(ch)
    x, info = op.solve(b, tol=1e-8, precond="sai")
    assert info["code"] == 0
    sig = rng.normal(size=mesh.n_panels)
    ref_mv = cdense.dense_rows(mesh.centroids, mesh.areas, sig, op.self_corr.cpu().numpy())
    print(f"torus {mesh.n_panels} panels, charge {ch}, split-FP16 solve: iters {info['iters']}, "
          f"refine {info['refine_steps']}, rel resid {info['rel_resid']:.2e}")
    return dict(mesh=mesh, charges=ch, x_ref=x, sig=sig, ref_mv=ref_mv)


@pytest.mark.parametrize("prec,mv_bound", [("bf16", 5e-3), ("fp16", 1e-3)])
def test_mixed_precision_solve_refines_to_1e8(torus_case, prec, mv_bound):
    from electrodrive_amd.core.operator import INNER_TOL, BemOperator
    from electrodrive_amd.fmm3d.plan import FmmPlan
    mesh = torus_case["mesh"]
    op = BemOperator(mesh, order=ORDER, leaf_size=LEAF, m2l_precision=prec, near=True,
                     self_depth=3, near_depth=3)
    op.enable_sai()
    b = op.rhs_point_charges(torus_case["charges"])
    x, info = op.solve(b, tol=1e-8, precond="sai")
    bn = float(torch.linalg.norm(b))
    rt = float(torch.linalg.norm(b - op.accurate(x)))
    rr = rt / bn
    xr = torus_case["x_ref"]
    dx = float(torch.linalg.norm(x - xr) / torch.linalg.norm(xr))
    # the 16-bit FMM matvec alone (no near correction) vs the oracle
    plan = FmmPlan(mesh.centroids, mesh.areas, leaf_size=LEAF, order=ORDER, m2l_precision=prec)
    sig = torch.as_tensor(torus_case["sig"], device="cuda")
    mv = plan.matvec(sig, op.self_corr).cpu().numpy()
    ref = torus_case["ref_mv"]
    mv_err = float(np.linalg.norm(mv - ref) / np.linalg.norm(ref))
    # the same solve with every refinement residual dense
    x0, info0 = op.solve(b, tol=1e-8, precond="sai", residual_update=False)
    rr0 = float(torch.linalg.norm(b - op.accurate(x0))) / bn
    print(f"{prec}: code {info['code']} iters {info['iters']} refine {info['refine_steps']} "
          f"inner tol {INNER_TOL[prec]:g}; dense residuals {info['n_accurate']} (without "
          f"updates {info0['n_accurate']}), updates {info['n_updated']}, update operator err "
          f"{info['update_err']:.2e}; FP64 resid {rr:.2e} (solver's {info['rel_resid']:.2e}, "
          f"bound {info['resid_bound'] / bn:.2e}; without updates {rr0:.2e}); vs split-FP16 "
          f"solve {dx:.2e}; matvec vs oracle {mv_err:.2e}")
    assert info["code"] == 0 and info0["code"] == 0
    assert rr <= 1e-8 and rr0 <= 1e-8
    assert 1 <= info["refine_steps"] <= 16
    assert dx <= 2e-6
    assert mv_err <= mv_bound
    # residual updates: fewer dense residuals, and the updated residual the
    # solver stopped on within its bound and a tenth of the tolerance of the
    # true one
    assert info["n_updated"] >= 1 and info["n_accurate"] < info0["n_accurate"]
    assert abs(info["resid"] - rt) <= info["resid_bound"]
    assert abs(info["resid"] - rt) <= 1e-9 * bn
    # the updates run the M2L in split FP16: FP32-level error on the iterate
    assert info["update_err"] <= 2e-5
