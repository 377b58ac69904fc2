"""GPU: what the trained network gains (held-out seeds, all four families).

The committed weights (electrodrive_amd/learn/weights/precond_mlp.npz) were
trained on 10,240 solver-produced solves with seed 0; these geometries and
charges come from seed 12345.  Each case is solved to the same 1e-8 relative
residual (FP64 dense + near-correction operator) four ways:

- plain: Jacobi GMRES from x0 = b/diag (the reference method)
- sai: sparse approximate inverse smoother, no network
- fixed: SAI + coarse space of the fixed columns only (monomials, Jacobi
  guess), x0 its coarse solution -- the method without the network
- learned: SAI + coarse space with the network channels, x0 the coarse
  solution on it

and the solutions must agree with plain GMRES's.  Plain GMRES is the judge.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _cases():
    from electrodrive_amd.learn.train import FAMILIES, make_geometry, random_charges
    rng = np.random.default_rng(12345)
    out = []
    for gi in range(4):
        mesh, g = make_geometry(rng, FAMILIES[gi], size="small")
        out.append((FAMILIES[gi], mesh, [random_charges(rng, g, 1), random_charges(rng, g, 2)]))
    return out


def test_learned_preconditioner_gains():
    from electrodrive_amd.core.operator import BemOperator
    from electrodrive_amd.learn import precond as lp_mod
    lp = lp_mod.LearnedPreconditioner()  # committed weights
    assert lp.n_out == lp_mod.N_OUT
    tot = {"plain": 0, "sai": 0, "fixed": 0, "learned": 0}
    for fam, mesh, charge_sets in _cases():
        op = BemOperator(mesh, order=10, ws=1, leaf_size=128)
        for ch in charge_sets:
            b = op.rhs_point_charges(ch)
            xp, ip = op.solve(b, tol=1e-8, check_every=1)
            xs, is_ = op.solve(b, tol=1e-8, precond="sai", check_every=1)
            xf, if_ = lp_mod.baseline_solve(op, b, tol=1e-8, precond="sai", check_every=1)
            xl, il = lp.solve(op, charges=ch, tol=1e-8, precond="sai", check_every=1)
            assert il["learned"], "learned solve fell back to plain GMRES"
            for x, info in ((xs, is_), (xf, if_), (xl, il)):
                assert info["code"] == 0
                rr = float(torch.linalg.norm(b - op.accurate(x)) / torch.linalg.norm(b))
                assert rr <= 1e-8
                assert float(torch.linalg.norm(x - xp) / torch.linalg.norm(xp)) < 1e-5
            print(fam, len(ch), "iters plain", ip["iters"], "sai", is_["iters"], "fixed",
                  if_["iters"], "learned", il["iters"])
            tot["plain"] += ip["iters"]
            tot["sai"] += is_["iters"]
            tot["fixed"] += if_["iters"]
            tot["learned"] += il["iters"]
        del op
    print("totals", tot)
    # the learned-preconditioned solve needs at least 2x fewer iterations than
    # plain GMRES, and the network beats the same method without it
    assert 2 * tot["learned"] <= tot["plain"]
    assert tot["learned"] < tot["fixed"]


def test_learned_x0_is_a_charge_density_prediction():
    """x0 = Z (Z^T A Z)^-1 Z^T b on the learned basis is a usable prediction:
    for a held-out sphere and one charge, within 5% of the solution and far
    closer than the Jacobi guess b/diag."""
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator
    from electrodrive_amd.learn import precond as lp_mod
    lp = lp_mod.LearnedPreconditioner()
    op = BemOperator(bem_mesh.icosphere(4, 1.3), order=10, ws=1, leaf_size=128)
    ch = [(0.8, (0.4, -1.6, 0.9))]
    b = op.rhs_point_charges(ch)
    x, _ = op.solve(b, tol=1e-10)
    x0 = lp.predict(op, b, lp_mod.rhs_components(op, ch))
    op.plan.set_coarse(None)
    e_learned = float(torch.linalg.norm(x0 - x) / torch.linalg.norm(x))
    e_jacobi = float(torch.linalg.norm(b / op.diag - x) / torch.linalg.norm(x))
    print("x0 error learned", e_learned, "jacobi", e_jacobi)
    assert e_learned < 0.05
    assert e_learned < 0.1 * e_jacobi
