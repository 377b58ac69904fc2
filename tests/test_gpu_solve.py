"""GPU parity of the solve: native FP64 GMRES (+ iterative refinement) against
the oracle's FP64 dense direct solve and the closed forms.

- cfg1 (icosphere s=3, unit potential): Q K_E / (V R) -> 1 within the
  discretisation error (2e-2), and sigma equal to the FP64 dense direct
  solution of the same operator to 1e-6 relative.
- point charge outside a grounded sphere: sigma vs the Kelvin image density
  (discretisation error, 5e-2 relative L2) and total charge -qR/d.
- gmres_restart (generic callable path) vs oracle.gmres.gmres_restart.
"""
import numpy as np
import pytest
import torch

from oracle import analytic
from oracle import dense as odense
from oracle import gmres as ogmres
from oracle import quadrature as oq
from oracle.constants import K_E

pytestmark = pytest.mark.gpu


def _dense_matrix(C, A, self_corr, pairs=None, w=None):
    d = C[:, None, :] - C[None, :, :]
    r = np.sqrt(np.maximum((d * d).sum(-1), 1e-24))
    K = K_E / r
    np.fill_diagonal(K, self_corr)
    M = K * A[None, :]
    if pairs is not None:
        np.add.at(M, (pairs[:, 0], pairs[:, 1]), w)
    return M


def test_cfg1_capacitance_and_dense_solution():
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator
    m = bem_mesh.icosphere(3)
    op = BemOperator(m, order=10, ws=1, near=True, self_depth=3, near_depth=3)
    b = torch.ones(m.n_panels, dtype=torch.float64, device="cuda")
    x, info = op.solve(b, tol=1e-8, restart=50)
    print("cfg1 info", info)
    assert info["code"] == 0 and info["rel_resid"] <= 1e-8
    Q = op.total_charge(x)
    cap = Q * K_E / (1.0 * 1.0)
    print("Q K_E/(V R) =", cap)
    assert abs(cap - 1.0) < 2e-2
    # exact FP64 solve of the same operator
    I, corr = oq.self_integrals(m.centroids, m.panel_vertices, m.areas, depth=3)
    pairs = oq.near_pairs(m.centroids, m.areas, 1.5)
    w = oq.near_weights(m.centroids, m.areas, m.panel_vertices, pairs, 2, 3)
    Mx = _dense_matrix(m.centroids, m.areas, corr, pairs, w)
    xs = np.linalg.solve(Mx, np.ones(m.n_panels))
    err = np.linalg.norm(x.cpu().numpy() - xs) / np.linalg.norm(xs)
    print("sigma vs dense direct", err)
    assert err < 1e-6


def test_kelvin_point_charge():
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator
    m = bem_mesh.icosphere(4)
    op = BemOperator(m, order=8, ws=1)
    rng = np.random.default_rng(0)
    rho = rng.uniform(1.2, 2.0)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    r0 = rho * u
    b = op.rhs_point_charges([(1.0, tuple(r0))])
    x, info = op.solve(b, tol=1e-8)
    assert info["code"] == 0
    sig_k = analytic.kelvin_sigma(m.centroids, 1.0, r0, 1.0)
    err = np.linalg.norm(x.cpu().numpy() - sig_k) / np.linalg.norm(sig_k)
    Q = op.total_charge(x)
    print("kelvin rel err", err, "Q", Q, "expected", -1.0 / rho)
    assert err < 5e-2
    assert abs(Q - (-1.0 / rho)) < 2e-2 / rho


def test_gmres_restart_generic_matches_oracle():
    from electrodrive_amd.core.bem_solver import gmres_restart
    rng = np.random.default_rng(0)
    n = 40
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    Amat = Q @ np.diag(rng.uniform(1.0, 10.0, n)) @ Q.T
    xt = rng.normal(size=n)
    b = Amat @ xt
    At = torch.tensor(Amat).cuda()

    def A(v):
        return At @ v

    x, info = gmres_restart(A, torch.tensor(b).cuda(), restart=20, tol=1e-10, maxiter=200,
                            precond="jacobi", A_diag=torch.tensor(np.diag(Amat)).cuda())
    xo, info_o = ogmres.gmres_restart(lambda v: Amat @ v, b, restart=20, tol=1e-10, maxiter=200,
                                      precond="jacobi", A_diag=np.diag(Amat))
    assert info["code"] == 0 and info_o["code"] == 0
    assert info["iters"] == info_o["iters"]
    assert np.linalg.norm(x.cpu().numpy() - xt) / np.linalg.norm(xt) < 1e-8
    assert np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo) < 1e-9


def test_gmres_restart_native_path():
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.bem_solver import gmres_restart
    from electrodrive_amd.core.operator import BemOperator
    m = bem_mesh.icosphere(3)
    op = BemOperator(m, order=10, ws=1)
    b = torch.ones(m.n_panels, dtype=torch.float64, device="cuda")
    x, info = gmres_restart(op.fmm, b, restart=50, tol=1e-8, maxiter=500, precond="jacobi",
                            A_diag=op.diag, x0=b / op.diag)
    print("native gmres_restart", info)
    assert info["code"] == 0
    r = b - op.approx(x)
    assert float(torch.linalg.norm(r)) <= info["tol_abs"] * 1.0001


def test_native_solve_writes_solver_trace(tmp_path):
    """The native solve records a row per host poll; gmres_restart writes them
    as SOLVER_TRACE rows (core/bem_solver.py:19-28) through the telemetry
    writer: initial and cycle-end rows carry the true residual, which ends at
    or below tol_abs; iteration counts never decrease; Arnoldi polls come every
    check_every (4) steps; a refined BemOperator solve adds refinement rows."""
    import csv
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.bem_solver import SOLVER_TRACE_HEADERS, gmres_restart
    from electrodrive_amd.core.operator import BemOperator
    from electrodrive_amd.utils import TelemetryWriter
    m = bem_mesh.icosphere(3)
    op = BemOperator(m, order=10, ws=1)
    b = torch.ones(m.n_panels, dtype=torch.float64, device="cuda")
    tw = TelemetryWriter(tmp_path / "trace.csv", SOLVER_TRACE_HEADERS)
    x, info = gmres_restart(op.fmm, b, restart=50, tol=1e-8, maxiter=500, precond="jacobi",
                            A_diag=op.diag, x0=b / op.diag, telemetry=tw, pass_index=3)
    with open(tmp_path / "trace.csv") as f:
        rows = list(csv.DictReader(f))
    assert list(rows[0].keys()) == SOLVER_TRACE_HEADERS
    phases = [r["phase"] for r in rows]
    print(len(rows), phases[:6], info["iters"])
    assert phases[0] == "initial" and "arnoldi" in phases and phases[-1] == "cycle_end"
    assert all(r["pass_idx"] == "3" for r in rows)
    its = [int(r["iter"]) for r in rows]
    assert its == sorted(its) and its[-1] == info["iters"]
    polls = [r for r in rows if r["phase"] == "arnoldi"]
    assert int(polls[0]["iter"]) == 4 and polls[0]["resid_true_l2"] == "nan"
    assert float(rows[-1]["resid_true_l2"]) <= info["tol_abs"] * 1.0001
    assert abs(float(rows[0]["resid_true_l2"]) - info["resid_true0"]) <= 1e-9 * info["resid_true0"]
    # refined solve: the accurate operator's residuals appear as refine rows,
    # residual updates as update rows (bound in the precond column); the last
    # of them met the tolerance at the last refinement step
    for upd in (False, True):
        x, info = op.solve(b, tol=1e-8, residual_update=upd)
        tr = op.plan.trace()
        ref = [r for r in tr if r["phase"] == "refine"]
        outer = [r for r in tr if r["phase"] in ("refine", "update")]
        assert len(ref) == info["n_accurate"] >= 1
        assert len(outer) - len(ref) == info["n_updated"]
        last = outer[-1]
        bound = last["resid_precond_l2"] if last["phase"] == "update" else 0.0
        assert last["resid_true_l2"] + bound <= 1e-8 * float(torch.linalg.norm(b)) * 1.0001
        assert last["refine_step"] == info["refine_steps"]


def test_sai_rows_match_oracle():
    """ed_plan_build_sai / ed_plan_sai_apply against oracle.sai (pinned by the
    exact-inverse case) on a 320-panel sphere with the near correction."""
    from oracle import sai as osai
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator
    m = bem_mesh.icosphere(2)
    op = BemOperator(m, order=8, ws=1, leaf_size=16, near=True, self_depth=3, near_depth=2)
    op.enable_sai()
    near = (op.near_pairs.cpu().numpy(), op.near_w.cpu().numpy())
    M = osai.sai_matrix(m.centroids, m.areas, op.self_corr.cpu().numpy(), near,
                        perm=op.plan.permutation())
    rng = np.random.default_rng(5)
    for _ in range(3):
        u = rng.normal(size=op.n)
        z = op.plan.sai_apply(torch.tensor(u, device=op.device)).cpu().numpy()
        ref = M @ u
        err = np.linalg.norm(z - ref) / np.linalg.norm(ref)
        print("sai apply rel err", err)
        assert err < 1e-9


def test_graph_replay_matches_direct_launches():
    """The Arnoldi step replayed from a captured hipGraph (use_graph=True)
    gives the solve of direct launches (the default) bit for bit, with the
    Jacobi and the SAI smoother, twice in a row (the second call replays the
    graph captured by the first)."""
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator
    m = bem_mesh.icosphere(4)
    op = BemOperator(m, order=10, ws=1, leaf_size=64)
    b = op.rhs_point_charges([(1.0, (0.3, -0.5, 1.4))])
    for precond in (True, "sai"):
        xd, idd = op.solve(b, tol=1e-8, precond=precond)
        for _ in range(2):
            xg, ig = op.solve(b, tol=1e-8, precond=precond, use_graph=True)
            print(precond, "iters direct", idd["iters"], "graph", ig["iters"])
            assert ig["code"] == 0 and ig["iters"] == idd["iters"]
            assert torch.equal(xg, xd)


def test_sai_solve_same_solution_fewer_iterations():
    """Jacobi and SAI smoothers (with and without a fixed coarse space) reach
    the same FP64-refined solution; the SAI takes far fewer iterations."""
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator
    from electrodrive_amd.learn import precond as lp
    m = bem_mesh.icosphere(4)
    op = BemOperator(m, order=10, ws=1, leaf_size=64)
    b = op.rhs_point_charges([(1.0, (0.3, -0.5, 1.4))])
    xj, ij = op.solve(b, tol=1e-8)
    xs, is_ = op.solve(b, tol=1e-8, precond="sai")
    xc, ic = lp.baseline_solve(op, b, tol=1e-8, precond="sai")
    print("iters jacobi", ij["iters"], "sai", is_["iters"], "sai+coarse", ic["iters"])
    for x, info in ((xs, is_), (xc, ic)):
        assert info["code"] == 0 and info["rel_resid"] <= 1e-8
        assert float(torch.linalg.norm(x - xj) / torch.linalg.norm(xj)) < 1e-6
    assert is_["iters"] * 1.5 <= ij["iters"]
