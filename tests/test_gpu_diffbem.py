"""GPU: differentiable kernels and the adjoint solve (core/bem_kernel.py:157-335,
:1016-1079; core/diffbem.py).

Gradients are checked against central finite differences of the oracle's
numpy restatements (oracle.dense.dense_matvec, direct_potential,
direct_E_field; FP64) and of numpy's dense solve of the oracle operator
(oracle.sai.operator_entries), and against the closed form
d/dalpha mean(sigma(alpha b)^2) = 2 mean(sigma(b)^2) at alpha = 1.
"""
import numpy as np
import pytest
import torch

from oracle import dense as od
from oracle import sai as osai

pytestmark = pytest.mark.gpu


def _geom(seed=0, n_sub=2):
    from electrodrive_amd.core import bem_mesh
    m = bem_mesh.icosphere(n_sub)
    rng = np.random.default_rng(seed)
    return m, rng


def _fd(f, x, h):
    g = np.zeros_like(x)
    for idx in np.ndindex(x.shape):
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        g[idx] = (f(xp) - f(xm)) / (2 * h)
    return g


def test_dense_matvec_gradients_match_finite_differences():
    from electrodrive_amd.core.bem_kernel import bem_matvec_diff
    m, rng = _geom(1, 1)  # 80 panels
    C, A = m.centroids, m.areas
    n = len(A)
    sig = rng.normal(size=n)
    selfc = 2 * np.pi * 8.9875517873681764e9 * np.sqrt(A / np.pi) / A
    w = rng.normal(size=n)
    dev = torch.device("cuda")
    t = {k: torch.tensor(v, device=dev, requires_grad=True)
         for k, v in dict(s=sig, C=C, A=A, si=selfc).items()}
    V = bem_matvec_diff(t["s"], t["C"], t["A"], self_integrals=t["si"])
    Vr = od.dense_matvec(C, A, sig, selfc)
    assert np.max(np.abs(V.detach().cpu().numpy() - Vr)) < 1e-12 * np.max(np.abs(Vr))
    (V * torch.tensor(w, device=dev)).sum().backward()
    L = lambda s=sig, c=C, a=A, si=selfc: float(w @ od.dense_matvec(c, a, s, si))
    checks = {"s": _fd(lambda x: L(s=x), sig, 1e-6),
              "A": _fd(lambda x: L(a=x), A, 1e-9),
              "si": _fd(lambda x: L(si=x), selfc, 1e-3 * selfc.mean()),
              "C": _fd(lambda x: L(c=x), C, 1e-7)}
    for k, ref in checks.items():
        got = t[k].grad.cpu().numpy()
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print("dense grad", k, err)
        assert err < 1e-5, k


def test_target_kernel_gradients_match_finite_differences():
    from electrodrive_amd.core.bem_kernel import (bem_E_field_targets_diff,
                                                  bem_potential_targets_diff)
    m, rng = _geom(2, 1)
    C, A = m.centroids, m.areas
    n = len(A)
    sig = rng.normal(size=n)
    X = rng.uniform(-2, 2, (17, 3))
    X[:3] = C[:3] * 1.3
    w = rng.normal(size=17)
    W = rng.normal(size=(17, 3))
    dev = torch.device("cuda")
    t = {k: torch.tensor(v, device=dev, requires_grad=True)
         for k, v in dict(x=X, C=C, A=A, s=sig).items()}
    V = bem_potential_targets_diff(t["x"], t["C"], t["A"], t["s"])
    (V * torch.tensor(w, device=dev)).sum().backward()
    L = lambda x=X, c=C, a=A, s=sig: float(w @ od.direct_potential(x, c, a, s))
    for k, ref in {"x": _fd(lambda v: L(x=v), X, 1e-7), "C": _fd(lambda v: L(c=v), C, 1e-7),
                   "A": _fd(lambda v: L(a=v), A, 1e-9), "s": _fd(lambda v: L(s=v), sig, 1e-6)}.items():
        err = np.max(np.abs(t[k].grad.cpu().numpy() - ref)) / np.max(np.abs(ref))
        print("target V grad", k, err)
        assert err < 1e-5, k
    ts = torch.tensor(sig, device=dev, requires_grad=True)
    ta = torch.tensor(A, device=dev, requires_grad=True)
    E = bem_E_field_targets_diff(torch.tensor(X, device=dev), torch.tensor(C, device=dev), ta, ts)
    (E * torch.tensor(W, device=dev)).sum().backward()
    LE = lambda a=A, s=sig: float((W * od.direct_E_field(X, C, a, s)).sum())
    for k, got, ref in (("s", ts.grad, _fd(lambda v: LE(s=v), sig, 1e-6)),
                        ("A", ta.grad, _fd(lambda v: LE(a=v), A, 1e-9))):
        err = np.max(np.abs(got.cpu().numpy() - ref)) / np.max(np.abs(ref))
        print("target E grad", k, err)
        assert err < 1e-5, k
    with pytest.raises(NotImplementedError):
        tx = torch.tensor(X, device=dev, requires_grad=True)
        bem_E_field_targets_diff(tx, torch.tensor(C, device=dev), ta, ts).sum().backward()


def test_solve_diffbem_adjoint_gradients():
    """sigma = A^-1 b on 320 panels; L = w . sigma.  dL/db and dL/dareas from
    the adjoint solve against finite differences of numpy's dense solve."""
    from electrodrive_amd.core.bem_kernel import bem_matvec_diff
    from electrodrive_amd.core.diffbem import solve_diffbem
    from electrodrive_amd.utils import BEMConfig
    m, rng = _geom(3, 2)
    C, A = m.centroids, m.areas
    n = len(A)
    selfc = 2 * np.pi * 8.9875517873681764e9 * np.sqrt(A / np.pi) / A
    b = rng.normal(size=n) * 1e10
    w = rng.normal(size=n)
    dev = torch.device("cuda")
    tb = torch.tensor(b, device=dev, requires_grad=True)
    tA = torch.tensor(A, device=dev, requires_grad=True)
    tC = torch.tensor(C, device=dev)
    tsc = torch.tensor(selfc, device=dev)
    cfg = BEMConfig()
    cfg.gmres_tol = 1e-13
    cfg.gmres_maxiter = 2000
    out = solve_diffbem(None, cfg, None, C=tC, N=tC, A=tA, rhs=tb,
                        matvec=lambda s: bem_matvec_diff(s, tC, tA, self_integrals=tsc))
    sig = out["sigma"]
    assert out["stats"]["success"]

    def solve_np(bb=b, aa=A):
        return np.linalg.solve(osai.operator_entries(C, aa, selfc), bb)
    s_ref = solve_np()
    assert np.linalg.norm(sig.detach().cpu().numpy() - s_ref) / np.linalg.norm(s_ref) < 1e-9
    (sig * torch.tensor(w, device=dev)).sum().backward()
    gb_ref = np.linalg.solve(osai.operator_entries(C, A, selfc).T, w)  # exact adjoint
    gA_ref = _fd(lambda a: float(w @ solve_np(aa=a)), A, 1e-9)
    eb = np.linalg.norm(tb.grad.cpu().numpy() - gb_ref) / np.linalg.norm(gb_ref)
    eA = np.max(np.abs(tA.grad.cpu().numpy() - gA_ref)) / np.max(np.abs(gA_ref))
    print("adjoint grad b", eb, "areas", eA)
    assert eb < 1e-8 and eA < 1e-5


def test_bem_solve_differentiable_end_to_end(tmp_path):
    """bem_solve(differentiable=True) on the reference's grounded-plane test
    spec: d/dalpha of mean(sigma(alpha b)^2) is
    2 mean(sigma^2) at alpha = 1; the solution matches the non-differentiable
    solve of the same operator; gradients reach the charge height through the
    evaluation kernels, and grad V = -E at probe points."""
    from electrodrive_amd.core.bem import bem_solve
    from electrodrive_amd.core.bem_kernel import (_bem_potential_targets_core_torch,
                                                  bem_matvec_diff)
    from electrodrive_amd.core.diffbem import solve_diffbem
    from electrodrive_amd.orchestration import CanonicalSpec
    from electrodrive_amd.utils import BEMConfig
    spec = CanonicalSpec.from_json({"domain": "R3", "BCs": "Dirichlet",
                                    "conductors": [{"type": "plane", "z": 0.0, "potential": 0.0}],
                                    "dielectrics": [],
                                    "charges": [{"type": "point", "q": 1.0, "pos": [0, 0, 0.5]}]})
    cfg = BEMConfig(fp64=True, initial_h=0.4, max_refine_passes=1, gmres_tol=1e-10,
                    gmres_maxiter=400)
    out = bem_solve(spec, cfg, None, differentiable=True)
    assert "error" not in out, out.get("error")
    sol = out["solution"]
    sigma = sol._S
    ref = bem_solve(spec, cfg, None, differentiable=False)["surface_charge_density"]
    d = np.linalg.norm(sigma.detach().cpu().numpy() - ref) / np.linalg.norm(ref)
    print("diff vs native solve", d)
    assert d < 1e-7
    # rhs scale: sigma(alpha b) = alpha sigma(b)
    C, A = sol._C.detach(), sol._A.detach()
    sc = 2 * np.pi * 8.9875517873681764e9 * torch.sqrt(A / np.pi) / A
    base = torch.tensor(np.random.default_rng(0).normal(size=C.shape[0]), device=C.device)
    alpha = torch.tensor(1.0, device=C.device, dtype=torch.float64, requires_grad=True)
    o2 = solve_diffbem(None, cfg, None, C=C, N=C, A=A, rhs=alpha * base,
                       matvec=lambda s: bem_matvec_diff(s, C, A, self_integrals=sc))
    loss = (o2["sigma"] ** 2).mean()
    loss.backward()
    expect = 2.0 * float(loss.detach())
    print("d/dalpha", float(alpha.grad), "expected", expect)
    assert abs(float(alpha.grad) - expect) < 1e-6 * abs(expect)
    # end to end: probes above the plane, charge height in the free-space part
    probes = torch.tensor([[0.0, 0.0, 0.25], [0.1, 0.0, 0.35], [-0.1, 0.0, 0.45]],
                          dtype=torch.float64, device=C.device)
    z = torch.tensor(0.5, dtype=torch.float64, device=C.device, requires_grad=True)
    pos = torch.stack([torch.zeros_like(z), torch.zeros_like(z), z])
    V_free = 8.9875517873681764e9 / torch.linalg.norm(probes - pos[None, :], dim=1)
    V_ind = _bem_potential_targets_core_torch(probes, C, A, sigma)
    (V_free + V_ind).mean().backward()
    assert z.grad is not None and np.isfinite(float(z.grad))
    # the evaluation is differentiable in the point: grad V = -E
    P = probes.clone().requires_grad_(True)
    V, E = sol.eval_V_E_batched(P)
    V.sum().backward()
    err = float(torch.linalg.norm(P.grad + E.detach()) / torch.linalg.norm(E.detach()))
    print("grad V vs -E", err)
    assert err < 1e-8
