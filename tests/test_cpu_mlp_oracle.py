"""Pin of the MLP oracle (oracle/mlp.py) by central finite differences."""
import numpy as np

from oracle import mlp as om


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = (5, 7, 6, 3)
    n_p = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))
    p = rng.normal(size=n_p) * 0.5
    X = rng.normal(size=(11, 5))
    T = rng.normal(size=(11, 3))
    w = rng.uniform(0.5, 2.0, size=(11, 3))
    loss, g = om.loss_and_grad(p, dims, X, T, w)
    h = 1e-6
    for k in rng.choice(n_p, 25, replace=False):
        e = np.zeros(n_p)
        e[k] = h
        fd = (om.loss_and_grad(p + e, dims, X, T, w)[0] - om.loss_and_grad(p - e, dims, X, T, w)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd))
    y, _, _ = om.forward(p, dims, X)
    assert abs(loss - np.sum(w * (y - T) ** 2)) < 1e-12


def test_proj_loss_oracle_gradient_matches_finite_differences():
    """Pins oracle.mlp.proj_loss_and_grad: ridge 0 makes the loss the exact
    least-squares minimum, whose derivative is the envelope-theorem gradient;
    central differences must agree.  Also: a target inside span(Z) has zero
    loss, and the loss is invariant to scaling t and to rescaling a column."""
    import numpy as np
    from oracle import mlp as om
    rng = np.random.default_rng(3)
    m, ky, kx, ns = 40, 3, 2, 2
    Y = rng.normal(size=(ns * m, ky))
    X = rng.normal(size=(ns * m, kx))
    T = rng.normal(size=ns * m)
    loss, dY = om.proj_loss_and_grad(Y, X, T, m, ridge=0.0)
    h = 1e-6
    for idx in [(0, 0), (5, 2), (m + 3, 1), (2 * m - 1, 0)]:
        Yp, Ym = Y.copy(), Y.copy()
        Yp[idx] += h
        Ym[idx] -= h
        fd = (om.proj_loss_and_grad(Yp, X, T, m, 0.0)[0].sum()
              - om.proj_loss_and_grad(Ym, X, T, m, 0.0)[0].sum()) / (2 * h)
        assert abs(fd - dY[idx]) < 1e-6 * max(1.0, abs(fd))
    Tin = (np.concatenate([Y, X], 1)[:m] @ rng.normal(size=ky + kx))
    l0, _ = om.proj_loss_and_grad(Y[:m], X[:m], Tin, m, ridge=0.0)
    assert l0[0] < 1e-20
    l1, _ = om.proj_loss_and_grad(Y * [1.0, 7.0, 0.1], X, 3.0 * T, m, ridge=0.0)
    assert np.allclose(l1, loss, rtol=1e-10)
