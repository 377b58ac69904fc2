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
