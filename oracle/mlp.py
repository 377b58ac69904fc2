"""FP64 numpy forward/backward of the learned-preconditioner MLP.

Test infrastructure only (see oracle/__init__.py).  No reference counterpart
(the reference has no learned preconditioner): this is the CPU autograd oracle
the native kernels' gradients are checked against.  Same layout as
electrodrive_amd/learn/mlp.py: flat parameters [W0 (out x in), b0, W1, b1,
...], SiLU between linear layers, none after the last.

Parity: pinned in tests by central finite differences of the loss.
"""
import numpy as np


def silu(x):
    return x / (1.0 + np.exp(-x))


def silu_grad(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def unpack(params, dims):
    out, o = [], 0
    for i in range(len(dims) - 1):
        fi, fo = dims[i], dims[i + 1]
        W = params[o:o + fo * fi].reshape(fo, fi)
        o += fo * fi
        b = params[o:o + fo]
        o += fo
        out.append((W, b))
    return out


def forward(params, dims, X):
    layers = unpack(np.asarray(params, dtype=np.float64), dims)
    h = np.asarray(X, dtype=np.float64)
    acts, pres = [h], []
    for i, (W, b) in enumerate(layers):
        z = h @ W.T + b
        pres.append(z)
        h = z if i == len(layers) - 1 else silu(z)
        acts.append(h)
    return h, acts, pres


def loss_and_grad(params, dims, X, T, w=None):
    """loss = sum w (y - t)^2 with y = MLP(X); returns (loss, dloss/dparams)."""
    y, acts, pres = forward(params, dims, X)
    T = np.asarray(T, dtype=np.float64).reshape(y.shape)
    w = np.ones_like(y) if w is None else np.asarray(w, dtype=np.float64).reshape(y.shape)
    d = y - T
    loss = float(np.sum(w * d * d))
    g = 2.0 * w * d
    layers = unpack(np.asarray(params, dtype=np.float64), dims)
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        W, b = layers[i]
        if i < len(layers) - 1:
            g = g * silu_grad(pres[i])
        grads.append((g.T @ acts[i], g.sum(0)))
        g = g @ W
    flat = []
    for gW, gb in reversed(grads):
        flat.append(gW.ravel())
        flat.append(gb.ravel())
    return loss, np.concatenate(flat)


def proj_loss_and_grad(Y, X, T, m, ridge=1e-6):
    """Projection loss of a learned coarse basis, per solve of m rows:
    Z = [Y | X], c = argmin |t - Zc|^2 + lam |c|^2 with lam = ridge tr(Z^T Z)/k,
    loss_s = |t - Zc|^2 / |t|^2, dloss_s/dY = -2 (t - Zc) c_Y^T / |t|^2 (the
    envelope theorem, lam held fixed).  Returns (loss (n_sol,), dY like Y).
    No reference counterpart; pinned in tests by finite differences."""
    Y = np.asarray(Y, dtype=np.float64)
    X = np.zeros((Y.shape[0], 0)) if X is None else np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64).reshape(-1)
    ky = Y.shape[1]
    n_sol = Y.shape[0] // m
    loss = np.zeros(n_sol)
    dY = np.zeros_like(Y)
    for s in range(n_sol):
        sl = slice(s * m, (s + 1) * m)
        Z = np.concatenate([Y[sl], X[sl]], axis=1)
        t = T[sl]
        G = Z.T @ Z
        k = G.shape[0]
        tr = np.trace(G)
        lam = ridge * (tr if tr > 0 else 1.0) / k
        c = np.linalg.solve(G + lam * np.eye(k), Z.T @ t)
        r = t - Z @ c
        tt = float(t @ t)
        inv = 1.0 / tt if tt > 0 else 0.0
        loss[s] = float(r @ r) * inv
        dY[sl] = -2.0 * np.outer(r, c[:ky]) * inv
    return loss, dY
