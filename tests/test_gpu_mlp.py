"""GPU: native MLP kernels (forward, backward, Adam) against the FP64 oracle
(oracle/mlp.py, pinned by finite differences).  FP32 tolerances."""
import numpy as np
import pytest
import torch

from oracle import mlp as om

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,B", [((12, 256, 256, 256, 1), 1000), ((7, 33, 65, 17, 5), 129)])
def test_native_mlp_matches_oracle(dims, B):
    from electrodrive_amd.learn.mlp import NativeMLP
    net = NativeMLP(dims, seed=3)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(B, dims[0])).astype(np.float32)
    T = rng.normal(size=(B, dims[-1])).astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=(B, dims[-1])).astype(np.float32)
    p = net.params.cpu().numpy().astype(np.float64)
    y = net.forward(torch.tensor(X).cuda(), keep=True)
    yo, _, _ = om.forward(p, dims, X)
    assert np.max(np.abs(y.cpu().numpy() - yo)) <= 1e-4 * (1 + np.max(np.abs(yo)))
    loss, g = net.wmse(y, torch.tensor(T).cuda(), torch.tensor(w).cuda())
    net.backward(g.view(B, dims[-1]))
    lo, go = om.loss_and_grad(p, dims, X, T, w)
    assert abs(loss - lo) <= 1e-4 * lo
    gn = net.grads.cpu().numpy()
    err = np.linalg.norm(gn - go) / np.linalg.norm(go)
    print("grad rel err", err)
    assert err < 1e-4
    # one Adam step: p' = p - lr * m/(sqrt(v)+eps) with bias correction -> lr*sign(g)
    before = net.params.cpu().numpy()
    net.adam(lr=1e-3)
    after = net.params.cpu().numpy()
    nz = np.abs(gn) > 1e-3  # |g| >> eps: the first Adam step is lr * sign(g)
    assert np.allclose((before - after)[nz], 1e-3 * np.sign(gn[nz]), rtol=1e-3, atol=1e-7)


@pytest.mark.parametrize("ky,kx,m,ns", [(8, 11, 512, 5), (3, 0, 77, 2), (16, 16, 300, 3)])
def test_native_proj_loss_matches_oracle(ky, kx, m, ns):
    """ed_proj_loss_f32 (loss and dY) against oracle.mlp.proj_loss_and_grad."""
    import ctypes
    from electrodrive_amd import _native as nat
    rng = np.random.default_rng(ky + kx)
    Y = rng.normal(size=(ns * m, ky)).astype(np.float32)
    X = rng.normal(size=(ns * m, kx)).astype(np.float32)
    T = (Y[:, :1] * 0.5 + rng.normal(size=(ns * m, 1))).astype(np.float32).reshape(-1)
    ridge = 1e-6
    lo, dYo = om.proj_loss_and_grad(Y, X, T, m, ridge=ridge)
    dev = torch.device("cuda")
    Yd, Xd, Td = (torch.tensor(a, device=dev) for a in (Y, X, T))
    dY = torch.empty_like(Yd)
    loss = torch.empty(ns, dtype=torch.float32, device=dev)
    nat.check(nat.lib().ed_proj_loss_f32(nat.ptr(Yd), ky, nat.ptr(Xd) if kx else None, kx,
                                         nat.ptr(Td), ns, m, ctypes.c_float(ridge),
                                         nat.ptr(dY), nat.ptr(loss),
                                         nat.stream_ptr(dev)), "ed_proj_loss_f32")
    l = loss.cpu().numpy()
    g = dY.cpu().numpy()
    print("loss", l, "oracle", lo)
    assert np.allclose(l, lo, rtol=1e-4, atol=1e-7)
    assert np.linalg.norm(g - dYo) <= 1e-4 * np.linalg.norm(dYo)
