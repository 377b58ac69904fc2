"""Training of the learned coarse space on solves produced by the solver.

Geometry families follow the batched workload: spheres (R ~ U[0.5, 2]),
tori (R = 1, r ~ U[0.2, 0.4]), sphere dimers (gap ~ U[0.05, 0.5]) and square
plates.  Every training solve has ONE RHS component -- a point charge placed
outside the surface (clearance >= 0.1) or, once per geometry, a unit
boundary potential -- because the learned coarse space is built per RHS
component.  Each geometry is solved for many right-hand sides with the native
FP64 solver; for every solve the features, fixed coarse columns and solution
sigma of ``panels_per_solve`` random panels are kept.  Training minimises the
projection loss of sigma onto span([network channels, fixed columns]) with
the native kernels (MLP forward/backward, ``ed_proj_loss_f32``, Adam).

``python -m electrodrive_amd.learn.train --solves 10240 --out <path.npz>``
"""
import argparse
import ctypes
import math
import time

import numpy as np
import torch

from .. import _native as nat
from ..core import bem_mesh
from ..core.operator import BemOperator
from .mlp import NativeMLP
from .precond import (DEFAULT_DIMS, LearnedPreconditioner, fixed_columns, panel_features,
                      rhs_components)

FAMILIES = ("sphere", "torus", "dimer", "plate")


def make_geometry(rng, family, size="small"):
    big = size == "large"
    if family == "sphere":
        R = rng.uniform(0.5, 2.0)
        sub = 5 if big else (4 if rng.uniform() < 0.7 else 3)
        return bem_mesh.icosphere(sub, R), dict(kind="sphere", R=R)
    if family == "torus":
        r = rng.uniform(0.2, 0.4)
        n = int(rng.integers(20000, 40000)) if big else int(rng.integers(2000, 8000))
        return bem_mesh.torus_for_panels(1.0, r, n), dict(kind="torus", R=1.0, r=r)
    if family == "dimer":
        a = rng.uniform(0.4, 0.8)
        gap = rng.uniform(0.05, 0.5)
        return bem_mesh.sphere_dimer(a, gap, 4 if big else 3), dict(kind="dimer", a=a, gap=gap)
    L = rng.uniform(1.0, 2.0)
    n = int(rng.integers(60, 100)) if big else int(rng.integers(24, 48))
    return bem_mesh.square_plate(L, n), dict(kind="plate", L=L)


def outside(p, g, clear=0.1):
    k = g["kind"]
    if k == "sphere":
        return np.linalg.norm(p) > g["R"] + clear
    if k == "torus":
        rho = math.hypot(p[0], p[1])
        return math.hypot(rho - g["R"], p[2]) > g["r"] + clear
    if k == "dimer":
        d = g["a"] + 0.5 * g["gap"]
        return (np.linalg.norm(p - [d, 0, 0]) > g["a"] + clear and
                np.linalg.norm(p + [d, 0, 0]) > g["a"] + clear)
    half = 0.5 * g["L"]
    return abs(p[2]) > clear or abs(p[0]) > half + clear or abs(p[1]) > half + clear


def random_charges(rng, g, count=1):
    out = []
    ext = {"sphere": g.get("R", 1.0) * 2.0, "torus": 2.0, "dimer": 2.5, "plate": g.get("L", 1.0)}[g["kind"]]
    while len(out) < count:
        p = rng.uniform(-ext, ext, 3)
        if outside(p, g):
            out.append((float(rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 1.5)), tuple(p)))
    return out


def generate(n_solves, seed=0, rhs_per_geometry=48, panels_per_solve=512, tol=1e-7, degree=2,
             families=FAMILIES, log_every=10):
    """Returns (F, X, T, meta): features (s*m, 12), fixed columns (s*m, kx),
    normalised solutions (s*m,) float32 device tensors, m = panels_per_solve."""
    rng = np.random.default_rng(seed)
    feats, fixed, targs = [], [], []
    done = 0
    iters = []
    t0 = time.time()
    gi = 0
    m = panels_per_solve
    while done < n_solves:
        fam = families[gi % len(families)]
        gi += 1
        mesh, g = make_geometry(rng, fam)
        if mesh.n_panels < m:
            continue
        op = BemOperator(mesh, order=8, ws=1, leaf_size=64)
        k = min(rhs_per_geometry, n_solves - done)
        for j in range(k):
            if j == 0:
                comps = rhs_components(op, (), 1.0)
            else:
                comps = rhs_components(op, random_charges(rng, g, 1), 0.0)
            b = comps[0].contiguous()
            x, info = op.solve(b, tol=tol)
            if info["code"] != 0:
                continue
            iters.append(info["iters"])
            idx = torch.as_tensor(rng.choice(op.n, m, replace=False), device=op.device)
            feats.append(panel_features(op, b)[idx])
            fixed.append(fixed_columns(op, b, degree)[idx].to(torch.float32))
            xs = x[idx]
            targs.append((xs / xs.abs().max().clamp_min(1e-300)).to(torch.float32))
            done += 1
        if gi % log_every == 0:
            print(f"[train-data] {done} solves, {gi} geometries, {time.time() - t0:.1f}s, "
                  f"mean iters {np.mean(iters):.1f}", flush=True)
        del op
    meta = dict(solves=done, geometries=gi, mean_iters=float(np.mean(iters)),
                seconds=time.time() - t0, panels_per_solve=m)
    return torch.cat(feats), torch.cat(fixed), torch.cat(targs), meta


class ProjLoss:
    """Native projection loss over a batch of solves (``ed_proj_loss_f32``)."""

    def __init__(self, ridge=1e-6):
        self.ridge = ridge

    def __call__(self, Y, X, T, m):
        ns = T.shape[0] // m
        dY = torch.empty_like(Y)
        loss = torch.empty(ns, dtype=torch.float32, device=Y.device)
        nat.check(nat.lib().ed_proj_loss_f32(nat.ptr(Y), Y.shape[1], nat.ptr(X), X.shape[1],
                                             nat.ptr(T), ns, m, ctypes.c_float(self.ridge),
                                             nat.ptr(dY), nat.ptr(loss),
                                             nat.stream_ptr(Y.device)), "ed_proj_loss_f32")
        return loss, dY


def train(F, X, T, m, steps=4000, batch_solves=32, lr=1e-3, seed=0, log_every=250,
          dims=DEFAULT_DIMS, holdout=0.05):
    """Adam on the projection loss; returns (net, history, holdout loss)."""
    net = NativeMLP(dims, seed=seed)
    ns = T.shape[0] // m
    n_hold = max(1, int(ns * holdout))
    n_train = ns - n_hold
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    loss_fn = ProjLoss()
    rows = torch.arange(m, device=F.device)
    hist = []

    def batch_rows(sol):
        return (sol[:, None].to(F.device) * m + rows[None, :]).reshape(-1)

    def evaluate(sols):
        tot = []
        for s0 in range(0, sols.numel(), 256):
            r = batch_rows(sols[s0:s0 + 256])
            y = net.forward(F[r].contiguous())
            l, _ = loss_fn(y.contiguous(), X[r].contiguous(), T[r].contiguous(), m)
            tot.append(l)
        return float(torch.cat(tot).mean())

    hold = torch.arange(n_train, ns)
    for s in range(1, steps + 1):
        sols = torch.randint(0, n_train, (batch_solves,), generator=g)
        r = batch_rows(sols)
        y = net.forward(F[r].contiguous(), keep=True)
        loss, dY = loss_fn(y.contiguous(), X[r].contiguous(), T[r].contiguous(), m)
        dY.mul_(1.0 / batch_solves)
        net.backward(dY)
        cur = lr * 0.5 * (1.0 + math.cos(math.pi * s / steps))
        net.adam(lr=cur)
        if s % log_every == 0 or s == 1:
            hl = evaluate(hold)
            hist.append((s, float(loss.mean()), hl))
            print(f"[train] step {s} proj loss {float(loss.mean()):.3e} holdout {hl:.3e}",
                  flush=True)
    return net, hist, evaluate(hold)


def baseline_holdout(X, T, m, holdout=0.05):
    """Projection loss of the fixed columns alone (what the channels add to)."""
    ns = T.shape[0] // m
    n_train = ns - max(1, int(ns * holdout))
    r = torch.arange(n_train * m, ns * m, device=T.device)
    Xh, Th = X[r].contiguous(), T[r].contiguous()
    # a zero network channel leaves the fixed-column projection
    Y = torch.zeros((Th.shape[0], 1), dtype=torch.float32, device=T.device)
    loss, _ = ProjLoss()(Y, Xh, Th, m)
    return float(loss.mean())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--solves", type=int, default=10240)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    F, X, T, meta = generate(a.solves, seed=a.seed)
    print("[train-data]", meta, "rows", F.shape[0], flush=True)
    m = meta["panels_per_solve"]
    base = baseline_holdout(X, T, m)
    print(f"[train] fixed columns alone: holdout proj loss {base:.3e}", flush=True)
    net, hist, hl = train(F, X, T, m, steps=a.steps, seed=a.seed)
    lp = LearnedPreconditioner(net=net)
    if a.out:
        lp.save(a.out, solves=meta["solves"], holdout_proj_loss=hl, fixed_only_proj_loss=base,
                train_seconds=meta["seconds"])
        print("saved", a.out, flush=True)


if __name__ == "__main__":
    main()
