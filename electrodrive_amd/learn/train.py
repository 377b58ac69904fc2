"""Training of the learned preconditioner on solves produced by the solver.

Geometry families follow the batched workload: spheres (R ~ U[0.5, 2]),
tori (R = 1, r ~ U[0.2, 0.4]), sphere dimers (gap ~ U[0.05, 0.5]) and square
plates, with point charges placed outside the surface (clearance >= 0.1).  Each
geometry is solved for many right-hand sides with the native FP64 solver; for
every solve the features of (up to) ``panels_per_solve`` random panels and the
normalised solution sigma / max|b/diag| are kept.  Training runs Adam on the
native MLP kernels.

``python -m electrodrive_amd.learn.train --solves 10240 --out <path.npz>``
"""
import argparse
import math
import time

import numpy as np
import torch

from ..core import bem_mesh
from ..core.operator import BemOperator
from .mlp import NativeMLP
from .precond import N_FEATURES, LearnedPreconditioner, jacobi_scale, panel_features

FAMILIES = ("sphere", "torus", "dimer", "plate")


def make_geometry(rng, family, size="small"):
    big = size == "large"
    if family == "sphere":
        R = rng.uniform(0.5, 2.0)
        return bem_mesh.icosphere(5 if big else int(rng.integers(3, 5)), R), dict(kind="sphere", R=R)
    if family == "torus":
        r = rng.uniform(0.2, 0.4)
        n = int(rng.integers(20000, 40000)) if big else int(rng.integers(2000, 6000))
        return bem_mesh.torus_for_panels(1.0, r, n), dict(kind="torus", R=1.0, r=r)
    if family == "dimer":
        a = rng.uniform(0.4, 0.8)
        gap = rng.uniform(0.05, 0.5)
        return bem_mesh.sphere_dimer(a, gap, 4 if big else 3), dict(kind="dimer", a=a, gap=gap)
    L = rng.uniform(1.0, 2.0)
    n = int(rng.integers(60, 100)) if big else int(rng.integers(20, 40))
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


def generate(n_solves, seed=0, rhs_per_geometry=64, panels_per_solve=256, tol=1e-7,
             families=FAMILIES, log_every=20):
    """Returns (F, T, meta): features (m, 12) float32, targets (m,) float32."""
    rng = np.random.default_rng(seed)
    feats, targs = [], []
    done = 0
    iters = []
    t0 = time.time()
    gi = 0
    while done < n_solves:
        fam = families[gi % len(families)]
        gi += 1
        mesh, g = make_geometry(rng, fam)
        op = BemOperator(mesh, order=8, ws=1, leaf_size=64)
        for _ in range(min(rhs_per_geometry, n_solves - done)):
            ch = random_charges(rng, g, count=int(rng.integers(1, 3)))
            b = op.rhs_point_charges(ch)
            x, info = op.solve(b, tol=tol)
            if info["code"] != 0:
                continue
            iters.append(info["iters"])
            F = panel_features(op, b)
            T = (x / jacobi_scale(op, b)).to(torch.float32)
            k = min(panels_per_solve, op.n)
            idx = torch.as_tensor(rng.choice(op.n, k, replace=False), device=F.device)
            feats.append(F[idx])
            targs.append(T[idx])
            done += 1
        if gi % log_every == 0:
            print(f"[train-data] {done} solves, {time.time() - t0:.1f}s, mean iters "
                  f"{np.mean(iters):.1f}", flush=True)
    return torch.cat(feats), torch.cat(targs), dict(solves=done, geometries=gi,
                                                     mean_iters=float(np.mean(iters)),
                                                     seconds=time.time() - t0)


def train(F, T, steps=3000, batch=8192, lr=2e-3, seed=0, log_every=500):
    net = NativeMLP((N_FEATURES, 256, 256, 256, 1), seed=seed)
    m = F.shape[0]
    g = torch.Generator(device=F.device)
    g.manual_seed(seed)
    hist = []
    for s in range(1, steps + 1):
        idx = torch.randint(0, m, (batch,), device=F.device, generator=g)
        xb, tb = F[idx].contiguous(), T[idx].contiguous()
        y = net.forward(xb, keep=True)
        loss, grad = net.wmse(y, tb)
        grad.mul_(1.0 / batch)
        net.backward(grad.view(batch, 1))
        cur = lr * 0.5 * (1.0 + math.cos(math.pi * s / steps))
        net.adam(lr=cur)
        if s % log_every == 0 or s == 1:
            hist.append((s, loss / batch))
            print(f"[train] step {s} mse {loss / batch:.3e}", flush=True)
    return net, hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--solves", type=int, default=10240)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    F, T, meta = generate(a.solves, seed=a.seed)
    print("[train-data]", meta, "samples", F.shape[0], flush=True)
    net, hist = train(F, T, steps=a.steps, seed=a.seed)
    lp = LearnedPreconditioner(net=net)
    if a.out:
        lp.save(a.out)
        print("saved", a.out)


if __name__ == "__main__":
    main()
