#!/usr/bin/env python3
"""Benchmark of the BEM/FMM solve hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): grounded
icosphere, subdivision 6 (81,920 flat panels, R = 1), point charge q = 1 at
r0 = rho * u with rho ~ U[1.2, 2.0] and u uniform on the sphere drawn from
numpy.random.default_rng(0); refined self integrals (depth 3) and near
quadrature correction (distance factor 1.5, depth 3); FMM in FP32 with FP32
MFMA M2L; FP64 GMRES with FP64 iterative refinement against the dense FP64
operator.

A step is one matvec of the solve operator (FMM far field + P2P + near
correction + self term) over the 81,920 panels with sigma resident in HBM.
``value`` = panels / s of that matvec.  The line also reports the time to a
1e-8 relative residual from the assembled geometry (set-up, first solve
with its workspace and graph capture; the set-up share is also reported
separately), per-stage
times, the roofline of the dominant kernel and the CPU baseline (the oracle's
C restatement of the dense reference matvec, OpenMP, timed on a bounded
sample of rows on this host).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3
MFMA_BF16_PEAK_TFLOPS = 2500.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--order", type=int, default=10)
    ap.add_argument("--ws", type=int, default=1)
    ap.add_argument("--leaf", type=int, default=128)
    ap.add_argument("--m2l", default="fp32")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def stage_roofline(name, ms, st, n, order):
    """Algorithmic work of one launch of a stage (documented in DESIGN.md)."""
    nc = (order + 1) ** 2
    s = ms * 1e-3
    if name == "m2l":
        flops = 2.0 * nc * nc * st["n_m2l_pairs"]
        return dict(bound="mfma", achieved=flops / s / 1e12, peak=MFMA_FP32_PEAK_TFLOPS,
                    unit="TFLOP/s", algorithmic=flops)
    if name == "leaf":
        # targets (x,y,z,q) + source tiles of every P2P leaf pair + local
        # coefficients + near CSR (col int32 + weight fp32) + sigma/self/perm/out
        byts = (16.0 * n + 16.0 * st["n_p2p_source_points"] + 4.0 * nc * st["n_leaves"]
                + 8.0 * st["n_near_pairs"] + 24.0 * n)
        return dict(bound="hbm", achieved=byts / s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    algorithmic=byts)
    if name == "m2l_reduce":
        byts = 4.0 * 96 * st["n_m2l_pairs"] + 8.0 * st["n_m2l_pairs"] + 4.0 * nc * st["n_nodes"]
        return dict(bound="hbm", achieved=byts / s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    algorithmic=byts)
    byts = 16.0 * n
    return dict(bound="hbm", achieved=byts / s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                algorithmic=byts)


def cpu_baseline(mesh, sigma_np, self_corr_np, budget_s):
    from oracle import cdense  # CPU baseline only
    n = mesh.n_panels
    threads = cdense.max_threads()
    t0 = time.perf_counter()
    cdense.dense_rows(mesh.centroids, mesh.areas, sigma_np, self_corr_np, 0, 64)
    t_probe = max(time.perf_counter() - t0, 1e-6)
    rows = int(min(n, max(64, 64 * budget_s / t_probe)))
    t0 = time.perf_counter()
    cdense.dense_rows(mesh.centroids, mesh.areas, sigma_np, self_corr_np, 0, rows)
    t = time.perf_counter() - t0
    return dict(value=rows / t, unit="panels/s", cores=threads, kind="port",
                sample=f"{rows} of {n} rows of the FP64 dense matvec (oracle/csrc/dense_ref.c, "
                       f"OpenMP {threads} threads), {t:.1f} s; full matvec extrapolated "
                       f"{n / (rows / t):.1f} s"), n / (rows / t)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (every box here has one "
              f"GPU and the path's multi-GPU sharding is design-only)", file=sys.stderr)
        return 2
    import numpy as np
    import torch
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    mesh = bem_mesh.icosphere(a.subdiv)
    n = mesh.n_panels
    rng = np.random.default_rng(a.seed)
    rho = rng.uniform(1.2, 2.0)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    r0 = rho * u

    def build():
        return BemOperator(mesh, order=a.order, ws=a.ws, leaf_size=a.leaf, m2l_precision=a.m2l,
                           near=True, self_depth=3, near_depth=3)

    # first construction loads the code objects; the timed path starts from a
    # second one (plan, tree, quadrature set-up, solver workspace and graph
    # capture all included)
    op = build()
    b = op.rhs_point_charges([(1.0, tuple(r0))])
    solve = None
    if not a.no_solve:
        op.solve(b, tol=1e-8)
        del op
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op = build()
        torch.cuda.synchronize()
        setup_s = time.perf_counter() - t0
        b = op.rhs_point_charges([(1.0, tuple(r0))])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x, info = op.solve(b, tol=1e-8)
        torch.cuda.synchronize()
        first_solve_s = time.perf_counter() - t1
        t1 = time.perf_counter()
        x, info = op.solve(b, tol=1e-8)
        torch.cuda.synchronize()
        solve_s = time.perf_counter() - t1
        from oracle import analytic  # checker only
        sig_k = analytic.kelvin_sigma(mesh.centroids, 1.0, r0, 1.0)
        err_k = float(np.linalg.norm(x.cpu().numpy() - sig_k) / np.linalg.norm(sig_k))
        r = b - op.accurate(x)
        rel = float(torch.linalg.norm(r) / torch.linalg.norm(b))
        solve = dict(time_to_solution_s=setup_s + first_solve_s, setup_s=setup_s,
                     first_solve_s=first_solve_s, repeat_solve_s=solve_s,
                     gmres_iters=info["iters"], refine_steps=info["refine_steps"],
                     code=info["code"], rel_resid_fp64_dense=rel, err_vs_kelvin=err_k,
                     total_charge=op.total_charge(x), expected_charge=-1.0 / rho)
        sigma = x
    else:
        sigma = torch.as_tensor(rng.normal(size=n), dtype=torch.float64, device=dev)

    st = op.plan.stats()
    out = torch.empty_like(sigma)
    for _ in range(a.warmup):
        op.plan.matvec(sigma, op.self_corr, out=out)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(a.steps):
        op.plan.matvec(sigma, op.self_corr, out=out)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / a.steps
    stages = op.plan.profile(sigma, op.self_corr, reps=max(5, a.steps // 2))
    dom = max(stages, key=stages.get)
    rf = stage_roofline(dom, stages[dom], st, n, a.order)
    roofline = dict(kernel=dom, bound=rf["bound"], achieved=rf["achieved"], peak=rf["peak"],
                    unit=rf["unit"], frac=rf["achieved"] / rf["peak"], traffic=None,
                    launch_ms=stages[dom], algorithmic_per_launch=rf["algorithmic"])
    m2l = stage_roofline("m2l", stages["m2l"], st, n, a.order)
    extra_roofline = dict(m2l_mfma=dict(achieved_tflops=m2l["achieved"],
                                        frac=m2l["achieved"] / MFMA_FP32_PEAK_TFLOPS,
                                        launch_ms=stages["m2l"]))

    cpu = None
    cpu_full = None
    if not a.no_cpu:
        cpu, cpu_full = cpu_baseline(mesh, sigma.cpu().numpy(), op.self_corr.cpu().numpy(),
                                     a.cpu_seconds)
        if solve is not None:
            cpu["time_to_solution_estimate_s"] = cpu_full * solve["gmres_iters"]

    value = n / (ms * 1e-3)
    line = {
        "metric": "time-to-1e-8 residual + FMM matvec panels/s; learned-precond vs plain GMRES",
        "value": value,
        "unit": "panels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"grounded icosphere s={a.subdiv} ({n} panels), point charge q=1 "
                               f"at rho={rho:.4f} (seed {a.seed}); FMM p={a.order} ws={a.ws} "
                               f"leaf<={a.leaf} M2L {a.m2l}; near quadrature on",
                   "panels": n, "order": a.order, "ws": a.ws, "leaf": a.leaf,
                   "m2l_precision": a.m2l, "krylov": "fp64 GMRES(50) + fp64 refinement"},
        "roofline": roofline,
        "roofline_extra": extra_roofline,
        "cpu_baseline": cpu,
        "stage_ms": stages,
        "plan_stats": st,
        "solve": solve,
        "timed_wall_s": wall,
    }
    if solve is not None:
        line["time_to_1e-8_s"] = solve["time_to_solution_s"]
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
