#!/usr/bin/env python3
"""Benchmark of the BEM/FMM solve hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): grounded
icosphere, subdivision 6 (81,920 flat panels, R = 1), point charge q = 1 at
r0 = rho * u with rho ~ U[1.2, 2.0] and u uniform on the sphere drawn from
numpy.random.default_rng(0); refined self integrals (depth 3) and near
quadrature correction (distance factor 1.5, depth 3); FMM in FP32 with FP32
MFMA M2L; FP64 GMRES (sparse approximate inverse smoother) with FP64
iterative refinement against the dense FP64 operator.

A step is one matvec of the solve operator (FMM far field + P2P + near
correction + self term) over the 81,920 panels with sigma resident in HBM.
``value`` = panels / s of that matvec.  The line also reports the time to a
1e-8 relative residual from the assembled geometry (set-up, first solve
with its workspace and graph capture; the set-up share is also reported
separately), the GMRES iterations of the plain (Jacobi) and the
learned-preconditioned solve at the same accuracy, per-stage times, the roofline of the dominant kernel and the CPU baseline (the oracle's
C restatement of the dense reference matvec, OpenMP, timed on a bounded
sample of rows on this host).
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3
MFMA_BF16_PEAK_TFLOPS = 2500.0
VALU_FP32_PEAK_TFLOPS = 157.3
# HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, the factors a 256 MiB device copy reads), keyed by
# (stage, panels, order, leaf, right-hand sides, M2L form), with the summary
# holding the commands; and the VALU-busy fraction of k_p2p_sym from SQ counters
PMC_SPHERE = "profiles/pmc_traffic_icosphere81920_p10_leaf256.md"
PMC_BATCH = "profiles/pmc_traffic_batch64_8rhs_p10_leaf512.md"
PMC_TORUS = "profiles/pmc_traffic_torus300192_p10_leaf256_bf16.md"
TRAFFIC_BYTES = {("p2p", 81920, 10, 256, 1, "fp16x3"): (2.129e7, PMC_SPHERE),
                 ("m2l", 81920, 10, 256, 1, "fp16x3"): (3.587e7, PMC_SPHERE),
                 ("m2l_reduce", 81920, 10, 256, 1, "fp16x3"): (2.823e7, PMC_SPHERE),
                 ("p2p", 877882, 10, 512, 8, "fp16x3"): (1.230e9, PMC_BATCH),
                 ("m2l", 877882, 10, 512, 8, "fp16x3"): (1.117e9, PMC_BATCH),
                 ("m2l_reduce", 877882, 10, 512, 8, "fp16x3"): (7.569e8, PMC_BATCH),
                 ("p2p", 300192, 10, 256, 1, "bf16"): (1.050e8, PMC_TORUS),
                 ("m2l", 300192, 10, 256, 1, "bf16"): (9.982e7, PMC_TORUS),
                 ("m2l_reduce", 300192, 10, 256, 1, "bf16"): (7.376e7, PMC_TORUS)}
P2P_VALU_BUSY = {(81920, 10, 256): 0.68}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--order", type=int, default=10)
    ap.add_argument("--ws", type=int, default=1)
    ap.add_argument("--leaf", type=int, default=128)
    ap.add_argument("--m2l", default="fp16x3",
                    help="M2L precision: fp16x3 (split FP16, FP32-level accuracy), fp32, bf16, fp16")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precond", default="sai", choices=["sai", "jacobi"])
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--workload", default="sphere", choices=["sphere", "batch", "torus", "pmc"],
                    help="sphere: 81,920-panel icosphere (default); batch: 64 geometries x 8 "
                         "RHS; torus: ~300k-panel torus with 16-bit M2L; pmc: a short run for "
                         "rocprofv3 --pmc (matvecs of the sphere workload, then a 256 MiB "
                         "device copy as the calibration kernel of known size)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --order/--leaf as given instead of the autotuner's choice")
    ap.add_argument("--tune-target", type=float, default=1e-5,
                    help="autotuner bound on the FMM's relative L2 error vs FP64 dense")
    ap.add_argument("--n-geoms", type=int, default=64)
    ap.add_argument("--rhs", type=int, default=8)
    ap.add_argument("--torus-panels", type=int, default=300312)
    ap.add_argument("--pmc-of", default="sphere", choices=["sphere", "batch", "torus"],
                    help="workload of --workload pmc")
    return ap.parse_args()


def stage_roofline(name, ms, st, n, order, m2l_peak=MFMA_FP32_PEAK_TFLOPS, R=1):
    """Algorithmic work of one launch of a stage for R right-hand sides: M2L
    2 (p+1)^4 flops per translation pair and RHS; leaf bytes below; other
    stages 16 B per panel."""
    nc = (order + 1) ** 2
    s = ms * 1e-3
    if name == "m2l":
        flops = 2.0 * nc * nc * st["n_m2l_pairs"] * R
        return dict(bound="mfma", achieved=flops / s / 1e12, peak=m2l_peak,
                    unit="TFLOP/s", algorithmic=flops)
    if name == "p2p":
        # target points + the source points of every P2P leaf pair (x, y, z,
        # q; with R > 1 the points and R charges) + the partial-sum writes;
        # one RHS evaluates r_ij and its rsqrt once per symmetric pair
        src = 16.0 if R == 1 else 16.0 + 4.0 * R
        sym = st.get("n_p2p_symmetric_interactions", 0) if R == 1 else 0
        parts = st.get("n_p2p_partials", n) if R == 1 else R * n
        evals = st["n_p2p_interactions"] - sym // 2
        byts = 16.0 * n + src * st["n_p2p_source_points"] + 4.0 * parts
        return dict(bound="hbm", achieved=byts / s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    algorithmic=byts, interactions_per_s=R * st["n_p2p_interactions"] / s,
                    pair_evaluations_per_s=evals / s,
                    # 18 flops of distance and 1/r per evaluated pair, shared
                    # by the R right-hand sides and both directions of a
                    # symmetric pair, + one FMA per interaction and RHS
                    valu_tflops=(18.0 * evals + 2.0 * R * st["n_p2p_interactions"]) / s / 1e12)
    if name == "l2p_out":
        # local coefficients once per leaf + near CSR (weight, column, gathered
        # sigma) + per panel: point, row pointer, perm and per RHS the P2P sum,
        # self term and out
        byts = (4.0 * nc * R * st["n_leaves"] + (8.0 + 4.0 * R) * st["n_near_pairs"]
                + (28.0 + 12.0 * R) * n)
        return dict(bound="hbm", achieved=byts / s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    algorithmic=byts)
    if name == "m2l_reduce":
        # the GEMM row of every M2L pair and its index, the node's own part
        # written
        byts = (4.0 * nc * R + 8.0) * st["n_m2l_pairs"] + 4.0 * nc * R * st["n_nodes"]
        return dict(bound="hbm", achieved=byts / s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    algorithmic=byts)
    byts = 16.0 * n * R
    return dict(bound="hbm", achieved=byts / s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                algorithmic=byts)


def host_cpu_info():
    """lscpu model, sockets, cores and threads of this host (the cores the CPU
    baseline runs on)."""
    import subprocess
    info = {"usable_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except Exception:
        out = ""
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus", "NUMA node(s)": "numa_nodes"}
    for ln in out.splitlines():
        k, _, v = ln.partition(":")
        if k.strip() in keys:
            info[keys[k.strip()]] = v.strip()
    return info


def cpu_baseline(mesh, sigma_np, self_corr_np, budget_s, repeats=3):
    """The oracle's C port of the dense FP64 reference matvec on this host's
    cores: a 64-row warm-up (which also sizes the sample), then ``repeats``
    timed passes over the same bounded row sample; value = median rate."""
    import statistics
    from oracle import cdense  # CPU baseline only
    n = mesh.n_panels
    threads = cdense.max_threads()
    t0 = time.perf_counter()
    cdense.dense_rows(mesh.centroids, mesh.areas, sigma_np, self_corr_np, 0, 64)
    t_probe = max(time.perf_counter() - t0, 1e-6)
    rows = int(min(n, max(64, 64 * budget_s / repeats / t_probe)))
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        cdense.dense_rows(mesh.centroids, mesh.areas, sigma_np, self_corr_np, 0, rows)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    host = host_cpu_info()
    return dict(value=rows / t, unit="panels/s", cores=threads, kind="port",
                sample=f"{rows} of {n} rows of the FP64 dense matvec (oracle/csrc/dense_ref.c, "
                       f"OpenMP {threads} threads), median of {repeats} timed passes after a "
                       f"64-row warm-up ({', '.join(f'{x:.2f}' for x in times)} s); full matvec "
                       f"extrapolated {n / (rows / t):.1f} s",
                host=host, repeats_s=times), n / (rows / t)


def cpu_reference_fmm(m2l_sample=48, subdiv=4, order=8, leaf=64, theta=0.5):
    """Second CPU-baseline leg: the reference's own FMM algorithm (oracle's
    numpy restatement: adaptive tree and dual-tree lists, every translation
    matrix built per pair by least squares at every matvec) at a small size,
    icosphere s = 4 (5,120 panels), the headline's p = 8 / leaf 64 / theta =
    0.5.  Bounded sample: every stage runs in full except M2L, of which the
    first ``m2l_sample`` pairs are built and applied and the rest
    extrapolated at their mean time; the dense leg at the same size is timed
    beside it (the faster of the two is the baseline at that size)."""
    import numpy as np
    from electrodrive_amd.core import bem_mesh
    from oracle import cdense, ref_fmm  # CPU baseline only
    try:
        from threadpoolctl import threadpool_info
        blas = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        blas = None
    mesh = bem_mesh.icosphere(subdiv)
    n = mesh.n_panels
    sig = np.random.default_rng(0).normal(size=n)
    st = {}
    t0 = time.perf_counter()
    ref_fmm.fmm_matvec(mesh.centroids, mesh.areas, sig, p=order, leaf_size=leaf, theta=theta,
                       stats=st, m2l_sample=m2l_sample)
    wall = time.perf_counter() - t0
    stages = st["stage_s"]
    t_full = wall - stages["m2l"] + st["m2l_s_per_pair"] * st["n_m2l_pairs"]
    t1 = time.perf_counter()
    cdense.dense_rows(mesh.centroids, mesh.areas, sig, np.zeros(n), 0, n)
    t_dense = time.perf_counter() - t1
    return dict(value=n / t_full, unit="panels/s", cores=blas, kind="port",
                panels=n, order=order, leaf=leaf, theta=theta,
                matvec_s_extrapolated=t_full, m2l_pairs=st["n_m2l_pairs"],
                m2l_s_per_pair=st["m2l_s_per_pair"], stage_s=stages,
                dense_leg_same_size=dict(value=n / t_dense, unit="panels/s",
                                         cores=cdense.max_threads(), matvec_s=t_dense),
                sample=f"restated reference FMM (oracle/ref_fmm.py, numpy/LAPACK) on {n} panels: "
                       f"tree, lists, P2M, M2M, L2L, L2P, P2P in full, {m2l_sample} of "
                       f"{st['n_m2l_pairs']} per-pair least-squares M2L builds timed "
                       f"({st['m2l_s_per_pair'] * 1e3:.0f} ms each) and the rest extrapolated; "
                       f"the dense leg at that size takes {t_dense:.3f} s, so it is the "
                       f"baseline there")


def quiesce_gc():
    """Collect, then move every live object out of the collector's
    generations: otherwise a full collection of the interpreter's
    (torch-sized) object graph lands inside a timed ~12 ms set-up in some runs
    (+30 ms), depending only on allocation counts since start-up."""
    gc.collect()
    gc.freeze()


def timed(fn):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    res = fn()
    torch.cuda.synchronize()
    return res, time.perf_counter() - t


def matvec_measure(op, sigma, a, n_units, nrhs=1):
    """Event-timed matvecs (``nrhs`` per step), stage times and roofline."""
    import torch
    st = op.plan.stats()
    sig = [sigma] if nrhs == 1 else list(sigma)
    out = torch.empty_like(sig[0])
    for _ in range(a.warmup):
        for sg in sig:
            op.plan.matvec(sg, op.self_corr, out=out)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(a.steps):
        for sg in sig:
            op.plan.matvec(sg, op.self_corr, out=out)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / a.steps
    multi = None
    if nrhs > 1:
        # the same right-hand sides in one multi-RHS matvec per step
        S = torch.stack(sig)
        outm = torch.empty_like(S)
        for _ in range(a.warmup):
            op.plan.matvec_multi(S, op.self_corr, out=outm)
        torch.cuda.synchronize()
        e2 = torch.cuda.Event(enable_timing=True)
        e3 = torch.cuda.Event(enable_timing=True)
        t1 = time.perf_counter()
        e2.record()
        for _ in range(a.steps):
            op.plan.matvec_multi(S, op.self_corr, out=outm)
        e3.record()
        torch.cuda.synchronize()
        ms_m = e2.elapsed_time(e3) / a.steps
        multi = dict(ms=ms_m, value=n_units / (ms_m * 1e-3), wall=time.perf_counter() - t1)
    # stage times of the launch the metric is quoted on (the multi-RHS matvec
    # when there are several right-hand sides)
    R = nrhs
    stages = (op.plan.profile(sig[0], op.self_corr, reps=max(5, a.steps // 2)) if R == 1 else
              op.plan.profile_multi(torch.stack(sig), op.self_corr, reps=max(5, a.steps // 2)))
    # P2P runs on a side stream beside the far-field chain (gather .. down),
    # the two branches meet before l2p_out and share the GPU, so the step is
    # neither their sum nor either one alone; the roofline is that of the
    # chain's largest kernel, the M2L GEMM (the upward stage is three
    # launches), and the P2P is rated against the VALU in roofline_extra
    chain = [k for k in stages if k != "p2p"]
    dom = "m2l" if "m2l" in stages else max(chain, key=stages.get)
    # 16-bit MFMA peak for the 16-bit forms; the split form executes three
    # FP16 products per FP32-equivalent one
    prec = op.plan.m2l_precision
    peak = MFMA_BF16_PEAK_TFLOPS if prec in ("bf16", "fp16", "fp16x3") else MFMA_FP32_PEAK_TFLOPS
    mfma_per_product = 3.0 if prec == "fp16x3" else 1.0
    rf = stage_roofline(dom, stages[dom], st, op.n, op.plan.order, peak, R)
    roofline = dict(kernel=dom, bound=rf["bound"], achieved=rf["achieved"], peak=rf["peak"],
                    unit=rf["unit"], frac=rf["achieved"] / rf["peak"], traffic=None,
                    launch_ms=stages[dom], algorithmic_per_launch=rf["algorithmic"], rhs=R)
    key = (op.n, op.plan.order, op.plan.leaf_size, R, prec)
    if (dom,) + key in TRAFFIC_BYTES:
        roofline["traffic"], roofline["traffic_source"] = TRAFFIC_BYTES[(dom,) + key]
    m2l = stage_roofline("m2l", stages["m2l"], st, op.n, op.plan.order, peak, R)
    p2p = stage_roofline("p2p", stages["p2p"], st, op.n, op.plan.order, R=R)
    red = stage_roofline("m2l_reduce", stages["m2l_reduce"], st, op.n, op.plan.order, R=R)
    extra = dict(m2l_mfma=dict(precision=prec, achieved_tflops=m2l["achieved"],
                               executed_mfma_tflops=m2l["achieved"] * mfma_per_product,
                               peak_tflops=peak,
                               frac=m2l["achieved"] * mfma_per_product / peak,
                               launch_ms=stages["m2l"]),
                 p2p=dict(achieved_gbs=p2p["achieved"], frac_hbm=p2p["achieved"] / HBM_PEAK_GBS,
                          interactions_per_s=p2p["interactions_per_s"],
                          valu_tflops=p2p["valu_tflops"],
                          frac_valu=p2p["valu_tflops"] / VALU_FP32_PEAK_TFLOPS,
                          valu_busy_pmc=P2P_VALU_BUSY.get(key[:3]) if R == 1 else None,
                          traffic=TRAFFIC_BYTES.get(("p2p",) + key, (None,))[0],
                          launch_ms=stages["p2p"]),
                 m2l_reduce=dict(achieved_gbs=red["achieved"], frac_hbm=red["achieved"] / HBM_PEAK_GBS,
                                 algorithmic_per_launch=red["algorithmic"],
                                 traffic=TRAFFIC_BYTES.get(("m2l_reduce",) + key, (None,))[0],
                                 launch_ms=stages["m2l_reduce"]))
    extra["m2l_mfma"]["traffic"] = TRAFFIC_BYTES.get(("m2l",) + key, (None,))[0]
    extra["critical_path"] = dict(chain_stages_ms=sum(stages[k] for k in chain),
                                  p2p_side_stream_ms=stages["p2p"],
                                  ms_per_matvec=ms if multi is None else multi["ms"],
                                  note="profile-mode stage times include their event pairs")
    return dict(ms=ms, wall=wall, stages=stages, roofline=roofline, extra=extra, stats=st,
                value=n_units / (ms * 1e-3), multi=multi)


# arithmetic of the matvec: FP32 P2P / expansions / quadrature apply, M2L on
# MFMA in the named form (FP32 accumulation), FP64 Krylov and refinement
DTYPE_LABEL = {"fp16x3": "fp32 (M2L split-fp16 MFMA)", "fp32": "fp32",
               "bf16": "fp32 (M2L bf16 MFMA)", "fp16": "fp32 (M2L fp16 MFMA)",
               "valu": "fp32"}


def matvec_accuracy(mesh, op, a, rng):
    """Relative L2 error of the FMM matvec (no near correction) against the FP64
    dense operator for the configuration measured and for the one the workload
    names (p = 8, leaf <= 64, theta = 0.5 -> ws = 2), with that one's time."""
    import numpy as np
    import torch
    from electrodrive_amd.core.bem_kernel import dense_matvec
    from electrodrive_amd.fmm3d.plan import FmmPlan
    sig = torch.as_tensor(np.random.default_rng(a.seed + 1).normal(size=op.n),
                          dtype=torch.float64, device=op.device)
    ref = dense_matvec(sig, op.C, op.A, op.self_corr)
    out = {}
    for name, (p, ws, leaf) in (("measured", (a.order, a.ws, a.leaf)),
                                ("workload_p8_leaf64_theta0.5", (8, 2, 64))):
        plan = FmmPlan(mesh.centroids, mesh.areas, leaf_size=leaf, order=p, ws=ws,
                       m2l_precision=a.m2l)
        v = plan.matvec(sig, op.self_corr)
        y = torch.empty_like(v)
        for _ in range(3):
            plan.matvec(sig, op.self_corr, out=y)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            plan.matvec(sig, op.self_corr, out=y)
        e1.record()
        torch.cuda.synchronize()
        out[name] = dict(order=p, ws=ws, leaf=leaf, m2l=a.m2l,
                         rel_err_vs_fp64_dense=float(torch.linalg.norm(v - ref)
                                                     / torch.linalg.norm(ref)),
                         ms_per_matvec=e0.elapsed_time(e1) / 20)
        del plan
    return out


def refine_record(op, b, info, solve_s, dense_s, precond, compare=True):
    """Refinement residuals of a solve: dense FP64 residuals and updated ones
    (r <- r - A_fmm d), the updated residual's bound, the dense residual's
    time and its share of the solve; with ``compare`` the same solve with
    every refinement residual dense beside it."""
    import torch
    bn = float(torch.linalg.norm(b))
    rec = dict(dense_residuals=info["n_accurate"], updated_residuals=info["n_updated"],
               update_operator_rel_err=info["update_err"],
               updated_resid_bound_rel=info["resid_bound"] / bn,
               dense_residual_s=dense_s,
               dense_residual_share=min(1.0, info["n_accurate"] * dense_s / max(solve_s, 1e-12)))
    if compare:
        op.solve(b, tol=1e-8, precond=precond, residual_update=False)  # warm
        (x0, i0), t0 = timed(lambda: op.solve(b, tol=1e-8, precond=precond,
                                              residual_update=False))
        rr0 = float(torch.linalg.norm(b - op.accurate(x0))) / bn
        rec["all_dense_refinement"] = dict(
            solve_s=t0, dense_residuals=i0["n_accurate"], refine_steps=i0["refine_steps"],
            iters=i0["iters"], code=i0["code"], rel_resid_fp64_dense=rr0,
            dense_residual_share=min(1.0, i0["n_accurate"] * dense_s / max(t0, 1e-12)))
    return rec


def base_line(a, value, ms, dtype, config):
    return {
        "metric": "time-to-1e-8 residual + FMM matvec panels/s; learned-precond vs plain GMRES",
        "value": value, "unit": "panels/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic", "config": config,
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (every box here has one "
              f"GPU and the path's multi-GPU sharding is design-only)", file=sys.stderr)
        return 2
    import torch
    torch.cuda.set_device(torch.device("cuda", 0))
    line = {"sphere": run_sphere, "batch": run_batch, "torus": run_torus,
            "pmc": run_pmc}[a.workload](a)
    print(json.dumps(line))
    return 0


def run_sphere(a):
    import numpy as np
    import torch
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator

    dev = torch.device("cuda", 0)
    mesh = bem_mesh.icosphere(a.subdiv)
    n = mesh.n_panels
    rng = np.random.default_rng(a.seed)
    rho = rng.uniform(1.2, 2.0)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    r0 = rho * u

    tune = None
    if not a.no_autotune:
        from electrodrive_amd.fmm3d import FmmConfig, autotune_for_problem
        base = FmmConfig(expansion_order=a.order, leaf_size=a.leaf,
                         mac_theta=0.8 if a.ws == 1 else 0.5, m2l_precision=a.m2l)
        tr, tune_s = timed(lambda: autotune_for_problem(
            base, n, dry_run=False, target_rel_error=a.tune_target, centroids=mesh.centroids,
            areas=mesh.areas, use_cache=False))
        a.order, a.leaf = tr.config.native_order, tr.config.leaf_size
        tune = dict(order=a.order, leaf=a.leaf, seconds=tune_s, target_rel_error=a.tune_target,
                    **{k: v for k, v in tr.metrics.items() if k in ("ms", "rel_err", "trials")})

    def build():
        return BemOperator(mesh, order=a.order, ws=a.ws, leaf_size=a.leaf, m2l_precision=a.m2l,
                           near=True, self_depth=3, near_depth=3)

    # first construction loads the code objects; the timed path starts from a
    # second one (plan, tree, quadrature and SAI set-up, solver workspace and
    # graph capture all included)
    op = build()
    b = op.rhs_point_charges([(1.0, tuple(r0))])
    solve = None
    if not a.no_solve:
        op.solve(b, tol=1e-8, precond=a.precond)
        del op
        quiesce_gc()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op = build()
        if a.precond == "sai":
            op.enable_sai()
        torch.cuda.synchronize()
        setup_s = time.perf_counter() - t0
        b = op.rhs_point_charges([(1.0, tuple(r0))])

        (x, info), first_solve_s = timed(lambda: op.solve(b, tol=1e-8, precond=a.precond))
        (x, info), solve_s = timed(lambda: op.solve(b, tol=1e-8, precond=a.precond))
        from oracle import analytic  # checker only
        sig_k = analytic.kelvin_sigma(mesh.centroids, 1.0, r0, 1.0)
        err_k = float(np.linalg.norm(x.cpu().numpy() - sig_k) / np.linalg.norm(sig_k))
        r, t_acc = timed(lambda: b - op.accurate(x))
        rel = float(torch.linalg.norm(r) / torch.linalg.norm(b))
        solve = dict(time_to_solution_s=setup_s + first_solve_s, setup_s=setup_s,
                     first_solve_s=first_solve_s, repeat_solve_s=solve_s, precond=a.precond,
                     gmres_iters=info["iters"], refine_steps=info["refine_steps"],
                     code=info["code"], rel_resid_fp64_dense=rel, err_vs_kelvin=err_k,
                     total_charge=op.total_charge(x), expected_charge=-1.0 / rho,
                     **refine_record(op, b, info, solve_s, t_acc, a.precond))
        if not a.no_learned:
            # learned-preconditioned vs plain GMRES at the same 1e-8 accuracy
            from electrodrive_amd.learn import precond as lp_mod
            lp = lp_mod.LearnedPreconditioner()
            ce = 1  # exact iteration counts

            def warm(fn):  # second of two calls: warm workspace and graph
                fn()
                return timed(fn)
            (xp, ip), tp = warm(lambda: op.solve(b, tol=1e-8, precond=True, check_every=ce))
            (xs, is_), ts = warm(lambda: op.solve(b, tol=1e-8, precond="sai", check_every=ce))
            (xf, if_), tf = warm(lambda: lp_mod.baseline_solve(op, b, tol=1e-8, precond="sai",
                                                               check_every=ce))
            (xl, il), tl = warm(lambda: lp.solve(op, charges=[(1.0, tuple(r0))], tol=1e-8,
                                                 precond="sai", check_every=ce))
            (xlj, ilj), tlj = warm(lambda: lp.solve(op, charges=[(1.0, tuple(r0))], tol=1e-8,
                                                    precond=True, check_every=ce))

            def leg(x_, i_, t_):
                rr = float(torch.linalg.norm(b - op.accurate(x_)) / torch.linalg.norm(b))
                return dict(iters=i_["iters"], seconds=t_, rel_resid_fp64_dense=rr,
                            diff_vs_plain=float(torch.linalg.norm(x_ - xp) / torch.linalg.norm(xp)))
            solve["learned_vs_plain"] = dict(
                plain_jacobi=leg(xp, ip, tp), sai=leg(xs, is_, ts),
                sai_fixed_coarse=leg(xf, if_, tf), sai_learned=leg(xl, il, tl),
                jacobi_learned=leg(xlj, ilj, tlj),
                iteration_ratio_plain_over_learned=ip["iters"] / max(1, il["iters"]),
                iteration_ratio_without_over_with_network=if_["iters"] / max(1, il["iters"]),
                weights="electrodrive_amd/learn/weights/precond_mlp.npz")
            pays = il["iters"] < if_["iters"] and tl <= ts
            solve["learned_vs_plain"]["network_pays_here"] = pays
            solve["learned_vs_plain"]["note"] = (
                "the network's coarse space beats the fixed one in iterations and SAI alone in "
                "wall time" if pays else
                "the iteration cut over plain Jacobi GMRES comes from the SAI smoother and the "
                "coarse space; the network (trained on 2k-8k-panel geometries) adds no "
                "iterations saved over the fixed coarse space and costs wall time at this size")
        sigma = x
    else:
        sigma = torch.as_tensor(rng.normal(size=n), dtype=torch.float64, device=dev)

    mm = matvec_measure(op, sigma, a, n)
    ms, stages, st = mm["ms"], mm["stages"], mm["stats"]
    accuracy = matvec_accuracy(mesh, op, a, rng)

    cpu = None
    cpu_full = None
    if not a.no_cpu:
        cpu, cpu_full = cpu_baseline(mesh, sigma.cpu().numpy(), op.self_corr.cpu().numpy(),
                                     a.cpu_seconds)
        cpu["reference_fmm_leg"] = cpu_reference_fmm()
        if solve is not None:
            # the reference method: Jacobi GMRES, one dense matvec per iteration
            lvp = solve.get("learned_vs_plain")
            its = lvp["plain_jacobi"]["iters"] if lvp else solve["gmres_iters"]
            cpu["time_to_solution_estimate_s"] = cpu_full * its
            cpu["time_to_solution_estimate"] = (f"{its} Jacobi GMRES iterations x the full "
                                                f"dense matvec time")

    line = base_line(a, mm["value"], ms, DTYPE_LABEL.get(a.m2l, "fp32"), {
        "workload": f"grounded icosphere s={a.subdiv} ({n} panels), point charge q=1 "
                    f"at rho={rho:.4f} (seed {a.seed}); FMM p={a.order} ws={a.ws} "
                    f"leaf<={a.leaf} M2L {a.m2l} (order and leaf from the autotuner at "
                    f"rel err <= {a.tune_target:g} vs FP64 dense); near quadrature on",
        "panels": n, "order": a.order, "ws": a.ws, "leaf": a.leaf,
        "m2l_precision": a.m2l, "krylov": "fp64 GMRES(50) + fp64 refinement"})
    line.update({"roofline": mm["roofline"], "roofline_extra": mm["extra"], "cpu_baseline": cpu,
                 "stage_ms": stages, "plan_stats": st, "solve": solve, "autotune": tune,
                 "matvec_accuracy": accuracy, "timed_wall_s": mm["wall"]})
    if solve is not None:
        # tuning is paid once per size bucket (cached on disk); both figures
        line["time_to_1e-8_s"] = solve["time_to_solution_s"]
        if tune is not None:
            line["time_to_1e-8_incl_autotune_s"] = solve["time_to_solution_s"] + tune["seconds"]
        if "learned_vs_plain" in solve:
            line["gmres_iters_plain_vs_learned"] = [
                solve["learned_vs_plain"]["plain_jacobi"]["iters"],
                solve["learned_vs_plain"]["sai_learned"]["iters"]]
    return line


def cpu_pair_rate(mesh, budget_s):
    """FP64 dense pair interactions per second of the oracle's C port on this
    host (bounded row sample)."""
    import numpy as np
    rng = np.random.default_rng(0)
    sig = rng.normal(size=mesh.n_panels)
    info, _ = cpu_baseline(mesh, sig, np.zeros(mesh.n_panels), budget_s)
    return info["value"] * mesh.n_panels, info


def batch_geometry(rng, family):
    """One geometry of the uneven batch: 2k-40k panels (log-uniform target)."""
    import numpy as np
    from electrodrive_amd.core import bem_mesh
    target = float(np.exp(rng.uniform(np.log(2000.0), np.log(40000.0))))
    if family == "sphere":
        R = rng.uniform(0.5, 2.0)
        return bem_mesh.icosphere(4 if target < 10000 else 5, R), dict(kind="sphere", R=R)
    if family == "torus":
        r = rng.uniform(0.2, 0.4)
        return bem_mesh.torus_for_panels(1.0, r, int(target)), dict(kind="torus", R=1.0, r=r)
    if family == "dimer":
        ra = rng.uniform(0.4, 0.8)
        gap = rng.uniform(0.05, 0.5)
        sub = 3 if target < 5000 else (4 if target < 20000 else 5)
        return bem_mesh.sphere_dimer(ra, gap, sub), dict(kind="dimer", a=ra, gap=gap)
    L = rng.uniform(1.0, 2.0)
    k = int(round((target / 2.0) ** 0.5))
    return bem_mesh.square_plate(L, k), dict(kind="plate", L=L)


def run_batch(a):
    """64 uneven geometries x 8 point-charge right-hand sides on one batched plan."""
    import numpy as np
    import torch
    from electrodrive_amd.core.operator import BemOperator
    from electrodrive_amd.learn.train import FAMILIES, random_charges

    rng = np.random.default_rng(a.seed)
    meshes, geos = [], []
    for g in range(a.n_geoms):
        m, gd = batch_geometry(rng, FAMILIES[g % 4])
        meshes.append(m)
        geos.append(gd)
    charges = [[random_charges(rng, gd, 1) for gd in geos] for _ in range(a.rhs)]
    sizes = np.array([m.n_panels for m in meshes])
    N = int(sizes.sum())

    tune = None
    if not a.no_autotune:
        # (order, leaf) of the fastest multi-RHS batched matvec within the
        # error bound on every geometry
        from electrodrive_amd.fmm3d import FmmConfig, autotune_for_problem
        base = FmmConfig(expansion_order=a.order, leaf_size=a.leaf,
                         mac_theta=0.8 if a.ws == 1 else 0.5, m2l_precision=a.m2l)
        off = np.concatenate([[0], np.cumsum(sizes)])
        tr, tune_s = timed(lambda: autotune_for_problem(
            base, N, dry_run=False, target_rel_error=a.tune_target,
            centroids=np.concatenate([m.centroids for m in meshes]),
            areas=np.concatenate([m.areas for m in meshes]), use_cache=False,
            geom_offsets=off, nrhs=a.rhs))
        a.order, a.leaf = tr.config.native_order, tr.config.leaf_size
        tune = dict(order=a.order, leaf=a.leaf, seconds=tune_s, target_rel_error=a.tune_target,
                    **{k: v for k, v in tr.metrics.items() if k in ("ms", "rel_err", "trials")})

    def build():
        op = BemOperator(meshes, order=a.order, ws=a.ws, leaf_size=a.leaf, m2l_precision=a.m2l,
                         near=True, self_depth=3, near_depth=3)
        if a.precond == "sai":
            op.enable_sai()
        return op

    def rhs(o):
        return torch.stack([o.rhs_point_charges(charges[r], per_geometry=True)
                            for r in range(a.rhs)])

    op = build()
    op.solve(rhs(op), tol=1e-8, precond=a.precond)
    del op
    quiesce_gc()
    # all R right-hand sides in one solve (one Krylov space per (RHS,
    # geometry) block, one multi-RHS matvec per Arnoldi step)
    op, setup_s = timed(build)
    B = rhs(op)
    (X, minfo), t_solve = timed(lambda: op.solve(B, tol=1e-8, precond=a.precond))
    multi_solve = dict(iters=minfo["iters"], code=minfo["code"], seconds=t_solve,
                       max_geometry_rel_resid=minfo.get("max_geometry_rel_resid"),
                       refine_steps=minfo["refine_steps"])
    # the same right-hand sides one solve at a time, for comparison
    solves = []
    t_one = 0.0
    for r in range(a.rhs):
        (x, info), t = timed(lambda: op.solve(B[r], tol=1e-8, precond=a.precond))
        t_one += t
        solves.append(dict(iters=info["iters"], code=info["code"], seconds=t,
                           max_geometry_rel_resid=info.get("max_geometry_rel_resid"),
                           diff_vs_multi=float(torch.linalg.norm(x - X[r])
                                               / torch.linalg.norm(x))))
    # FMM operator against the FP64 dense operator, per geometry, random sigma
    sig = torch.as_tensor(rng.normal(size=N), dtype=torch.float64, device=op.device)
    fa, fd = op.approx(sig), op.accurate(sig)
    errs = []
    for g in range(len(meshes)):
        lo, hi = int(op.geom_offsets[g]), int(op.geom_offsets[g + 1])
        errs.append(float(torch.linalg.norm(fa[lo:hi] - fd[lo:hi]) / torch.linalg.norm(fd[lo:hi])))
    sigmas = [torch.as_tensor(rng.normal(size=N), dtype=torch.float64, device=op.device)
              for _ in range(a.rhs)]
    mm = matvec_measure(op, sigmas, a, N * a.rhs, nrhs=a.rhs)
    cpu = None
    if not a.no_cpu:
        rate, info = cpu_pair_rate(meshes[int(np.argmax(sizes))], a.cpu_seconds)
        pairs = float((sizes.astype(np.float64) ** 2).sum())
        cpu = dict(value=N / (pairs / rate), unit="panels/s", cores=info["cores"], kind="port",
                   host=info["host"],
                   sample=f"dense FP64 pair rate {rate:.3e}/s of the oracle's C port measured on "
                          f"the largest geometry ({info['sample']}); batch matvec = sum N_g^2 "
                          f"pairs")
    # metric: sum_g N_g R / t for one matvec of all R right-hand sides (the
    # multi-RHS matvec); the R one-RHS matvecs in a row are reported beside it
    mv = mm["multi"] if mm["multi"] is not None else dict(ms=mm["ms"], value=mm["value"],
                                                          wall=mm["wall"])
    line = base_line(a, mv["value"], mv["ms"], DTYPE_LABEL.get(a.m2l, "fp32"), {
        "workload": f"uneven batch: {a.n_geoms} seeded geometries (spheres, tori, sphere dimers, "
                    f"plates; {int(sizes.min())}-{int(sizes.max())} panels, {N} total) x {a.rhs} "
                    f"point-charge RHS; one batched launch per FMM stage covering all RHS "
                    f"(multi-RHS matvec); FMM p={a.order} "
                    f"ws={a.ws} leaf<={a.leaf} M2L {a.m2l}"
                    + (" (order and leaf from the autotuner at rel err <= "
                       f"{a.tune_target:g} vs FP64 dense on every geometry)" if tune else "")
                    + f"; GMRES {a.precond} + fp64 refinement",
        "geometries": a.n_geoms, "rhs": a.rhs, "panels_total": N, "order": a.order,
        "leaf": a.leaf, "m2l_precision": a.m2l})
    line.update({"roofline": mm["roofline"], "roofline_extra": mm["extra"], "cpu_baseline": cpu,
                 "stage_ms": mm["stages"], "plan_stats": mm["stats"], "autotune": tune,
                 "fmm_vs_fp64_dense_max_rel_err": max(errs),
                 "solve": dict(time_to_solution_s=setup_s + t_solve, setup_s=setup_s,
                               solve_all_rhs_s=t_solve, all_rhs=multi_solve,
                               one_rhs_at_a_time_s=t_one, per_rhs=solves),
                 "time_to_1e-8_s": setup_s + t_solve,
                 "time_to_1e-8_incl_autotune_s": setup_s + t_solve + (tune["seconds"] if tune
                                                                      else 0.0),
                 "timed_wall_s": mv["wall"],
                 "rhs_one_at_a_time": dict(ms_per_step=mm["ms"], value=mm["value"])})
    return line


def torus_charge(rng):
    import numpy as np
    while True:
        p = rng.uniform(-2.0, 2.0, 3)
        rho = np.hypot(p[0], p[1])
        if np.hypot(rho - 1.0, p[2]) > 0.3 + 0.05:
            return [(1.0, tuple(p))]


def run_torus(a):
    """~300k-panel torus, M2L in BF16 and FP16 (FP32 accumulation) with FP64
    iterative refinement to 1e-8; errors against the FP32-M2L solve and the
    FP64 dense operator."""
    import numpy as np
    import torch
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator

    rng = np.random.default_rng(a.seed)
    mesh = bem_mesh.torus_for_panels(1.0, 0.3, a.torus_panels)
    n = mesh.n_panels
    ch = torus_charge(rng)
    sig = torch.as_tensor(rng.normal(size=n), dtype=torch.float64, device="cuda")
    tune = None
    if not a.no_autotune:
        # one (order, leaf) for all four M2L forms: the fastest split-FP16
        # matvec within the error bound (the 16-bit forms run on the same tree)
        from electrodrive_amd.fmm3d import FmmConfig, autotune_for_problem
        base = FmmConfig(expansion_order=a.order, leaf_size=a.leaf,
                         mac_theta=0.8 if a.ws == 1 else 0.5, m2l_precision="fp16x3")
        tr, tune_s = timed(lambda: autotune_for_problem(
            base, n, dry_run=False, target_rel_error=a.tune_target, centroids=mesh.centroids,
            areas=mesh.areas, use_cache=False))
        a.order, a.leaf = tr.config.native_order, tr.config.leaf_size
        tune = dict(order=a.order, leaf=a.leaf, seconds=tune_s, target_rel_error=a.tune_target,
                    **{k: v for k, v in tr.metrics.items() if k in ("ms", "rel_err", "trials")})
    res = {}
    ref_x = None
    dense = None
    for prec in ("fp32", "fp16x3", "bf16", "fp16"):
        def build():
            op = BemOperator(mesh, order=a.order, ws=a.ws, leaf_size=a.leaf, m2l_precision=prec,
                             near=True, self_depth=3, near_depth=3)
            op.enable_sai()
            return op
        quiesce_gc()
        op, setup_s = timed(build)
        b = op.rhs_point_charges(ch)
        (x, info), t = timed(lambda: op.solve(b, tol=1e-8, precond="sai"))
        if dense is None:
            dense = op.accurate(sig)
        fa = op.approx(sig)
        rv, t_acc = timed(lambda: b - op.accurate(x))
        rr = float(torch.linalg.norm(rv) / torch.linalg.norm(b))
        if ref_x is None:
            ref_x = x
        mm = matvec_measure(op, sig, a, n)
        res[prec] = dict(setup_s=setup_s, solve_s=t, time_to_solution_s=setup_s + t,
                         iters=info["iters"], refine_steps=info["refine_steps"], code=info["code"],
                         rel_resid_fp64_dense=rr,
                         solution_diff_vs_fp32_m2l=float(torch.linalg.norm(x - ref_x)
                                                          / torch.linalg.norm(ref_x)),
                         matvec_rel_err_vs_fp64_dense=float(torch.linalg.norm(fa - dense)
                                                            / torch.linalg.norm(dense)),
                         ms_per_matvec=mm["ms"], panels_per_s=mm["value"], stage_ms=mm["stages"],
                         m2l=mm["extra"]["m2l_mfma"],
                         **refine_record(op, b, info, t, t_acc, "sai",
                                         compare=prec in ("bf16", "fp16")))
        if prec == "bf16":
            head = mm
        del op
    cpu = None
    if not a.no_cpu:
        rate, info = cpu_pair_rate(mesh, a.cpu_seconds)
        cpu = dict(value=n / (float(n) ** 2 / rate), unit="panels/s", cores=info["cores"],
                   host=info["host"],
                   kind="port", sample=info["sample"])
    line = base_line(a, head["value"], head["ms"], DTYPE_LABEL["bf16"], {
        "workload": f"torus R=1 r=0.3 ({n} panels), point charge ~U[-2,2]^3 outside the "
                    f"surface (seed {a.seed}); FMM p={a.order} ws={a.ws} leaf<={a.leaf}"
                    + (" (order and leaf from the autotuner at rel err <= "
                       f"{a.tune_target:g} with split-FP16 M2L)" if tune else "") + "; M2L "
                    f"BF16/FP16 MFMA (FP32 accumulation), FP64 GMRES (SAI) + FP64 refinement",
        "panels": n, "order": a.order, "leaf": a.leaf, "m2l_precision": "bf16"})
    line.update({"roofline": head["roofline"], "roofline_extra": head["extra"],
                 "cpu_baseline": cpu, "stage_ms": head["stages"], "plan_stats": head["stats"],
                 "precisions": res, "time_to_1e-8_s": res["bf16"]["time_to_solution_s"],
                 "time_to_1e-8_incl_autotune_s": res["bf16"]["time_to_solution_s"]
                 + (tune["seconds"] if tune else 0.0),
                 "autotune": tune,
                 "timed_wall_s": head["wall"]})
    return line


def run_pmc(a):
    """Counter run: --steps matvecs of one workload at the given order / leaf
    (--pmc-of sphere: the 81,920-panel icosphere; batch: the 64-geometry
    batch, one 8-RHS matvec per step; torus: the ~300k-panel torus), then
    three copies of a 256 MiB device buffer (the calibration kernel of known
    size); nothing else runs on the GPU in between."""
    import numpy as np
    import torch
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core.operator import BemOperator
    rng = np.random.default_rng(a.seed)
    if a.pmc_of == "batch":
        from electrodrive_amd.learn.train import FAMILIES
        meshes = [batch_geometry(rng, FAMILIES[g % 4])[0] for g in range(a.n_geoms)]
        op = BemOperator(meshes, order=a.order, ws=a.ws, leaf_size=a.leaf, m2l_precision=a.m2l,
                         near=True, self_depth=3, near_depth=3)
        S = torch.as_tensor(rng.normal(size=(a.rhs, op.n)), dtype=torch.float64, device="cuda")
        out = torch.empty_like(S)
        step = lambda: op.plan.matvec_multi(S, op.self_corr, out=out)  # noqa: E731
    else:
        mesh = (bem_mesh.torus_for_panels(1.0, 0.3, a.torus_panels) if a.pmc_of == "torus"
                else bem_mesh.icosphere(a.subdiv))
        op = BemOperator(mesh, order=a.order, ws=a.ws, leaf_size=a.leaf, m2l_precision=a.m2l,
                         near=True, self_depth=3, near_depth=3)
        sig = torch.as_tensor(rng.normal(size=op.n), dtype=torch.float64, device="cuda")
        out = torch.empty_like(sig)
        step = lambda: op.plan.matvec(sig, op.self_corr, out=out)  # noqa: E731
    (_, t_mv) = timed(lambda: [step() for _ in range(a.steps)])
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    (_, t_cp) = timed(lambda: [dst.copy_(src) for _ in range(3)])
    return {"workload": "pmc", "of": a.pmc_of, "panels": op.n, "order": a.order, "leaf": a.leaf,
            "m2l": a.m2l, "stats": op.plan.stats(), "matvecs": a.steps, "matvec_s": t_mv,
            "copy_bytes": 2 * src.numel() * 4, "copies": 3, "copy_s": t_cp}


if __name__ == "__main__":
    sys.exit(main())
