"""Leaf size / expansion order autotuner (mirror of fmm3d/autotune.py:66-230).

Same public names and arguments as the reference: ``autotune_for_problem(
base_cfg, problem_size, *, dry_run=True, target_rel_error=1e-2,
max_trials=24, device=None, use_cache=True) -> TuningResult(config,
metrics)`` with a JSON cache keyed by a hardware fingerprint and a problem-size
bucket (fmm3d/autotune.py:670-830).  What differs is what is measured: every
candidate is a native plan on the device, timed with HIP events, and its error
is the relative L2 difference to the FP64 dense operator on the same panels
(the reference times its CPU pipeline and compares with a direct sum).

Candidates: expansion order in ``SUPPORTED_ORDERS`` at or above 6 and leaf
sizes 32 .. 512, with ``ws`` and the M2L precision of ``base_cfg``.  The fastest
candidate whose error meets ``target_rel_error`` wins; when none does, the most
accurate one.  Not every candidate is measured: the error is set by the order
(the leaf size moves it by a few percent), so an order whose first candidate
misses the target twice over is dropped, the search stops at an order whose
first candidate is twice as slow as the best one that met the target, and
for the other orders the leaf size walks from 256 towards the faster
neighbour while the matvec gets faster (its time is unimodal in the leaf
size).  On the 82k sphere, the 300k torus and the 64-geometry batch about 10
of the 24 candidates are measured, and the small-leaf plans, the slowest to
build and to run, are skipped.  Pass ``centroids`` / ``areas`` to tune on the actual geometry
(otherwise a unit sphere of ``problem_size`` panels is used).

The cache file (``$ELECTRODRIVE_AMD_AUTOTUNE_CACHE`` or
``~/.cache/electrodrive_amd/fmm3d_autotune.json``) is the only shared mutable
state; writers hold an exclusive ``flock`` on it.
"""
import dataclasses
import fcntl
import json
import math
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Dict, List, Optional

import numpy as np
import torch

from .. import _native as nat
from .config import SUPPORTED_ORDERS, FmmConfig

AUTOTUNE_CACHE_ENV_VAR = "ELECTRODRIVE_AMD_AUTOTUNE_CACHE"
AUTOTUNE_CACHE_FILENAME = "fmm3d_autotune.json"
LEAF_SIZES = (32, 64, 128, 256, 384, 512)


@dataclass
class TuningResult:
    config: FmmConfig
    metrics: Dict[str, float]


def _bucket_problem_size(n: int) -> int:
    """Power-of-two bucket (fmm3d/autotune.py:670-674)."""
    return 1 << max(0, int(math.ceil(math.log2(max(1, n)))))


_BUILD_ID = None


def _build_id() -> str:
    """Content hash of the loaded native library: a rebuilt kernel set gets
    fresh cache entries."""
    global _BUILD_ID
    if _BUILD_ID is None:
        import hashlib
        with open(nat.LIB_PATH, "rb") as f:
            _BUILD_ID = hashlib.sha1(f.read()).hexdigest()[:12]
    return _BUILD_ID


def _hardware_fingerprint(device=None) -> str:
    dev = torch.device(device or "cuda")
    props = torch.cuda.get_device_properties(dev)
    arch = getattr(props, "gcnArchName", "") or ""
    return (f"{props.name}|{arch}|cu{props.multi_processor_count}|lib{nat.lib().ed_abi_version()}"
            f"|build{_build_id()}")


def _get_cache_path() -> Path:
    env = os.environ.get(AUTOTUNE_CACHE_ENV_VAR)
    if env:
        return Path(env)
    return Path.home() / ".cache" / "electrodrive_amd" / AUTOTUNE_CACHE_FILENAME


def _cache_key(fp, bucket, cfg, target):
    return f"{fp}|n{bucket}|ws{cfg.ws}|{cfg.m2l_precision}|err{target:.1e}"


def _load_best_from_cache(key: str) -> Optional[Dict[str, Any]]:
    path = _get_cache_path()
    if not path.exists():
        return None
    try:
        with open(path, "r") as f:
            fcntl.flock(f, fcntl.LOCK_SH)
            data = json.load(f)
    except (OSError, ValueError):
        return None
    return data.get(key)


def _append_record_to_cache(key: str, record: Dict[str, Any]) -> None:
    path = _get_cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        with open(path, "a+") as f:
            fcntl.flock(f, fcntl.LOCK_EX)
            f.seek(0)
            txt = f.read()
            try:
                data = json.loads(txt) if txt.strip() else {}
            except ValueError:
                data = {}
            data[key] = record
            f.seek(0)
            f.truncate()
            json.dump(data, f, indent=1, sort_keys=True)
    except OSError:
        pass  # best effort, as the reference (fmm3d/autotune.py:767-803)


def _synthetic_geometry(n):
    """Unit sphere with about n panels (icosphere of the nearest size)."""
    from ..core import bem_mesh
    s = max(1, int(round(math.log(max(n, 20) / 20.0, 4))))
    m = bem_mesh.icosphere(s)
    return m.centroids, m.areas


def _measure(cfg, C, A, sigma, ref, reps, dev, geom_offsets=None):
    """Mean matvec time and relative L2 error against the FP64 dense ``ref``
    (the worst geometry's when ``geom_offsets`` is given).  ``sigma`` of shape
    (R, n) times the multi-right-hand-side matvec; row 0 is checked."""
    from .plan import FmmPlan
    plan = FmmPlan(C, A, leaf_size=cfg.leaf_size, order=cfg.native_order, ws=cfg.ws,
                   max_level=cfg.max_level, m2l_precision=cfg.m2l_precision,
                   geom_offsets=geom_offsets)
    out = torch.empty_like(sigma)
    mv = (lambda: plan.matvec_multi(sigma, out=out)) if sigma.dim() == 2 else \
        (lambda: plan.matvec(sigma, out=out))
    # warm-up, then a run long enough that the start-up gap (the host ahead
    # of an idle device by one matvec) stays small: with 5 timed matvecs the
    # 82k sphere read 0.102 ms at leaf 256 and 384 alike, against 0.087 and
    # 0.100 ms in a long run
    for _ in range(3):
        mv()
    torch.cuda.synchronize(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        mv()
    e1.record()
    torch.cuda.synchronize(dev)
    ms = e0.elapsed_time(e1) / reps
    o = out[0] if sigma.dim() == 2 else out
    off = [0, o.shape[0]] if geom_offsets is None else [int(v) for v in geom_offsets]
    err = max(float(torch.linalg.norm(o[b:e] - ref[b:e]) / torch.linalg.norm(ref[b:e]))
              for b, e in zip(off[:-1], off[1:]) if e > b)
    st = plan.stats()
    plan.close()
    return dict(ms=ms, rel_err=err, n_leaves=float(st["n_leaves"]),
                n_m2l_pairs=float(st["n_m2l_pairs"]),
                n_p2p_interactions=float(st["n_p2p_interactions"]))


def autotune_for_problem(base_cfg: FmmConfig, problem_size: int, *, dry_run: bool = True,
                         target_rel_error: float = 1e-2, max_trials: int = 24, device=None,
                         use_cache: bool = True, centroids=None, areas=None,
                         reps: int = 20, geom_offsets=None, nrhs: int = 1) -> TuningResult:
    """Picks (order, leaf) with the fastest matvec whose relative L2 error
    against the FP64 dense operator is at most ``target_rel_error``.  With
    ``geom_offsets`` the plans are batched over those geometries and the error
    is the worst geometry's; with ``nrhs`` > 1 the multi-right-hand-side
    matvec over that many vectors is timed."""
    if problem_size <= 0:
        raise ValueError("problem_size must be positive")
    base_cfg.validate()
    if dry_run:
        return TuningResult(config=base_cfg,
                            metrics={"mode": 0.0, "problem_size": float(problem_size)})
    # cache records describe the synthetic geometry of a size bucket with one
    # right-hand side: a caller's own geometry is tuned but never cached
    use_cache = use_cache and geom_offsets is None and nrhs == 1 and centroids is None
    nat.require_gpu()
    dev = torch.device(device or "cuda")
    if dev.type != "cuda":
        raise ValueError("the native autotuner measures on a GPU device")
    fp = _hardware_fingerprint(dev)
    bucket = _bucket_problem_size(problem_size)
    key = _cache_key(fp, bucket, base_cfg, target_rel_error)
    if use_cache:
        rec = _load_best_from_cache(key)
        if rec is not None:
            cfg = dataclasses.replace(base_cfg, expansion_order=int(rec["expansion_order"]),
                                      leaf_size=int(rec["leaf_size"]))
            m = dict(rec["metrics"])
            m["from_cache"] = 1.0
            return TuningResult(config=cfg, metrics=m)

    if centroids is None:
        C, A = _synthetic_geometry(problem_size)
    else:
        C = np.asarray(centroids, dtype=np.float64).reshape(-1, 3)
        A = np.asarray(areas, dtype=np.float64).reshape(-1)
    n = C.shape[0]
    rng = np.random.default_rng(0)
    sigma = torch.as_tensor(rng.normal(size=n), dtype=torch.float64, device=dev)
    from ..core.bem_kernel import dense_matvec
    Ct, At = torch.as_tensor(C, device=dev), torch.as_tensor(A, device=dev)
    if geom_offsets is None:
        ref = dense_matvec(sigma, Ct, At)
    else:
        off = [int(v) for v in geom_offsets]
        ref = torch.cat([dense_matvec(sigma[b:e], Ct[b:e], At[b:e])
                         for b, e in zip(off[:-1], off[1:])])
    if nrhs > 1:
        sigma = torch.cat([sigma[None], torch.as_tensor(rng.normal(size=(nrhs - 1, n)),
                                                        dtype=torch.float64, device=dev)])

    trials = []
    budget = [max_trials]

    def run(p, leaf):
        if budget[0] <= 0:
            return None
        budget[0] -= 1
        cfg = dataclasses.replace(base_cfg, expansion_order=p, leaf_size=leaf)
        try:
            m = _measure(cfg, C, A, sigma, ref, reps, dev, geom_offsets)
        except RuntimeError:
            return None
        trials.append((cfg, m))
        return m

    leaves = list(LEAF_SIZES)
    start = leaves.index(256) if 256 in leaves else len(leaves) // 2
    for p in (o for o in SUPPORTED_ORDERS if o >= 6):
        first = run(p, leaves[start])
        ok_ms = [t[1]["ms"] for t in trials if t[1]["rel_err"] <= target_rel_error]
        if first is not None and ok_ms and first["ms"] > 2.0 * min(ok_ms):
            break  # a higher order only costs more once a lower one meets the target
        if first is None or first["rel_err"] > 2.0 * target_rel_error:
            continue
        best = first["ms"]
        for step in (1, -1):  # towards larger leaves first, else smaller
            i, moved = start + step, False
            while 0 <= i < len(leaves):
                m = run(p, leaves[i])
                if m is None or m["ms"] >= best:
                    break
                best, moved = m["ms"], True
                i += step
            if moved:
                break
    if not trials:
        return TuningResult(config=base_cfg, metrics={"mode": 1.0, "problem_size": float(n)})
    ok = [t for t in trials if t[1]["rel_err"] <= target_rel_error]
    best_cfg, best = (min(ok, key=lambda t: t[1]["ms"]) if ok
                      else min(trials, key=lambda t: t[1]["rel_err"]))
    metrics = dict(best)
    metrics.update(mode=2.0, problem_size=float(n), trials=float(len(trials)),
                   met_target=1.0 if ok else 0.0,
                   panels_per_s=n / (best["ms"] * 1e-3))
    if use_cache:
        _append_record_to_cache(key, dict(expansion_order=best_cfg.expansion_order,
                                          leaf_size=best_cfg.leaf_size, metrics=metrics))
    return TuningResult(config=best_cfg, metrics=metrics)


__all__ = ["TuningResult", "autotune_for_problem", "AUTOTUNE_CACHE_ENV_VAR"]
