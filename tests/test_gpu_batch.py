"""GPU: batched plans over several independent geometries.

A batched plan (one merged tree, one launch per stage over all geometries)
must reproduce the per-geometry plans to FP32 rounding, never couple
geometries, and its block-diagonal solve must meet the same residual as
separate solves.  Oracle: oracle.dense (core/bem_kernel.py:157-230) per block.
"""
import numpy as np
import pytest
import torch

from oracle import dense as odense

pytestmark = pytest.mark.gpu


def _meshes():
    from electrodrive_amd.core import bem_mesh
    return [bem_mesh.icosphere(3, 0.7), bem_mesh.torus(1.0, 0.3, 40, 14),
            bem_mesh.sphere_dimer(0.5, 0.2, 2), bem_mesh.square_plate(1.0, 14),
            bem_mesh.icosphere(1, 1.3)]


def test_batched_matvec_equals_separate_plans():
    from electrodrive_amd.fmm3d.plan import FmmPlan
    ms = _meshes()
    C = np.concatenate([m.centroids for m in ms])
    A = np.concatenate([m.areas for m in ms])
    off = np.concatenate([[0], np.cumsum([m.n_panels for m in ms])])
    rng = np.random.default_rng(4)
    sig = rng.normal(size=C.shape[0])
    bp = FmmPlan(C, A, leaf_size=32, order=10, ws=1, geom_offsets=off)
    assert np.array_equal(bp.geometries(), off)
    out = bp.matvec(torch.tensor(sig).cuda()).cpu().numpy()
    for g, m in enumerate(ms):
        b, e = off[g], off[g + 1]
        p = FmmPlan(m.centroids, m.areas, leaf_size=32, order=10, ws=1)
        sep = p.matvec(torch.tensor(sig[b:e]).cuda()).cpu().numpy()
        assert np.max(np.abs(out[b:e] - sep)) <= 1e-6 * np.max(np.abs(sep))
        ref = odense.dense_matvec(m.centroids, m.areas, sig[b:e])
        assert np.linalg.norm(out[b:e] - ref) / np.linalg.norm(ref) < 2e-5


def test_batched_solve():
    from electrodrive_amd.core.operator import BemOperator
    ms = _meshes()
    op = BemOperator(ms, order=10, ws=1, leaf_size=32)
    charges = [[(1.0, (0.0, 0.0, 1.5))], [(1.0, (0.0, 0.0, 1.0))], [(-1.0, (0.0, 1.0, 0.0))],
               [(1.0, (0.2, 0.1, 0.5))], [(2.0, (3.0, 0.0, 0.0))]]
    b = op.rhs_point_charges(charges, per_geometry=True)
    x, info = op.solve(b, tol=1e-8)
    print("batched solve", info)
    assert info["code"] == 0
    r = b - op.accurate(x)
    for g in range(len(ms)):
        s, e = int(op.geom_offsets[g]), int(op.geom_offsets[g + 1])
        assert float(torch.linalg.norm(r[s:e]) / torch.linalg.norm(b[s:e])) <= 1e-8
    assert info["max_geometry_rel_resid"] <= 1e-8
    xs, infos = op.solve(b, tol=1e-8, precond="sai")
    assert infos["code"] == 0 and infos["max_geometry_rel_resid"] <= 1e-8
    assert float(torch.linalg.norm(xs - x) / torch.linalg.norm(x)) < 1e-6
    # each block equals its own single-geometry solve; one Krylov space per
    # geometry: the batch takes about as many Arnoldi steps as its slowest
    # geometry alone, not the steps of the joint spectrum
    single = []
    for g, (m, ch) in enumerate(zip(ms, charges)):
        s, e = int(op.geom_offsets[g]), int(op.geom_offsets[g + 1])
        og = BemOperator(m, order=10, ws=1, leaf_size=32)
        xg, ig = og.solve(og.rhs_point_charges(ch), tol=1e-8)
        assert float(torch.linalg.norm(x[s:e] - xg) / torch.linalg.norm(xg)) < 1e-6
        single.append(og.solve(og.rhs_point_charges(ch), tol=1e-8, precond="sai")[1]["iters"])
    print("batched SAI steps", infos["iters"], "single-geometry SAI steps", single)
    assert infos["iters"] <= 1.25 * max(single) + 4
