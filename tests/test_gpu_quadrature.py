"""GPU parity: self integrals, near pairs, near weights, near apply, Morton keys.

Oracles: oracle.quadrature (core/bem.py:475-734, :1777-1816,
core/bem_quadrature.py) and oracle.morton (fmm3d/geometry.py:671-799).
Near pairs and Morton keys are index work: bit-exact.  Quadrature values are
FP64: 1e-12 relative.
"""
import numpy as np
import pytest
import torch

from oracle import morton as omorton
from oracle import quadrature as oq

pytestmark = pytest.mark.gpu


def _mesh():
    from electrodrive_amd.core import bem_mesh
    return bem_mesh.icosphere(3)


def test_self_integrals_match_oracle():
    from electrodrive_amd.core import bem_quadrature as bq
    m = _mesh()
    for depth in (0, 1, 3, 4):
        ref, ref_corr = oq.self_integrals(m.centroids, m.panel_vertices, m.areas, depth=depth)
        I, corr = bq.self_integrals(m.centroids, m.panel_vertices, m.areas, depth=depth)
        err = np.max(np.abs(I.cpu().numpy() - ref) / np.abs(ref))
        print("depth", depth, "self integral rel err", err)
        assert err < 1e-12
        assert np.max(np.abs(corr.cpu().numpy() - ref_corr) / np.abs(ref_corr)) < 1e-12


def test_near_pairs_bit_exact():
    from electrodrive_amd.core import bem_quadrature as bq
    from electrodrive_amd.core import bem_mesh
    for m in (_mesh(), bem_mesh.torus(1.0, 0.3, 24, 10), bem_mesh.square_plate(1.0, 12)):
        for f in (1.5, 2.5):
            ref = oq.near_pairs(m.centroids, m.areas, f)
            got = bq.build_near_pairs(m.centroids, m.areas, f).cpu().numpy()
            assert got.shape == ref.shape
            assert np.array_equal(got, ref)


def _cloud(seed, n):
    rng = np.random.default_rng(seed)
    C = rng.uniform(-1.0, 1.0, (n, 3))
    A = np.exp(rng.normal(np.log(4.0 / n), 1.0, n))  # radii spread over a decade
    return C, A


@pytest.mark.parametrize("case", ["cloud", "edge_values", "far_clusters", "one_big_panel"])
def test_near_pairs_bit_exact_edge_cases(case):
    """Cell-grid search against the brute-force scan of core/bem.py:475-555 on
    inputs that stress the grid: mixed radii, NaN / zero / negative / inf areas
    and non-finite centroids, clusters 1e6 apart (cell-count cap), and one panel
    whose radius sets a cell larger than the whole cloud."""
    from electrodrive_amd.core import bem_quadrature as bq
    C, A = _cloud(7, 3000)
    if case == "edge_values":
        A[[3, 40]] = np.nan
        A[[5, 41]] = 0.0
        A[[7, 42]] = -1e-3
        A[9] = np.inf
        C[11] = [np.nan, 0.0, 0.0]
        C[13] = [np.inf, 0.0, 0.0]
        C[15] = C[16]  # coincident centroids are never paired
    elif case == "far_clusters":
        C[1500:] += 1e6
    elif case == "one_big_panel":
        A[17] = 50.0
    for f in (1.5, 0.0, -1.0) if case == "edge_values" else (1.5, 3.0):
        ref = oq.near_pairs(C, A, f)
        got = bq.build_near_pairs(C, A, f).cpu().numpy()
        print(case, f, "pairs", ref.shape[0])
        assert got.shape == ref.shape
        assert np.array_equal(got, ref)


def test_near_weights_and_apply():
    from electrodrive_amd.core import bem_quadrature as bq
    m = _mesh()
    pairs = oq.near_pairs(m.centroids, m.areas, 1.5)
    for depth in (None, 1, 3):
        ref = oq.near_weights(m.centroids, m.areas, m.panel_vertices, pairs, 2, depth)
        got = bq.near_correction_weights(m.centroids, m.areas, m.panel_vertices, pairs,
                                         quad_order=2, max_depth=depth).cpu().numpy()
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print("near weights depth", depth, "err", err)
        assert err < 1e-11
    rng = np.random.default_rng(3)
    vfar = rng.normal(size=m.n_panels)
    sig = rng.normal(size=m.n_panels)
    w = oq.near_weights(m.centroids, m.areas, m.panel_vertices, pairs, 2, 3)
    ref = oq.near_apply(vfar, sig, pairs, w)
    got = bq.near_quadrature_matvec(torch.tensor(vfar).cuda(), torch.tensor(sig).cuda(),
                                    m.centroids, m.areas, m.panel_vertices, pairs,
                                    weights=w).cpu().numpy()
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12
    # on-the-fly rule (CUDA-extension semantics) == precomputed depth-0 weights
    w0 = oq.near_weights(m.centroids, m.areas, m.panel_vertices, pairs, 2, None)
    ref0 = oq.near_apply(vfar, sig, pairs, w0)
    got0 = bq.near_quadrature_matvec(torch.tensor(vfar).cuda(), torch.tensor(sig).cuda(),
                                     m.centroids, m.areas, m.panel_vertices, pairs,
                                     quad_order=2).cpu().numpy()
    assert np.max(np.abs(got0 - ref0)) / np.max(np.abs(ref0)) < 1e-11


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_morton_keys_bit_exact(dtype):
    from electrodrive_amd.fmm3d import BoundingBox, morton_keys_from_points
    rng = np.random.default_rng(42)
    for pts in (rng.uniform(-3, 5, size=(5000, 3)), rng.normal(size=(777, 3)) * [1, 1e-3, 0],
                np.zeros((1, 3))):
        p = pts.astype(dtype)
        for level in (1, 10, 16, 21):
            ref = omorton.morton_keys(p, level)
            got, bb = morton_keys_from_points(torch.tensor(p).cuda(), level=level)
            assert np.array_equal(got.cpu().numpy(), ref), (level, p.shape)
        # the returned box is BoundingBox.from_min_max of the padded tight box
        # (fmm3d/geometry.py:790-793)
        mn, mx = omorton.global_bbox(p)
        assert isinstance(bb, BoundingBox) and bb.center.dtype == torch.tensor(p).dtype
        assert np.array_equal(bb.center.cpu().numpy(), (dtype(0.5) * (mn + mx)).astype(dtype))
        assert np.array_equal(bb.half_extent.cpu().numpy(),
                              np.maximum(dtype(0.5) * (mx - mn), 0).astype(dtype))
        c = np.array([0.5, 0.5, 0.5], dtype=dtype)
        h = np.array([4.0, 4.0, 4.0], dtype=dtype)
        ref = omorton.morton_keys(p, 16, c, h)
        got, _ = morton_keys_from_points(torch.tensor(p).cuda(), level=16, global_bbox=(c, h))
        assert np.array_equal(got.cpu().numpy(), ref)
        box = BoundingBox(torch.tensor(c), torch.tensor(h))
        got, bb2 = morton_keys_from_points(torch.tensor(p).cuda(), level=16, global_bbox=box)
        assert np.array_equal(got.cpu().numpy(), ref) and bb2 is box


def test_attach_near_device_equals_host_path():
    """ed_plan_attach_near_device builds the same tree-ordered CSR as the host
    path: the FMM matvec with either attached correction is bit-identical;
    unsorted pairs are rejected."""
    from electrodrive_amd.core import bem_mesh
    from electrodrive_amd.core import bem_quadrature as bq
    from electrodrive_amd.fmm3d.plan import FmmPlan
    m = bem_mesh.torus(1.0, 0.3, 40, 12)
    dev = torch.device("cuda")
    C = torch.tensor(m.centroids, device=dev)
    A = torch.tensor(m.areas, device=dev)
    PV = torch.tensor(m.panel_vertices, device=dev)
    pairs = bq.build_near_pairs(C, A, 1.5)
    w = bq.near_correction_weights(C, A, PV, pairs, quad_order=2, max_depth=2)
    p1 = FmmPlan(m.centroids, m.areas, leaf_size=32, order=8)
    p2 = FmmPlan(m.centroids, m.areas, leaf_size=32, order=8)
    p1.attach_near(pairs.cpu().numpy(), w.cpu().numpy())
    p2.attach_near_device(pairs, w)
    sig = torch.tensor(np.random.default_rng(2).normal(size=m.n_panels), device=dev)
    assert torch.equal(p1.matvec(sig), p2.matvec(sig))
    assert p1.stats()["n_near_pairs"] == p2.stats()["n_near_pairs"] == pairs.shape[0]
    with pytest.raises(ValueError):
        p2.attach_near_device(pairs.flip(0).contiguous(), w)
