"""Reference-mode octree and interaction lists (CPU; the builder is host code).

Oracle (oracle/ref_tree.py, restating fmm3d/tree.py:801-960 and
fmm3d/interaction_lists.py:273-528) pinned here by:
- the FmmTree.verify invariants (fmm3d/tree.py:618-793) with exact boxes;
- a brute-force walk of every point down from the root: at each internal
  node the point's child is the octant (x > c_x) + 2 (y > c_y) + 4 (z > c_z),
  checked on lattice points that lie exactly on split planes (the strict '>'
  sends them to the lower child; a '>=' oracle fails this);
- the lists' partition property: every ordered (source point, target point)
  pair is covered exactly once by P2P u M2L, every M2L pair meets the MAC,
  every P2P pair is leaf-leaf and fails it (or is a self pair).
Native (ed_ref_tree_* / ed_ref_lists_*) against the oracle: node SoA,
index_map and (src, tgt) pair arrays bit-exact, on the reference's own
seeded generators (fmm3d/tests/test_stress.py:25-53, seeds 123, 555, 42),
mesh centroids, FP32 points, two-tree lists, ties and degenerate inputs.
"""
import math

import numpy as np
import pytest
import torch

from oracle import ref_tree as rt


def _ref_points(n, seed, mode):
    """The reference stress tests' generator (fmm3d/tests/test_stress.py:25-53)."""
    torch.manual_seed(seed)
    if mode == "uniform":
        return 2.0 * torch.rand(n, 3, dtype=torch.float64) - 1.0
    if mode == "clusters":
        n1 = n // 2
        c1 = torch.tensor([0.5, 0.0, 0.0], dtype=torch.float64)
        c2 = torch.tensor([-0.5, 0.0, 0.0], dtype=torch.float64)
        return torch.cat([c1 + 0.1 * torch.randn(n1, 3, dtype=torch.float64),
                          c2 + 0.1 * torch.randn(n - n1, 3, dtype=torch.float64)])
    u = torch.randn(n, 3, dtype=torch.float64)
    u = u / u.norm(dim=1, keepdim=True)
    return 0.8 * u + 0.05 * torch.randn(n, 3, dtype=torch.float64)


def _lattice(k):
    g = np.arange(k, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def _check_tree_invariants(P, T, leaf_size, max_depth):
    n, m = P.shape[0], T["centers"].shape[0]
    im = T["index_map"]
    assert sorted(im.tolist()) == list(range(n))
    R, ch, par, lev, leaf = T["ranges"], T["children"], T["parents"], T["levels"], T["is_leaf"]
    lr = R[leaf]
    lr = lr[np.argsort(lr[:, 0])]
    assert lr[0, 0] == 0 and lr[-1, 1] == n and np.all(lr[1:, 0] == lr[:-1, 1])
    assert par[0] == -1 and lev[0] == 0
    dt = P.dtype.type
    Pt = P[im]
    for i in range(m):
        kids = [c for c in ch[i] if c >= 0]
        assert bool(leaf[i]) == (not kids)
        for c in kids:
            assert par[c] == i and lev[c] == lev[i] + 1
            assert c > i  # pre-order ids
        if kids:
            cr = sorted((R[c, 0], R[c, 1]) for c in kids)
            assert cr[0][0] == R[i, 0] and cr[-1][1] == R[i, 1]
            assert all(a[1] == b[0] for a, b in zip(cr, cr[1:]))
            assert [c for c in ch[i] if c >= 0] == sorted(kids)  # octant order = id order
        pts = Pt[R[i, 0]:R[i, 1]]
        mn, mx = pts.min(0), pts.max(0)
        assert np.array_equal(T["centers"][i], (dt(0.5) * (mn + mx)).astype(dt))
        assert np.array_equal(T["half"][i], np.maximum(dt(0.5) * (mx - mn), 0).astype(dt))
        npts = R[i, 1] - R[i, 0]
        if leaf[i]:
            assert npts <= leaf_size or lev[i] >= max_depth or float(T["half"][i].max()) < 1e-7
        else:
            assert npts > leaf_size and lev[i] < max_depth


def _check_split_sides(P, T):
    """Walk every point from the root: the child holding it is its octant."""
    im, ch, R, C = T["index_map"], T["children"], T["ranges"], T["centers"]
    pos = np.empty(P.shape[0], dtype=np.int64)
    pos[im] = np.arange(P.shape[0])
    for j in range(P.shape[0]):
        node = 0
        while True:
            assert R[node, 0] <= pos[j] < R[node, 1]
            if T["is_leaf"][node]:
                break
            p, c = P[j], C[node]
            o = int(p[0] > c[0]) + 2 * int(p[1] > c[1]) + 4 * int(p[2] > c[2])
            node = ch[node, o]
            assert node >= 0


def _check_partition(S, Tt, p2p, m2l, theta, same):
    ns, nt = S["index_map"].shape[0], Tt["index_map"].shape[0]
    cover = np.zeros((nt, ns), dtype=np.int32)
    for s, t in np.concatenate([p2p, m2l]):
        cover[Tt["ranges"][t, 0]:Tt["ranges"][t, 1], S["ranges"][s, 0]:S["ranges"][s, 1]] += 1
    assert cover.min() == 1 and cover.max() == 1
    for s, t in m2l:
        d = float(np.sqrt(((S["centers"][s] - Tt["centers"][t]) ** 2).sum()))
        assert max(2 * float(S["radii"][s]), 2 * float(Tt["radii"][t])) / d < theta
    for s, t in p2p:
        assert S["is_leaf"][s] and Tt["is_leaf"][t]
        if not (same and s == t):
            d = float(np.sqrt(((S["centers"][s] - Tt["centers"][t]) ** 2).sum()))
            assert d == 0 or max(2 * float(S["radii"][s]), 2 * float(Tt["radii"][t])) / d >= theta
    key = lambda a: [(t, s) for s, t in a.tolist()]  # noqa: E731
    assert key(p2p) == sorted(key(p2p)) and key(m2l) == sorted(key(m2l))


@pytest.mark.parametrize("case", ["uniform", "clusters", "shell", "lattice", "coincident"])
def test_oracle_tree_and_lists_pins(case):
    if case == "lattice":
        P = _lattice(9)  # centres land exactly on lattice planes
    elif case == "coincident":
        P = np.concatenate([np.zeros((40, 3)), _ref_points(300, 2, "uniform").numpy()])
    else:
        P = _ref_points(900, {"uniform": 123, "clusters": 555, "shell": 42}[case], case).numpy()
    for leaf in (1, 16, 64):
        T = rt.build_tree(P, leaf, 32)
        _check_tree_invariants(P, T, leaf, 32)
        _check_split_sides(P, T)
    T = rt.build_tree(P, 16, 32)
    p2p, m2l, _ = rt.build_lists(T, None, 0.5)
    _check_partition(T, T, p2p, m2l, 0.5, True)


def test_oracle_strict_split_discriminates():
    """On lattice points a '>=' split would move the points on each centre
    plane to the upper child: the split-side walk must reject that tree."""
    P = _lattice(5)
    T = rt.build_tree(P, 4, 32)
    assert T["centers"][0].tolist() == [2.0, 2.0, 2.0]
    ch0 = T["children"][0]
    assert (ch0 >= 0).all()
    # the lower child along x holds x in {0, 1, 2}: 3 of 5 planes
    assert T["ranges"][ch0[0], 1] - T["ranges"][ch0[0], 0] == 27
    bad = dict(T)
    bad["children"] = T["children"].copy()
    bad["children"][0] = ch0[[1, 0, 3, 2, 5, 4, 7, 6]]  # x bit flipped
    with pytest.raises(AssertionError):
        _check_split_sides(P, bad)


def test_oracle_two_tree_partition():
    S = _ref_points(700, 123, "uniform").numpy()
    Tp = _ref_points(500, 42, "shell").numpy() * 1.7
    TS, TT = rt.build_tree(S, 32, 32), rt.build_tree(Tp, 32, 32)
    p2p, m2l, _ = rt.build_lists(TS, TT, 0.6)
    _check_partition(TS, TT, p2p, m2l, 0.6, False)


def _native(P, leaf=64, depth=32):
    from electrodrive_amd.fmm3d.tree import build_fmm_tree
    return build_fmm_tree(torch.as_tensor(P), leaf_size=leaf, max_depth=depth)


def _assert_tree_equal(tr, T):
    assert np.array_equal(tr.index_map.numpy(), T["index_map"])
    for name, key in (("node_centers", "centers"), ("node_half_extents", "half"),
                      ("node_radii", "radii"), ("node_levels", "levels"),
                      ("node_parents", "parents"), ("node_children", "children"),
                      ("node_ranges", "ranges"), ("node_is_leaf", "is_leaf")):
        got = getattr(tr, name).numpy()
        assert got.dtype == T[key].dtype or name == "node_is_leaf", name
        assert np.array_equal(got, T[key]), name


def _inputs():
    from electrodrive_amd.core import bem_mesh
    yield "uniform123", _ref_points(4000, 123, "uniform").numpy(), 64
    yield "clusters555", _ref_points(3000, 555, "clusters").numpy(), 50
    yield "shell42", _ref_points(3000, 42, "shell").numpy(), 64
    yield "icosphere4", bem_mesh.icosphere(4).centroids, 64
    yield "torus", bem_mesh.torus(1.0, 0.3, 60, 20).centroids, 32
    yield "plate", bem_mesh.square_plate(2.0, 30).centroids, 16
    yield "lattice", _lattice(10), 8
    yield "fp32", _ref_points(3000, 123, "uniform").numpy().astype(np.float32), 64
    yield "coincident", np.concatenate([np.ones((100, 3)), _ref_points(200, 7, "shell").numpy()]), 16


def test_native_tree_and_lists_bit_exact():
    for name, P, leaf in _inputs():
        T = rt.build_tree(P, leaf, 32)
        tr = _native(P, leaf)
        _assert_tree_equal(tr, T)
        tr.verify()
        from electrodrive_amd.fmm3d.interaction_lists import (build_interaction_lists,
                                                              verify_interaction_lists)
        for theta in (0.5, 0.8):
            p2p, m2l, st = rt.build_lists(T, None, theta)
            L = build_interaction_lists(tr, mac_theta=theta)
            a, b = L.as_tensors()
            assert np.array_equal(a.numpy(), p2p) and np.array_equal(b.numpy(), m2l), name
            assert (L.stats.num_mac_accept, L.stats.num_mac_reject, L.stats.num_leaf_leaf,
                    L.stats.num_stack_pops, L.stats.max_stack_size) == (
                st["num_mac_accept"], st["num_mac_reject"], st["num_leaf_leaf"],
                st["num_stack_pops"], st["max_stack_size"]), name
            verify_interaction_lists(tr, tr, L)
        print(name, "nodes", tr.n_nodes, "p2p", len(p2p), "m2l", len(m2l))


def test_native_two_trees_and_unsorted_pairs():
    from electrodrive_amd.fmm3d.interaction_lists import build_interaction_lists
    S = _ref_points(2500, 123, "uniform").numpy()
    Tp = _ref_points(1500, 42, "shell").numpy() * 1.5
    TS, TT = rt.build_tree(S, 40, 32), rt.build_tree(Tp, 40, 32)
    p2p, m2l, _ = rt.build_lists(TS, TT, 0.5)
    ts, tt = _native(S, 40), _native(Tp, 40)
    L = build_interaction_lists(ts, tt, mac_theta=0.5)
    a, b = L.as_tensors()
    assert np.array_equal(a.numpy(), p2p) and np.array_equal(b.numpy(), m2l)
    Lu = build_interaction_lists(ts, tt, mac_theta=0.5, sort_pairs=False)
    assert Lu.p2p_pairs == sorted(map(tuple, p2p.tolist()))
    assert Lu.m2l_pairs == sorted(map(tuple, m2l.tolist()))
    assert sum(len(v) for v in L.v_list) == len(m2l) and L.w_list == L.v_list


def test_native_edge_cases_and_errors():
    from electrodrive_amd.fmm3d.interaction_lists import build_interaction_lists
    from electrodrive_amd.fmm3d.tree import build_fmm_tree
    # empty and one point
    t0 = build_fmm_tree(torch.zeros((0, 3), dtype=torch.float64))
    assert t0.n_nodes == 0 and t0.root == -1
    L0 = build_interaction_lists(t0)
    assert L0.p2p_pairs == [] and L0.m2l_pairs == []
    t1 = build_fmm_tree(torch.ones((1, 3), dtype=torch.float64))
    assert t1.n_nodes == 1 and bool(t1.node_is_leaf[0])
    assert build_interaction_lists(t1).p2p_pairs == [(0, 0)]
    # max_depth forces leaves above leaf_size
    P = _ref_points(500, 123, "uniform").numpy()
    T = rt.build_tree(P, 4, 2)
    tr = _native(P, 4, 2)
    _assert_tree_equal(tr, T)
    assert int(tr.node_levels.max()) == 2
    # permutation helpers
    v = torch.arange(500, dtype=torch.float64)
    assert torch.equal(tr.map_to_original_order(tr.map_to_tree_order(v)), v)
    st = tr.statistics()
    assert st.n_leaves == int(tr.node_is_leaf.sum()) and st.max_depth == 2
    # errors (fmm3d/tree.py:1018-1037, interaction_lists.py:392-395, :521-526)
    pts = torch.as_tensor(P)
    with pytest.raises(ValueError):
        build_fmm_tree(pts, leaf_size=0)
    with pytest.raises(ValueError):
        build_fmm_tree(pts, max_depth=0)
    with pytest.raises(ValueError):
        build_fmm_tree(pts, max_depth=65)
    with pytest.raises(ValueError):
        build_fmm_tree(pts[:, :2])
    with pytest.raises(TypeError):
        build_fmm_tree(pts.to(torch.int64))
    with pytest.raises(ValueError):
        build_interaction_lists(tr, mac_theta=1.0)
    with pytest.raises(ValueError):
        build_interaction_lists(tr, mac_theta=0.0)
    with pytest.raises(RuntimeError):
        build_interaction_lists(tr, mac_theta=0.5, max_pairs=3)
    _ = math
