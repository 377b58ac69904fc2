"""Reference-mode octree and interaction lists restated (numpy).

Test infrastructure only (see oracle/__init__.py).

- ``build_tree``   fmm3d/tree.py:801-960 (``_build_adaptive_octree``) and the
                   SoA of fmm3d/tree.py:286-339 (``FmmTree._build_soa``):
                   tight per-node bounding box (geometry.py:125-157, 240-252),
                   centre 0.5 (min + max), half = max(0.5 (max - min), 0);
                   leaf iff n <= leaf_size, level >= max_depth or
                   max(half) < 1e-7 (:895-908); split by the strict test
                   point > centre, octant = x + 2y + 4z (:914-919); stable
                   partition by octant (:921-939); children built
                   depth-first in octant order, node ids in pre-order
                   (:878-889, :941-951); radius = |half| (1 + 1e-12) (:331).
- ``build_lists``  fmm3d/interaction_lists.py:374-528 with the MAC of
                   :273-297 (dist > 0 and max(2 r_s, 2 r_t) / dist < theta),
                   the refinement side of :306-337 (leaf rules, larger
                   diameter with 1 % hysteresis, ties by point count),
                   LIFO traversal from (root, root), self pairs never
                   admissible, leaf-leaf pairs to P2P, dedup and sort by
                   (tgt, src) (:474-485).

Arithmetic is in the points' dtype (float32 or float64) for the tree; norms
are sqrt(x^2 + y^2 + z^2) summed in that order in the dtype and the MAC is
evaluated in double, as the reference does with Python floats.

Parity: pinned by the invariants of FmmTree.verify (fmm3d/tree.py:618-793:
leaf ranges partition [0, N), parent/child/level consistency, children cover
their parent's range contiguously, stored boxes equal the recomputed tight
boxes) and of verify_interaction_lists (interaction_lists.py:580-621), by a
brute-force check that every point lies in its leaf's box on the correct
side of every ancestor's split planes, and by the partition property: every
ordered (source point, target point) pair is covered exactly once by
P2P u M2L (tests/test_cpu_ref_tree.py).
"""
import numpy as np


def build_tree(points, leaf_size=64, max_depth=32):
    P = np.asarray(points)
    if P.ndim != 2 or P.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    dt = P.dtype.type
    N = P.shape[0]
    idx = np.arange(N, dtype=np.int64)
    nodes = []  # dicts in pre-order
    if N == 0:
        return _soa([], idx, dt)

    def node(start, end, level, parent):
        loc = idx[start:end]
        lp = P[loc]
        mins = lp.min(axis=0)
        maxs = lp.max(axis=0)
        center = (dt(0.5) * (mins + maxs)).astype(dt)
        half = np.maximum(dt(0.5) * (maxs - mins), dt(0.0)).astype(dt)
        me = len(nodes)
        nodes.append(dict(center=center, half=half, level=level, parent=parent,
                          start=start, end=end, children=[-1] * 8))
        n = end - start
        degenerate = float(half.max()) < 1e-7
        if n <= leaf_size or level >= max_depth or degenerate:
            return me
        rel = lp > center[None, :]
        octs = rel[:, 0].astype(np.int64) + 2 * rel[:, 1] + 4 * rel[:, 2]
        order = np.argsort(octs, kind="stable")
        idx[start:end] = loc[order]
        counts = np.bincount(octs, minlength=8)
        off = start
        bounds = []
        for o in range(8):
            bounds.append((off, off + int(counts[o])))
            off += int(counts[o])
        for o in range(8):
            s, e = bounds[o]
            if e > s:
                nodes[me]["children"][o] = node(s, e, level + 1, me)
        return me

    node(0, N, 0, -1)
    return _soa(nodes, idx, dt)


def _norm3(v, dt):
    x, y, z = (dt(c) for c in v)
    return dt(np.sqrt(dt(dt(x * x + y * y) + z * z)))


def _soa(nodes, idx, dt):
    M = len(nodes)
    out = dict(
        index_map=idx,
        centers=np.zeros((M, 3), dtype=dt), half=np.zeros((M, 3), dtype=dt),
        radii=np.zeros(M, dtype=dt), levels=np.zeros(M, dtype=np.int32),
        parents=np.full(M, -1, dtype=np.int64), children=np.full((M, 8), -1, dtype=np.int64),
        ranges=np.zeros((M, 2), dtype=np.int64), is_leaf=np.zeros(M, dtype=bool))
    for i, nd in enumerate(nodes):
        out["centers"][i] = nd["center"]
        out["half"][i] = nd["half"]
        out["radii"][i] = dt(_norm3(nd["half"], dt) * dt(1.0 + 1e-12))
        out["levels"][i] = nd["level"]
        out["parents"][i] = nd["parent"]
        out["children"][i] = nd["children"]
        out["ranges"][i] = (nd["start"], nd["end"])
        out["is_leaf"][i] = all(c < 0 for c in nd["children"])
    return out


def build_lists(src, tgt=None, theta=0.5):
    """Dual-tree traversal; returns (p2p_pairs, m2l_pairs, stats) with pairs
    as int64 (K, 2) arrays of (src, tgt) sorted by (tgt, src)."""
    same = tgt is None
    if same:
        tgt = src
    if theta <= 0.0 or theta >= 1.0:
        raise ValueError("mac_theta must be in (0, 1)")
    stats = dict(num_mac_accept=0, num_mac_reject=0, num_leaf_leaf=0, num_stack_pops=0,
                 max_stack_size=1)
    ms, mt = src["centers"].shape[0], tgt["centers"].shape[0]
    if ms == 0 or mt == 0:
        z = np.zeros((0, 2), dtype=np.int64)
        return z, z, stats
    dt_s = src["centers"].dtype.type
    leaf_s, leaf_t = src["is_leaf"], tgt["is_leaf"]

    def diam(tr, i):
        return float(2.0 * tr["radii"][i])

    def dist(i, j):
        d = (src["centers"][i] - tgt["centers"][j]).astype(dt_s)
        return float(_norm3(d, dt_s))

    def count(tr, i):
        return int(tr["ranges"][i, 1] - tr["ranges"][i, 0])

    p2p, m2l = set(), set()
    stack = [(0, 0)]
    while stack:
        s, t = stack.pop()
        stats["num_stack_pops"] += 1
        mac = False
        if not (same and s == t):
            dd = dist(s, t)
            if dd != 0.0:
                mac = max(diam(src, s), diam(tgt, t)) / dd < theta
            if mac:
                stats["num_mac_accept"] += 1
            else:
                stats["num_mac_reject"] += 1
        if mac:
            m2l.add((s, t))
            continue
        if leaf_s[s] and leaf_t[t]:
            stats["num_leaf_leaf"] += 1
            p2p.add((s, t))
            continue
        # refinement side (interaction_lists.py:306-337)
        if leaf_s[s] and not leaf_t[t]:
            side = "tgt"
        elif leaf_t[t] and not leaf_s[s]:
            side = "src"
        else:
            ds, dtg = diam(src, s), diam(tgt, t)
            if ds > dtg * 1.01:
                side = "src"
            elif dtg > ds * 1.01:
                side = "tgt"
            else:
                side = "src" if count(src, s) >= count(tgt, t) else "tgt"
        if side == "src":
            for c in src["children"][s]:
                if c >= 0:
                    stack.append((int(c), t))
        else:
            for c in tgt["children"][t]:
                if c >= 0:
                    stack.append((s, int(c)))
        stats["max_stack_size"] = max(stats["max_stack_size"], len(stack))

    def arr(pairs):
        a = sorted(pairs, key=lambda st: (st[1], st[0]))
        return np.array(a, dtype=np.int64).reshape(-1, 2)
    return arr(p2p), arr(m2l), stats
