"""Lattice octree and U/V interaction lists of the native FMM (numpy).

Test infrastructure only (see oracle/__init__.py).

The native FMM replaces the reference's tight-bbox adaptive tree
(fmm3d/tree.py:801-960) by a cubic lattice octree so that every M2L
translation depends only on an integer offset.  Points are ordered by the
reference's own Morton key (fmm3d/geometry.py:736-799) computed in FP64 at
level 21 inside the root cube (centre = bbox centre, half = max half-extent
enlarged by 2^-20 relative plus 1e-300), ties broken by original index.

Nodes: level 0 is the root; a node is split into its non-empty children iff
count > leaf_size and level < max_level.  Nodes are stored level by level,
Morton-sorted within a level.

Lists (ws = well-separation radius in boxes, 1 by default):
  adjacent(A, B)  same level and max |coord(A) - coord(B)| <= ws
  V(T)  nodes S at T's level, children of a non-leaf node adjacent to
        parent(T), and not adjacent to T                       -> M2L S->T
  U(T)  (T a leaf) leaves S with adjacent(anc_m(T), anc_m(S)), m = min levels
                                                               -> P2P S->T
Every ordered pair of points (i, j) is then covered exactly once, either by a
P2P pair of leaves or by one M2L pair of ancestors; ``coverage_counts``
checks that by brute force.
"""
import numpy as np

from .morton import interleave

KEY_LEVEL = 21


def root_cube(points):
    p = np.asarray(points, dtype=np.float64)
    mins = p.min(axis=0)
    maxs = p.max(axis=0)
    center = 0.5 * (mins + maxs)
    half = float(np.max(0.5 * (maxs - mins)))
    half = half * (1.0 + 2.0 ** -20) + 1e-300
    return center, half


def lattice_keys(points, center, half):
    """Morton keys at level 21 with the reference normalisation
    (geometry.py:696-703, 785-789) inside the cube (center, half)."""
    p = np.asarray(points, dtype=np.float64)
    mins = center - half
    maxs = center + half
    extent = np.maximum(maxs - mins, np.finfo(np.float64).eps)
    nrm = np.minimum(np.maximum((p - mins) / extent, 0.0), 1.0 - 1e-9)
    ijk = np.clip(np.floor(nrm * float(1 << KEY_LEVEL)).astype(np.int64), 0,
                  (1 << KEY_LEVEL) - 1)
    return interleave(ijk[:, 0], ijk[:, 1], ijk[:, 2], KEY_LEVEL), ijk


class LatticeTree:
    pass


def build(points, leaf_size=64, max_level=16, ws=1):
    p = np.asarray(points, dtype=np.float64)
    N = p.shape[0]
    center, half = root_cube(p)
    keys, ijk = lattice_keys(p, center, half)
    perm = np.lexsort((np.arange(N), keys))
    skeys = keys[perm]
    sijk = ijk[perm]
    t = LatticeTree()
    t.center, t.half, t.perm, t.keys = center, half, perm, skeys
    t.leaf_size, t.max_level, t.ws = leaf_size, max_level, ws
    levels, begins, ends, parents, coords, is_leaf = [], [], [], [], [], []
    first_child, n_child = [], []
    # level 0
    frontier = [(0, N, -1)]
    lvl = 0
    while frontier:
        nxt = []
        base = len(levels)
        for k, (b, e, par) in enumerate(frontier):
            idx = len(levels)
            levels.append(lvl)
            begins.append(b)
            ends.append(e)
            parents.append(par)
            shift = KEY_LEVEL - lvl
            coords.append(tuple(int(c) >> shift for c in sijk[b]))
            leaf = (e - b) <= leaf_size or lvl >= max_level
            is_leaf.append(leaf)
            first_child.append(-1)
            n_child.append(0)
        for k, (b, e, par) in enumerate(frontier):
            idx = base + k
            if is_leaf[idx]:
                continue
            shift = 3 * (KEY_LEVEL - lvl - 1)
            pref = skeys[b:e] >> shift
            cuts = np.flatnonzero(np.diff(pref)) + 1
            starts = np.concatenate([[0], cuts]) + b
            stops = np.concatenate([cuts, [e - b]]) + b
            first_child[idx] = base + len(frontier) + len(nxt)
            n_child[idx] = len(starts)
            for s0, s1 in zip(starts, stops):
                nxt.append((int(s0), int(s1), idx))
        frontier = nxt
        lvl += 1
    t.level = np.array(levels)
    t.begin = np.array(begins)
    t.end = np.array(ends)
    t.parent = np.array(parents)
    t.coord = np.array(coords, dtype=np.int64).reshape(-1, 3)
    t.is_leaf = np.array(is_leaf)
    t.first_child = np.array(first_child)
    t.n_child = np.array(n_child)
    t.n_nodes = len(levels)
    t.box_side = lambda l: 2.0 * half / (1 << l)
    _build_lists(t)
    return t


def _adjacent(t, a, b):
    return np.max(np.abs(t.coord[a] - t.coord[b])) <= t.ws


def _colleagues(t, by_level, n):
    l = t.level[n]
    cand = by_level[l]
    d = np.max(np.abs(t.coord[cand] - t.coord[n][None, :]), axis=1)
    return cand[d <= t.ws]


def _leaf_descendants(t, n):
    if t.is_leaf[n]:
        return [n]
    out = []
    for c in range(t.first_child[n], t.first_child[n] + t.n_child[n]):
        out.extend(_leaf_descendants(t, c))
    return out


def _build_lists(t):
    by_level = {}
    for l in np.unique(t.level):
        by_level[int(l)] = np.flatnonzero(t.level == l)
    V = [[] for _ in range(t.n_nodes)]
    U = [[] for _ in range(t.n_nodes)]
    for n in range(t.n_nodes):
        par = t.parent[n]
        if par >= 0:
            for c in _colleagues(t, by_level, par):
                if t.is_leaf[c]:
                    continue
                for s in range(t.first_child[c], t.first_child[c] + t.n_child[c]):
                    if not _adjacent(t, s, n):
                        V[n].append(s)
        if t.is_leaf[n]:
            us = []
            a = n
            while a >= 0:
                for c in _colleagues(t, by_level, a):
                    if t.is_leaf[c]:
                        us.append(c)
                    elif a == n:
                        us.extend(_leaf_descendants(t, c))
                a = t.parent[a]
            U[n] = sorted(set(us))
        V[n] = sorted(V[n])
    t.V = V
    t.U = U


def point_leaf(t):
    """Leaf node index of each point in tree order."""
    out = np.empty(t.end[0], dtype=np.int64)
    for n in np.flatnonzero(t.is_leaf):
        out[t.begin[n]:t.end[n]] = n
    return out


def coverage_counts(t):
    """(N, N) count of how often each (target, source) point pair in tree
    order is covered by P2P or M2L; every entry must be 1."""
    N = t.end[0]
    cnt = np.zeros((N, N), dtype=np.int64)
    for T in np.flatnonzero(t.is_leaf):
        for S in t.U[T]:
            cnt[t.begin[T]:t.end[T], t.begin[S]:t.end[S]] += 1
    for T in range(t.n_nodes):
        for S in t.V[T]:
            cnt[t.begin[T]:t.end[T], t.begin[S]:t.end[S]] += 1
    return cnt
