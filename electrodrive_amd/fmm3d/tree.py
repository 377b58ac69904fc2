"""Reference-mode adaptive octree (mirror of fmm3d/tree.py of the reference).

``build_fmm_tree(points, *, leaf_size=64, max_depth=32)`` builds the
reference's tight-bbox octree natively (``ed_ref_tree_create``) and returns an
``FmmTree`` with the reference's SoA fields, node ids, ranges and permutation
bit for bit (fmm3d/tree.py:801-1050, SoA :286-339).  The FMM matvec of this
package runs on its own lattice octree (``FmmPlan``); this tree serves callers
of the reference's tree API and its interaction lists
(``interaction_lists.build_interaction_lists``).

Errors follow the reference: points not (N, 3) -> ValueError, not float32 /
float64 -> TypeError, points on a GPU -> RuntimeError (the tree is built on
the CPU and moved with ``FmmTree.to``), leaf_size <= 0 or max_depth outside
[1, 64] -> ValueError.
"""
import ctypes
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np
import torch

from .. import _native as nat
from .geometry import BoundingBox, compute_bounding_box


@dataclass
class TreeNode:
    """Object view of one node (fmm3d/tree.py:60-110)."""
    index: int
    parent: Optional[int]
    level: int
    center: torch.Tensor
    half_extent: torch.Tensor
    start: int
    end: int
    children: Tuple[Optional[int], ...]

    @property
    def is_leaf(self) -> bool:
        return all(c is None for c in self.children)

    @property
    def n_points(self) -> int:
        return self.end - self.start

    @property
    def half_size(self):
        return self.half_extent


@dataclass
class TreeStatistics:
    """fmm3d/tree.py:118-162."""
    n_points: int
    n_nodes: int
    n_leaves: int
    max_depth: int
    nodes_per_level: torch.Tensor
    leaf_size_min: int
    leaf_size_max: int
    leaf_size_mean: float

    def as_dict(self) -> dict:
        return {"n_points": self.n_points, "n_nodes": self.n_nodes, "n_leaves": self.n_leaves,
                "max_depth": self.max_depth, "nodes_per_level": self.nodes_per_level.clone(),
                "leaf_size_min": self.leaf_size_min, "leaf_size_max": self.leaf_size_max,
                "leaf_size_mean": self.leaf_size_mean}


class FmmTree:
    """Adaptive octree in SoA form: ``points`` in tree order, ``index_map``
    (tree -> original), ``inverse_index_map``, ``node_centers``,
    ``node_half_extents``, ``node_radii`` (points' dtype), ``node_levels``
    (int32), ``node_parents``, ``node_children`` (M, 8), ``node_ranges``
    (M, 2) (int64, -1 for none), ``node_is_leaf`` (bool)."""

    def __init__(self, points, *, index_map, soa, leaf_size, max_depth, global_bbox):
        self.points = points
        self.index_map = index_map
        self.inverse_index_map = torch.empty_like(index_map)
        if index_map.numel():
            self.inverse_index_map[index_map] = torch.arange(index_map.shape[0],
                                                             dtype=index_map.dtype)
        (self.node_centers, self.node_half_extents, self.node_radii, self.node_levels,
         self.node_parents, self.node_children, self.node_ranges, self.node_is_leaf) = soa
        self.leaf_size = int(leaf_size)
        self.max_depth = int(max_depth)
        self.global_bbox = global_bbox
        self._nodes = None

    # -- basic properties (fmm3d/tree.py:238-285) --------------------------
    @property
    def device(self):
        return self.points.device

    @property
    def dtype(self):
        return self.points.dtype

    @property
    def n_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.node_levels.shape[0])

    @property
    def root(self) -> int:
        return 0 if self.n_nodes > 0 else -1

    @property
    def tree_to_original(self):
        return self.index_map

    @property
    def original_to_tree(self):
        return self.inverse_index_map

    @property
    def nodes(self) -> List[TreeNode]:
        """TreeNode objects built on first use from the SoA arrays."""
        if self._nodes is None:
            ch = self.node_children.cpu().tolist()
            par = self.node_parents.cpu().tolist()
            lev = self.node_levels.cpu().tolist()
            rng = self.node_ranges.cpu().tolist()
            self._nodes = [TreeNode(index=i, parent=None if par[i] < 0 else par[i], level=lev[i],
                                    center=self.node_centers[i], half_extent=self.node_half_extents[i],
                                    start=rng[i][0], end=rng[i][1],
                                    children=tuple(None if c < 0 else c for c in ch[i]))
                           for i in range(self.n_nodes)]
        return self._nodes

    # -- convenience (fmm3d/tree.py:404-540) -------------------------------
    def leaf_indices(self):
        return torch.nonzero(self.node_is_leaf, as_tuple=False).view(-1).to(torch.int64)

    def leaves(self) -> List[TreeNode]:
        return [self.nodes[int(i)] for i in self.leaf_indices()]

    def map_to_tree_order(self, values):
        if values.shape[0] != self.n_points:
            raise ValueError(f"values.shape[0] = {values.shape[0]} does not match "
                             f"n_points = {self.n_points}")
        return values if self.index_map.numel() == 0 else values[self.index_map.to(values.device)]

    def map_to_original_order(self, values):
        if values.shape[0] != self.n_points:
            raise ValueError(f"values.shape[0] = {values.shape[0]} does not match "
                             f"n_points = {self.n_points}")
        if self.index_map.numel() == 0:
            return values
        return values[self.inverse_index_map.to(values.device)]

    map_from_tree_order = map_to_original_order

    def to(self, device, dtype=None) -> "FmmTree":
        device = torch.device(device)
        dt = self.points.dtype if dtype is None else dtype
        self.points = self.points.to(device=device, dtype=dt)
        self.node_centers = self.node_centers.to(device=device, dtype=dt)
        self.node_half_extents = self.node_half_extents.to(device=device, dtype=dt)
        self.node_radii = self.node_radii.to(device=device, dtype=dt)
        for name in ("index_map", "inverse_index_map", "node_levels", "node_parents",
                     "node_children", "node_ranges", "node_is_leaf"):
            setattr(self, name, getattr(self, name).to(device))
        self.global_bbox = self.global_bbox.to(device=device, dtype=dt)
        self._nodes = None
        return self

    def statistics(self) -> TreeStatistics:
        if self.n_nodes == 0:
            return TreeStatistics(self.n_points, 0, 0, -1, torch.zeros(0, dtype=torch.int64),
                                  0, 0, 0.0)
        lev = self.node_levels.cpu().to(torch.int64)
        sizes = (self.node_ranges[:, 1] - self.node_ranges[:, 0]).cpu()[self.node_is_leaf.cpu()]
        return TreeStatistics(self.n_points, self.n_nodes, int(sizes.numel()), int(lev.max()),
                              torch.bincount(lev), int(sizes.min()), int(sizes.max()),
                              float(sizes.double().mean()))

    def verify(self, *, check_bboxes=True, atol=1e-6, rtol=1e-5) -> None:
        """Invariants of fmm3d/tree.py:618-793; raises ValueError."""
        n, m = self.n_points, self.n_nodes
        if m == 0:
            if n != 0:
                raise ValueError("Empty tree must have zero points.")
            return
        im, inv = self.index_map.cpu(), self.inverse_index_map.cpu()
        if im.shape[0] != n or inv.shape[0] != n:
            raise ValueError("Permutation arrays do not match n_points")
        if not torch.equal(im[inv], torch.arange(n)):
            raise ValueError("index_map and inverse_index_map are not true inverses")
        R = self.node_ranges.cpu()
        s, e = R[:, 0], R[:, 1]
        if (s < 0).any() or (e > n).any() or (s > e).any():
            raise ValueError("Invalid node range detected")
        leaf = self.node_is_leaf.cpu()
        lr = R[leaf]
        lr = lr[torch.argsort(lr[:, 0])]
        if lr.shape[0] == 0 or int(lr[0, 0]) != 0 or int(lr[-1, 1]) != n:
            raise ValueError("Leaf ranges do not cover full [0, N) interval")
        if (lr[1:, 0] != lr[:-1, 1]).any():
            raise ValueError("Leaf ranges are not contiguous / disjoint")
        par, ch, lev = self.node_parents.cpu(), self.node_children.cpu(), self.node_levels.cpu()
        if int(par[0]) != -1:
            raise ValueError("Root node must have parent = -1")
        for i in range(m):
            p = int(par[i])
            if p != -1 and not (0 <= p < m and bool((ch[p] == i).any())):
                raise ValueError(f"Node {i} is not listed among children of its parent {p}")
            kids = [int(c) for c in ch[i] if int(c) >= 0]
            for c in kids:
                if c >= m or int(par[c]) != i:
                    raise ValueError(f"Parent mismatch: node {c} does not point back to {i}")
                if int(lev[c]) != int(lev[i]) + 1:
                    raise ValueError("Child level must be parent level + 1")
            if bool(leaf[i]) == bool(kids):
                raise ValueError(f"Node {i}: leaf flag and children disagree")
            if kids:
                cr = sorted((int(R[c, 0]), int(R[c, 1])) for c in kids)
                if cr[0][0] != int(s[i]) or cr[-1][1] != int(e[i]) or any(
                        a[1] != b[0] or b[1] <= b[0] for a, b in zip(cr, cr[1:])):
                    raise ValueError(f"Children of node {i} do not partition its range")
        if check_bboxes and n > 0:
            P = self.points.cpu()
            C, H = self.node_centers.cpu(), self.node_half_extents.cpu()
            for i in range(m):
                pts = P[int(s[i]):int(e[i])]
                if pts.shape[0] == 0:
                    continue
                bb = BoundingBox.from_min_max(*compute_bounding_box(pts))
                if not torch.allclose(bb.center, C[i], atol=atol, rtol=rtol):
                    raise ValueError(f"Center mismatch for node {i}")
                if not torch.allclose(bb.half_extent, H[i], atol=atol, rtol=rtol):
                    raise ValueError(f"Half-extent mismatch for node {i}")


class _RefTreeHandle:
    """Owns a native ed_ref_tree (kept for building interaction lists)."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            if self.h is not None:
                nat.lib().ed_ref_tree_destroy(self.h)
        except Exception:
            pass


def build_fmm_tree(points, *, leaf_size: int = 64, max_depth: int = 32) -> FmmTree:
    if max_depth <= 0:
        raise ValueError(f"max_depth must be positive, got {max_depth}")
    if max_depth > 64:
        raise ValueError(f"max_depth={max_depth} is too large (at most 64)")
    if leaf_size <= 0:
        raise ValueError(f"leaf_size must be positive, got {leaf_size}")
    if points.device.type != "cpu":
        raise RuntimeError("Tree building must run on CPU. "
                           "Call build_fmm_tree(points.cpu()) and then use tree.to('cuda').")
    if points.ndim != 2 or points.shape[-1] != 3:
        raise ValueError(f"points must have shape (N, 3), got {tuple(points.shape)}")
    if not torch.is_floating_point(points) or points.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"points must be float32 or float64, got {points.dtype}")
    L = nat.lib()
    is32 = points.dtype == torch.float32
    npdt = np.float32 if is32 else np.float64
    P = np.ascontiguousarray(points.detach().numpy().astype(npdt, copy=False))
    if not np.isfinite(P).all():
        raise ValueError("points must be finite")
    h = ctypes.c_void_p()
    nat.check(L.ed_ref_tree_create(P.ctypes.data_as(ctypes.c_void_p), int(is32), P.shape[0],
                                   int(leaf_size), int(max_depth), ctypes.byref(h)),
              "ed_ref_tree_create")
    handle = _RefTreeHandle(h)
    sz = np.zeros(4, dtype=np.int64)
    nat.check(L.ed_ref_tree_sizes(h, sz.ctypes.data_as(ctypes.c_void_p)), "ed_ref_tree_sizes")
    n, m = int(sz[0]), int(sz[1])
    im = np.zeros(n, dtype=np.int64)
    cen = np.zeros((m, 3), dtype=npdt)
    half = np.zeros((m, 3), dtype=npdt)
    rad = np.zeros(m, dtype=npdt)
    lev = np.zeros(m, dtype=np.int32)
    par = np.zeros(m, dtype=np.int64)
    chl = np.zeros((m, 8), dtype=np.int64)
    rng = np.zeros((m, 2), dtype=np.int64)
    leaf = np.zeros(m, dtype=np.uint8)
    vp = ctypes.c_void_p
    nat.check(L.ed_ref_tree_export(h, *(a.ctypes.data_as(vp) for a in
                                        (im, cen, half, rad, lev, par, chl, rng, leaf))),
              "ed_ref_tree_export")
    idx = torch.from_numpy(im)
    pts = points[idx] if n else points
    if n:
        gb = BoundingBox.from_min_max(*compute_bounding_box(points))
    else:
        z = torch.zeros(3, dtype=points.dtype)
        gb = BoundingBox(z, z.clone())
    soa = (torch.from_numpy(cen), torch.from_numpy(half), torch.from_numpy(rad),
           torch.from_numpy(lev), torch.from_numpy(par), torch.from_numpy(chl),
           torch.from_numpy(rng), torch.from_numpy(leaf.astype(bool)))
    tree = FmmTree(pts, index_map=idx, soa=soa, leaf_size=leaf_size, max_depth=max_depth,
                   global_bbox=gb)
    tree._native = handle
    return tree


__all__ = ["TreeNode", "TreeStatistics", "FmmTree", "build_fmm_tree"]
