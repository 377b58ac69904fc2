"""Reference-mode interaction lists (mirror of fmm3d/interaction_lists.py).

``build_interaction_lists(source_tree, target_tree=None, mac_theta=0.5, *,
max_pairs=None, sort_pairs=True)`` runs the reference's dual-tree traversal
natively (``ed_ref_lists_create``) over trees from ``tree.build_fmm_tree``
and returns ``InteractionLists`` with the same (src, tgt) pair sets and order
(interaction_lists.py:374-577): M2L pairs by the MAC
max(2 r_s, 2 r_t) / dist < theta, leaf-leaf pairs (self pairs included) to
P2P, per-target u/v/w/x lists (v = w = x, as in the reference) and the
self-pair mask.  mac_theta outside (0, 1) -> ValueError; a truncated
traversal (max_pairs reached) -> RuntimeError; trees not on the CPU ->
RuntimeError.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from .. import _native as nat
from .tree import FmmTree


@dataclass
class InteractionStats:
    """interaction_lists.py:59-88."""
    num_p2p_pairs: int = 0
    num_m2l_pairs: int = 0
    num_stack_pops: int = 0
    max_stack_size: int = 0
    num_mac_accept: int = 0
    num_mac_reject: int = 0
    num_leaf_leaf: int = 0
    truncated: bool = False

    def as_dict(self) -> Dict[str, object]:
        return dict(self.__dict__)


@dataclass
class InteractionLists:
    """interaction_lists.py:96-191: node-pair lists as (src, tgt)."""
    p2p_pairs: List[Tuple[int, int]]
    m2l_pairs: List[Tuple[int, int]]
    mac_theta: float
    u_list: List[List[int]]
    v_list: List[List[int]]
    w_list: List[List[int]]
    x_list: List[List[int]]
    u_self_mask: Optional[List[bool]] = field(default=None, repr=False)
    truncated: bool = False
    stats: Optional[InteractionStats] = field(default=None, repr=False)
    _p2p_arr: Optional[np.ndarray] = field(default=None, repr=False)
    _m2l_arr: Optional[np.ndarray] = field(default=None, repr=False)

    @property
    def num_p2p_pairs(self) -> int:
        return len(self.p2p_pairs)

    @property
    def num_m2l_pairs(self) -> int:
        return len(self.m2l_pairs)

    def as_tensors(self, device=None):
        """(p2p_pairs, m2l_pairs) as int64 tensors of shape (K, 2)."""
        out = []
        for arr, pairs in ((self._p2p_arr, self.p2p_pairs), (self._m2l_arr, self.m2l_pairs)):
            t = (torch.from_numpy(arr.copy()) if arr is not None
                 else torch.tensor(pairs, dtype=torch.long).reshape(-1, 2))
            out.append(t if device is None else t.to(torch.device(device)))
        return out[0], out[1]

    def to(self, device) -> "InteractionLists":
        self.as_tensors(device=device)
        return self


def _assert_cpu(tree: FmmTree, name: str):
    if tree.node_centers.device.type != "cpu":
        raise RuntimeError(f"interaction_lists: {name} must reside on CPU "
                           f"(got device={tree.node_centers.device}).")
    if getattr(tree, "_native", None) is None:
        raise ValueError(f"interaction_lists: {name} must come from build_fmm_tree")


def build_interaction_lists(source_tree: FmmTree, target_tree: Optional[FmmTree] = None,
                            mac_theta: float = 0.5, *, max_pairs: Optional[int] = None,
                            sort_pairs: bool = True) -> InteractionLists:
    if target_tree is None:
        target_tree = source_tree
    _assert_cpu(source_tree, "source_tree")
    _assert_cpu(target_tree, "target_tree")
    if mac_theta <= 0.0:
        raise ValueError("mac_theta must be positive")
    if mac_theta >= 1.0:
        raise ValueError("mac_theta should be < 1.0 for a meaningful MAC")
    same = source_tree is target_tree
    L = nat.lib()
    h = ctypes.c_void_p()
    nat.check(L.ed_ref_lists_create(source_tree._native.h,
                                    None if same else target_tree._native.h, float(mac_theta),
                                    -1 if max_pairs is None else int(max_pairs),
                                    int(bool(sort_pairs)), ctypes.byref(h)),
              "ed_ref_lists_create")
    try:
        sz = np.zeros(8, dtype=np.int64)
        nat.check(L.ed_ref_lists_sizes(h, sz.ctypes.data_as(ctypes.c_void_p)), "ed_ref_lists_sizes")
        p2p = np.zeros((int(sz[0]), 2), dtype=np.int64)
        m2l = np.zeros((int(sz[1]), 2), dtype=np.int64)
        nat.check(L.ed_ref_lists_export(h, p2p.ctypes.data_as(ctypes.c_void_p),
                                        m2l.ctypes.data_as(ctypes.c_void_p)),
                  "ed_ref_lists_export")
    finally:
        L.ed_ref_lists_destroy(h)
    stats = InteractionStats(num_p2p_pairs=int(sz[0]), num_m2l_pairs=int(sz[1]),
                             num_stack_pops=int(sz[5]), max_stack_size=int(sz[6]),
                             num_mac_accept=int(sz[2]), num_mac_reject=int(sz[3]),
                             num_leaf_leaf=int(sz[4]), truncated=bool(sz[7]))
    if stats.truncated:
        raise RuntimeError("Interaction list truncated—MAC is likely misconfigured or pair "
                           "count exploded.")
    m_t = target_tree.n_nodes
    u = [[] for _ in range(m_t)]
    v = [[] for _ in range(m_t)]
    mask = [False] * m_t
    for s, t in p2p.tolist():
        u[t].append(s)
        if same and s == t:
            mask[t] = True
    for s, t in m2l.tolist():
        v[t].append(s)
    return InteractionLists(
        p2p_pairs=[tuple(x) for x in p2p.tolist()], m2l_pairs=[tuple(x) for x in m2l.tolist()],
        mac_theta=float(mac_theta), u_list=u, v_list=v, w_list=[list(x) for x in v],
        x_list=[list(x) for x in v], u_self_mask=mask, truncated=False, stats=stats,
        _p2p_arr=p2p, _m2l_arr=m2l)


def verify_interaction_lists(source_tree: FmmTree, target_tree: FmmTree,
                             lists: InteractionLists) -> None:
    """interaction_lists.py:580-621; raises AssertionError."""
    ms, mt = source_tree.n_nodes, target_tree.n_nodes
    for s, t in lists.p2p_pairs:
        assert 0 <= s < ms and 0 <= t < mt, "p2p index out of range"
    for s, t in lists.m2l_pairs:
        assert 0 <= s < ms and 0 <= t < mt, "m2l index out of range"
    p2p, m2l = set(lists.p2p_pairs), set(lists.m2l_pairs)
    for t in range(mt):
        assert all((s, t) in p2p for s in lists.u_list[t]), "u_list not subset of p2p_pairs"
        for lst in (lists.v_list, lists.w_list, lists.x_list):
            assert all((s, t) in m2l for s in lst[t]), "v/w/x list not subset of m2l_pairs"
    if source_tree is target_tree and lists.u_self_mask is not None:
        for t in range(mt):
            assert lists.u_self_mask[t] == ((t, t) in p2p), "u_self_mask mismatch"


__all__ = ["InteractionStats", "InteractionLists", "build_interaction_lists",
           "verify_interaction_lists"]
