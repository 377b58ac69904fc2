"""Morton (Z-order) keys restated from fmm3d/geometry.py:671-799 (numpy).

Test infrastructure only (see oracle/__init__.py).

Arithmetic follows the reference step by step in the points' own dtype
(float32 or float64): tight bbox (geometry.py:125-157), degenerate-axis padding
eps_rel*(1+max_extent) (:678-693), normalisation (p-min)/max(extent, eps)
clamped to [0, 1-1e-9] in that dtype (:696-703), floor(x*2^level) clamped to
[0, 2^level-1] (:785-789), bit interleave x->3b, y->3b+1, z->3b+2 (:705-733).
When a bounding box (center, half) is given, mins/maxs are center -/+ half
(geometry.py:261-268), as the reference does for ``global_bbox``.

Parity: pinned by hand-computed keys and the FP32 clamp case
(tests/test_cpu_oracle.py::test_morton_hand_keys) and by the ordering property
(keys sort points into nested octants).
"""
import numpy as np


def interleave(ix, iy, iz, level):
    """geometry.py:705-733."""
    if level <= 0 or level > 21:
        raise ValueError(f"level must be in [1, 21], got {level}")
    ix = np.asarray(ix, dtype=np.int64)
    iy = np.asarray(iy, dtype=np.int64)
    iz = np.asarray(iz, dtype=np.int64)
    keys = np.zeros_like(ix)
    for bit in range(level):
        keys |= ((ix >> bit) & 1) << (3 * bit)
        keys |= ((iy >> bit) & 1) << (3 * bit + 1)
        keys |= ((iz >> bit) & 1) << (3 * bit + 2)
    return keys


def global_bbox(points):
    """geometry.py:678-693: (mins, padded maxs) in the points' dtype."""
    p = np.asarray(points)
    dt = p.dtype.type
    mins = p.min(axis=0)
    maxs = p.max(axis=0)
    extent = (maxs - mins).astype(p.dtype)
    max_extent = np.max(np.abs(extent))
    padded = dt(1e-12) * (dt(1.0) + max_extent)
    extent = np.where(extent > 0, extent, padded).astype(p.dtype)
    return mins, (mins + extent).astype(p.dtype)


def morton_keys(points, level=16, bbox_center=None, bbox_half=None):
    """geometry.py:736-799.  ``points`` float32 or float64 (N,3)."""
    p = np.asarray(points)
    if p.dtype not in (np.float32, np.float64):
        raise TypeError("points must be float32 or float64")
    dt = p.dtype.type
    if p.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    if bbox_center is None:
        mins, maxs = global_bbox(p)
    else:
        c = np.asarray(bbox_center, dtype=p.dtype)
        h = np.asarray(bbox_half, dtype=p.dtype)
        mins, maxs = (c - h).astype(p.dtype), (c + h).astype(p.dtype)
    extent = (maxs - mins).astype(p.dtype)
    eps = np.finfo(p.dtype).eps
    extent = np.maximum(extent, dt(eps))
    normalized = ((p - mins) / extent).astype(p.dtype)
    hi = dt(1.0 - 1e-9)
    normalized = np.minimum(np.maximum(normalized, dt(0.0)), hi)
    scale = dt(float(1 << level))
    ijk = np.floor(normalized * scale).astype(np.int64)
    ijk = np.clip(ijk, 0, (1 << level) - 1)
    return interleave(ijk[:, 0], ijk[:, 1], ijk[:, 2], level)
