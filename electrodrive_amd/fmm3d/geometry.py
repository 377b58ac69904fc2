"""Bounding boxes and Morton keys (mirror of fmm3d/geometry.py of the reference).

- ``BoundingBox`` (geometry.py:206-270): centre and half extent in the points'
  dtype, ``from_min_max`` (centre 0.5 (min + max), half clamped >= 0),
  ``from_points``, ``mins`` / ``maxs`` = centre -/+ half.
- ``compute_bounding_box`` (geometry.py:125-157): tight (mins, maxs).
- ``morton_keys_from_points(points, level=16, global_bbox=None)``
  (geometry.py:736-799) returns ``(keys, bbox)``: keys bit-exact with the
  reference, computed by the native kernel; ``bbox`` is the normalisation box
  (``global_bbox`` when given, else the padded tight box of the points).  A
  (center, half_extent) pair is also accepted for ``global_bbox``.
"""
import ctypes
from dataclasses import dataclass
from typing import Tuple

import numpy as np
import torch

from .. import _native as nat


def _check_points(points):
    if points.ndim != 2 or points.shape[-1] != 3:
        raise ValueError(f"points must have shape (N, 3), got {tuple(points.shape)}")
    if points.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"points must have dtype float32 or float64, got {points.dtype}")


@dataclass
class BoundingBox:
    center: torch.Tensor
    half_extent: torch.Tensor

    def __post_init__(self):
        if self.center.shape != (3,) or self.half_extent.shape != (3,):
            raise ValueError("BoundingBox center and half_extent must have shape (3,)")

    @classmethod
    def from_min_max(cls, mins, maxs):
        if mins.shape != (3,) or maxs.shape != (3,):
            raise ValueError(f"mins/maxs must have shape (3,), got {tuple(mins.shape)}, "
                             f"{tuple(maxs.shape)}")
        center = 0.5 * (mins + maxs)
        half = torch.clamp(0.5 * (maxs - mins), min=0.0)
        return cls(center=center, half_extent=half)

    @classmethod
    def from_points(cls, points):
        return cls.from_min_max(*compute_bounding_box(points))

    @property
    def mins(self):
        return self.center - self.half_extent

    @property
    def maxs(self):
        return self.center + self.half_extent

    @property
    def device(self):
        return self.center.device

    @property
    def dtype(self):
        return self.center.dtype

    def to(self, device=None, dtype=None):
        return BoundingBox(self.center.to(device=device, dtype=dtype),
                           self.half_extent.to(device=device, dtype=dtype))


def compute_bounding_box(points) -> Tuple[torch.Tensor, torch.Tensor]:
    _check_points(points)
    if points.shape[0] == 0:
        z = torch.zeros(3, dtype=points.dtype, device=points.device)
        return z.clone(), z.clone()
    if not bool(torch.isfinite(points).all()):
        raise ValueError("points must be finite")
    return points.min(dim=0).values, points.max(dim=0).values


def _padded_bbox(points):
    """The normalisation box the reference derives when none is given
    (geometry.py:671-693): degenerate axes padded by 1e-12 (1 + max extent)."""
    mins, maxs = compute_bounding_box(points)
    ext = maxs - mins
    pad = points.new_tensor(1e-12) * (1.0 + ext.abs().max())
    ext = torch.where(ext > 0, ext, pad)
    return BoundingBox.from_min_max(mins, mins + ext)


def morton_keys_from_points(points, *, level: int = 16, global_bbox=None):
    _check_points(points)
    if not (1 <= level <= 21):
        raise ValueError(f"level must be in [1, 21], got {level}")
    if global_bbox is not None and not isinstance(global_bbox, BoundingBox):
        c, h = global_bbox
        global_bbox = BoundingBox(torch.as_tensor(c, dtype=points.dtype),
                                  torch.as_tensor(h, dtype=points.dtype))
    if points.shape[0] == 0:
        if global_bbox is None:
            z = torch.zeros(3, dtype=points.dtype, device=points.device)
            global_bbox = BoundingBox(z, z.clone())
        return torch.zeros(0, dtype=torch.int64, device=points.device), global_bbox
    nat.require_gpu()
    dev = points.device if points.is_cuda else torch.device("cuda", torch.cuda.current_device())
    P = points.to(dev).contiguous()
    if not bool(torch.isfinite(P).all()):
        raise ValueError("points must be finite")
    n = P.shape[0]
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    bc = bh = None
    if global_bbox is not None:
        npdt = np.float64 if points.dtype == torch.float64 else np.float32
        bc = np.ascontiguousarray(global_bbox.center.detach().cpu().numpy().astype(npdt))
        bh = np.ascontiguousarray(global_bbox.half_extent.detach().cpu().numpy().astype(npdt))
    fn = nat.lib().ed_morton_keys_f64 if points.dtype == torch.float64 else nat.lib().ed_morton_keys_f32
    nat.check(fn(nat.ptr(P), n, int(level),
                 None if bc is None else bc.ctypes.data_as(ctypes.c_void_p),
                 None if bh is None else bh.ctypes.data_as(ctypes.c_void_p),
                 nat.ptr(keys), nat.stream_ptr(dev)), "ed_morton_keys")
    if global_bbox is None:
        global_bbox = _padded_bbox(points)
    return keys.to(points.device), global_bbox


__all__ = ["BoundingBox", "compute_bounding_box", "morton_keys_from_points"]
