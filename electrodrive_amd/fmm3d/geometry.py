"""Morton keys on the GPU (mirror of fmm3d/geometry.py:736-799).

``morton_keys_from_points(points, level=16, global_bbox=None)`` returns
``(keys, (center, half_extent))`` with keys bit-exact to the reference
(same dtype arithmetic, clamps and bit interleave).  ``global_bbox`` may be a
(center, half_extent) pair; mins/maxs are then center -/+ half as in the
reference BoundingBox (geometry.py:261-268).
"""
import torch

from .. import _native as nat


def morton_keys_from_points(points, *, level: int = 16, global_bbox=None):
    if points.ndim != 2 or points.shape[-1] != 3:
        raise ValueError(f"points must have shape (N, 3), got {tuple(points.shape)}")
    if points.dtype not in (torch.float32, torch.float64):
        raise TypeError("points must be float32 or float64")
    if not (1 <= level <= 21):
        raise ValueError(f"level must be in [1, 21], got {level}")
    nat.require_gpu()
    dev = points.device if points.is_cuda else torch.device("cuda", torch.cuda.current_device())
    P = points.to(dev).contiguous()
    if not bool(torch.isfinite(P).all()):
        raise ValueError("points must be finite")
    n = P.shape[0]
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    bc = bh = None
    if global_bbox is not None:
        c, h = global_bbox
        import numpy as np
        npdt = np.float64 if points.dtype == torch.float64 else np.float32
        bc = np.ascontiguousarray(torch.as_tensor(c).cpu().numpy().astype(npdt))
        bh = np.ascontiguousarray(torch.as_tensor(h).cpu().numpy().astype(npdt))
    fn = nat.lib().ed_morton_keys_f64 if points.dtype == torch.float64 else nat.lib().ed_morton_keys_f32
    import ctypes
    nat.check(fn(nat.ptr(P), n, int(level),
                 None if bc is None else bc.ctypes.data_as(ctypes.c_void_p),
                 None if bh is None else bh.ctypes.data_as(ctypes.c_void_p),
                 nat.ptr(keys), nat.stream_ptr(dev)), "ed_morton_keys")
    return keys.to(points.device), global_bbox
