"""Drop-in FMM backend for the BEM matvec (mirror of fmm3d/bem_fmm.py).

``LaplaceFmm3D`` / ``make_laplace_fmm_backend`` keep the reference signatures
(fmm3d/bem_fmm.py:127-140, :888-940) and argument rules:

- geometry tensors must be CPU, shapes (N,3) and (N,) (bem_fmm.py:169-181);
- ``matvec(*, sigma, src_centroids, areas, tile_size, self_integrals, logger,
  kernel, **kw)`` requires the very same geometry tensors (data_ptr identity,
  bem_fmm.py:329-349), sigma dtype == geometry dtype (:810-814), CUDA sigma on
  the plan's device (:794-807), only the Laplace single-layer kernel (:783-786),
  self_integrals with sigma's shape/device/dtype (:815-827);
- it returns a new (N,) tensor on sigma's device, in original panel order:
  K_E * sum_{j != i} sigma_j A_j / r_ij + self_integrals * sigma * A, with
  non-finite entries clamped to 0 (bem_fmm.py:718-749).

The work runs in the native plan (HIP kernels for gfx950).  ``backend="cpu"``
is rejected: there is no CPU implementation in this package.
"""
from typing import Any, Optional

import torch

from .. import _native as nat
from ..core.kernels import LaplaceSingleLayerKernel
from .config import FmmConfig
from .plan import FmmPlan


class LaplaceFmm3D:
    def __init__(self, src_centroids, areas, *, max_leaf_size: int = 64, theta: float = 0.5,
                 use_dipole: bool = True, logger: Optional[Any] = None,
                 expansion_order: Optional[int] = None, backend: str = "auto",
                 device=None, ws: Optional[int] = None, m2l_precision: str = "fp16x3",
                 max_level: int = 16):
        if src_centroids.ndim != 2 or src_centroids.shape[1] != 3:
            raise ValueError("src_centroids must have shape (N, 3)")
        if areas.ndim != 1 or areas.shape[0] != src_centroids.shape[0]:
            raise ValueError("areas must have shape (N,) matching src_centroids")
        if src_centroids.device.type != "cpu" or areas.device.type != "cpu":
            raise ValueError("LaplaceFmm3D currently supports only CPU tensors for "
                             "src_centroids and areas.")
        if backend not in ("auto", "gpu", "cpu"):
            raise ValueError(f"unknown backend {backend!r}")
        if backend == "cpu":
            raise RuntimeError("electrodrive_amd has no CPU FMM backend; use backend='gpu'")
        if not torch.cuda.is_available():
            raise RuntimeError("GPU backend requested but no HIP device is available.")
        self.src_centroids = src_centroids
        self.areas = areas
        self.N = int(src_centroids.shape[0])
        self.max_leaf_size = int(max_leaf_size)
        self.theta = float(theta)
        self.use_dipole = bool(use_dipole)
        self.logger = logger
        self.device = src_centroids.device
        self.dtype = src_centroids.dtype
        kw = dict(mac_theta=self.theta, leaf_size=self.max_leaf_size, dtype=self.dtype,
                  backend="gpu", ws=ws, m2l_precision=m2l_precision, max_level=max_level)
        if expansion_order is not None:
            kw["expansion_order"] = int(expansion_order)
        self.cfg = FmmConfig(**kw)
        self.backend = "gpu"
        self.fmm_device = torch.device(device or "cuda")
        if self.fmm_device.index is None:
            self.fmm_device = torch.device("cuda", torch.cuda.current_device())
        self.plan: Optional[FmmPlan] = None
        if self.N > 0:
            with torch.cuda.device(self.fmm_device):
                self.plan = FmmPlan(src_centroids.detach().double().numpy(),
                                    areas.detach().double().numpy(),
                                    leaf_size=self.cfg.leaf_size,
                                    order=self.cfg.native_order, ws=self.cfg.ws,
                                    max_level=self.cfg.max_level,
                                    m2l_precision=self.cfg.m2l_precision)
        self._areas_gpu = None

    # -- helpers ------------------------------------------------------------
    def _check_geometry_consistency(self, src_centroids, areas):
        if src_centroids.data_ptr() != self.src_centroids.data_ptr():
            raise ValueError("LaplaceFmm3D.matvec called with different src_centroids than the "
                             "ones used to build the FMM tree.")
        if areas.data_ptr() != self.areas.data_ptr():
            raise ValueError("LaplaceFmm3D.matvec called with different areas than the ones used "
                             "to build the FMM tree.")
        if src_centroids.device.type != "cpu" or areas.device.type != "cpu":
            raise ValueError("LaplaceFmm3D currently requires matvec geometry tensors to live on "
                             "the CPU.")
        if src_centroids.dtype != self.dtype or areas.dtype != self.dtype:
            raise ValueError("LaplaceFmm3D geometry dtype mismatch: expected "
                             f"{self.dtype}, got {src_centroids.dtype} / {areas.dtype}.")

    def stats(self):
        return self.plan.stats() if self.plan is not None else {}

    # -- public matvec --------------------------------------------------------
    def matvec(self, *, sigma, src_centroids, areas, tile_size, self_integrals,
               logger=None, kernel=None, **kwargs):
        del kwargs, tile_size
        self._check_geometry_consistency(src_centroids, areas)
        if kernel is not None and not isinstance(kernel, LaplaceSingleLayerKernel):
            raise ValueError("LaplaceFmm3D currently only supports LaplaceSingleLayerKernel.")
        N = int(sigma.shape[0])
        if N != self.N:
            raise ValueError(f"sigma has length {N}, but FMM backend was built for {self.N} panels.")
        dev = sigma.device
        if dev.type == "cuda":
            if dev != self.fmm_device:
                raise ValueError(f"LaplaceFmm3D.matvec CUDA sigma must live on {self.fmm_device}, "
                                 f"got {dev}.")
        elif dev.type != "cpu":
            raise ValueError("LaplaceFmm3D.matvec only supports CPU or CUDA sigma.")
        if sigma.dtype != self.dtype:
            raise ValueError(f"LaplaceFmm3D sigma dtype mismatch: expected {self.dtype}, "
                             f"got {sigma.dtype}.")
        if self_integrals is not None:
            if self_integrals.shape != sigma.shape:
                raise ValueError("self_integrals must have shape (N,) matching sigma.")
            if self_integrals.device != dev:
                raise ValueError("self_integrals must be on the same device as sigma "
                                 f"(sigma={dev}, self_integrals={self_integrals.device}).")
            if self_integrals.dtype != sigma.dtype:
                raise ValueError("self_integrals must have the same dtype as sigma "
                                 f"(sigma={sigma.dtype}, self_integrals={self_integrals.dtype}).")
        if N == 0:
            return torch.zeros_like(sigma)
        s_dev = sigma.to(self.fmm_device, non_blocking=True).contiguous()
        si_dev = (None if self_integrals is None
                  else self_integrals.to(self.fmm_device, non_blocking=True).contiguous())
        out = self.plan.matvec(s_dev, si_dev)
        if dev.type == "cpu":
            out = out.to(dev)
        return out


def make_laplace_fmm_backend(src_centroids, areas, *, max_leaf_size: int = 64, theta: float = 0.5,
                             use_dipole: bool = True, logger: Optional[Any] = None,
                             expansion_order: Optional[int] = None, backend: str = "auto",
                             device=None, **kw) -> LaplaceFmm3D:
    """fmm3d/bem_fmm.py:888-940."""
    return LaplaceFmm3D(src_centroids, areas, max_leaf_size=max_leaf_size, theta=theta,
                        use_dipole=use_dipole, logger=logger, expansion_order=expansion_order,
                        backend=backend, device=device, **kw)


__all__ = ["LaplaceFmm3D", "make_laplace_fmm_backend"]
_ = nat
