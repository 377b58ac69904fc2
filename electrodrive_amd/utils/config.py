"""``BEMConfig`` of ``bem_solve`` (electrodrive/utils/config.py:79-153): the
reference's fields with the same defaults, so configurations written for the
reference run unchanged, plus the knobs of the native FMM path.

Reference fields that size the reference's dense tiles (vram_autotune,
target_vram_fraction, vram_cap_gb, target_peak_gb, min_tile, max_tile,
tile_size, tile_mem_divisor) are accepted and recorded; the native operator
has no tiles to size.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

from ..core.kernels import EPS_0, K_E  # noqa: F401  (one definition of the constants)


@dataclass
class BEMConfig:
    use_gpu: bool = True
    fp64: bool = False
    initial_h: float = 0.2
    max_refine_passes: int = 3
    target_bc_inf_norm: float = 1e-7
    linear_tol_factor: float = 0.1
    gmres_tol: float = 5e-8
    gmres_maxiter: int = 2000
    gmres_restart: int = 256
    gmres_log_every: int = 10
    vram_autotune: bool = True
    target_vram_fraction: float = 0.95
    vram_cap_gb: float = 24.0
    target_peak_gb: float = 20.0
    min_tile: int = 1024
    max_tile: int = 1 << 18
    tile_size: int = 0
    tile_mem_divisor: float = 2.0
    use_near_quadrature: bool = False
    use_near_quadrature_matvec: bool = False
    near_quadrature_distance_factor: float = 1.5
    near_quadrature_order: int = 2
    # refinement control read by bem_solve through getattr in the reference
    # (core/bem.py:1580-1625, :2440-2470)
    refine_factor: float = 0.5
    min_panels: int = 0
    min_refine_passes: int = 1
    plateau_max: Optional[int] = None
    plateau_rel_improvement: float = 0.05
    use_precond: bool = True
    # failure recovery of bem_solve (core/bem.py:1604-1607): after a failed
    # preconditioned solve, one retry without the preconditioner, then the
    # damped Jacobi fall-back solve
    allow_jacobi_fallback: bool = True
    jacobi_maxiter: int = 256
    jacobi_tol: float = 1e-5
    jacobi_omega: float = 0.8
    # native FMM operator: expansion order and leaf size (None: the
    # autotuner's choice for the mesh at a 1e-5 relative error bound, or the
    # FP64 dense operator below fmm_min_panels panels), M2L precision
    fmm_order: Optional[int] = None
    fmm_leaf_size: Optional[int] = None
    fmm_min_panels: int = 4096
    m2l_precision: str = "fp16x3"
    precond: str = "sai"
