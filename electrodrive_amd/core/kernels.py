"""Kernel descriptors and constants (core/bem_kernel.py:123-141,
electrodrive/utils/config.py:34-35)."""
import math
from dataclasses import dataclass

EPS_0: float = 8.854187817e-12
K_E: float = 1.0 / (4.0 * math.pi * EPS_0)


@dataclass
class LaplaceSingleLayerKernel:
    """Free-space Laplace single-layer kernel G(r) = K_E / |r|."""

    name: str = "laplace_single_layer"


DEFAULT_SINGLE_LAYER_KERNEL = LaplaceSingleLayerKernel()
