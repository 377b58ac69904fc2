"""Reference-compatible BEM layer (core/bem_kernel.py, core/bem_solver.py,
core/bem_quadrature.py, core/bem_mesh.py of Electrodrive)."""
from .kernels import DEFAULT_SINGLE_LAYER_KERNEL, EPS_0, K_E, LaplaceSingleLayerKernel  # noqa: F401
