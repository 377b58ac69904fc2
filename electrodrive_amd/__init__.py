"""MI355X-native (gfx950) BEM/FMM solve path of Electrodrive.

The hot path — the GMRES-driven single-layer Laplace collocation matvec
evaluated by a 3D Laplace FMM plus singular / near-singular panel quadrature —
runs in hand-written HIP kernels (``libednative.so``, C ABI in
include/ednative.h).  Python keeps the reference call signatures:

- ``electrodrive_amd.fmm3d``: FmmConfig, LaplaceFmm3D, make_laplace_fmm_backend,
  morton_keys_from_points, BoundingBox, build_fmm_tree / FmmTree (the
  reference's adaptive octree), build_interaction_lists / InteractionLists
- ``electrodrive_amd.core.bem_kernel``: bem_matvec_gpu, bem_matvec_batched
- ``electrodrive_amd.core.bem_solver``: gmres_restart, setup_preconditioner
- ``electrodrive_amd.core.bem_quadrature``: self integrals, near pairs,
  near weights, near_quadrature_matvec
- ``electrodrive_amd.core.operator.BemOperator``: GPU set-up + FP64 GMRES with
  iterative refinement
"""
from ._native import LIB_PATH, lib  # noqa: F401

__version__ = "0.1.0"
