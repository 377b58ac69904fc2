"""Reference-compatible FMM package (electrodrive/fmm3d/__init__.py:21-24)."""
from .config import FmmConfig  # noqa: F401
from .bem_fmm import LaplaceFmm3D, make_laplace_fmm_backend  # noqa: F401
from .geometry import BoundingBox, compute_bounding_box, morton_keys_from_points  # noqa: F401
from .tree import FmmTree, TreeNode, TreeStatistics, build_fmm_tree  # noqa: F401
from .interaction_lists import (InteractionLists, InteractionStats,  # noqa: F401
                                build_interaction_lists, verify_interaction_lists)
from .plan import FmmPlan  # noqa: F401
from .autotune import TuningResult, autotune_for_problem  # noqa: F401

__all__ = ["FmmConfig", "LaplaceFmm3D", "make_laplace_fmm_backend", "morton_keys_from_points",
           "BoundingBox", "compute_bounding_box", "FmmTree", "TreeNode", "TreeStatistics",
           "build_fmm_tree", "InteractionLists", "InteractionStats", "build_interaction_lists",
           "verify_interaction_lists", "FmmPlan", "TuningResult", "autotune_for_problem"]
