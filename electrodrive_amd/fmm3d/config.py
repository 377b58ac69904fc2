"""FMM configuration (mirror of fmm3d/config.py:32-180 of the reference).

Same field names, defaults and validation rules (p > 0, theta in [0.1, 0.9],
leaf_size > 0, 2p <= 32, dtype float32/float64).  Two fields are added for the
lattice FMM of this package:

- ``ws``: well-separation radius in boxes (1 or 2).  When left as None it is
  derived from ``mac_theta``: the reference MAC max(2r_s, 2r_t)/dist < theta
  (fmm3d/interaction_lists.py:273-297) is met by lattice boxes one box apart
  only for theta > 0.866, so theta >= 0.75 maps to ws = 1 and smaller theta to
  ws = 2.
- ``m2l_precision``: "fp32" (MFMA fp32), "bf16" / "fp16" (MFMA with fp32
  accumulation), "fp16x3" (split FP16: table and multipoles as hi + lo FP16
  pairs, three FP16 MFMAs per product, FP32-level accuracy) or "valu" (plain
  fp32 vector path).
"""
from dataclasses import dataclass
from typing import Literal, Optional

import torch

FMM_L_MAX_HARD_LIMIT: int = 32
FMM_M2L_LMAX_FACTOR: int = 2

BackendKind = Literal["cpu", "gpu", "auto"]
PrecisionKind = Literal["single", "double", "mixed"]

# orders 14 and 16 run the VALU M2L (their tables do not fit the LDS)
SUPPORTED_ORDERS = (4, 6, 8, 10, 12, 14, 16)
M2L_PRECISIONS = {"valu": -1, "fp32": 0, "bf16": 1, "fp16": 2, "fp16x3": 3}


@dataclass
class FmmConfig:
    kernel: str = "laplace_single_layer"
    expansion_order: int = 8
    mac_theta: float = 0.5
    leaf_size: int = 64

    backend: BackendKind = "auto"
    precision: PrecisionKind = "mixed"
    use_fft_m2l: bool = True
    use_gpu: bool = True
    use_multi_gpu: bool = False
    use_mpi: bool = False

    dtype: Optional[torch.dtype] = None
    p2p_batch_size: Optional[int] = None

    ws: Optional[int] = None
    max_level: int = 16
    m2l_precision: str = "fp16x3"

    def __post_init__(self) -> None:
        if self.dtype is None:
            self.dtype = torch.float64 if self.precision == "double" else torch.float32
        if self.ws is None:
            self.ws = 1 if self.mac_theta >= 0.75 else 2
        self.validate()

    def validate(self) -> None:
        if self.expansion_order <= 0:
            raise ValueError("expansion_order must be positive")
        if not (0.1 <= self.mac_theta <= 0.9):
            raise ValueError("mac_theta should be in [0.1, 0.9] for stability")
        if self.leaf_size <= 0:
            raise ValueError("leaf_size must be positive")
        if FMM_M2L_LMAX_FACTOR * self.expansion_order > FMM_L_MAX_HARD_LIMIT:
            raise ValueError("expansion_order too large: 2p must be <= 32")
        if self.dtype not in (torch.float32, torch.float64):
            raise ValueError(f"dtype must be torch.float32 or torch.float64; got {self.dtype!r}")
        if self.p2p_batch_size is not None and self.p2p_batch_size <= 0:
            raise ValueError("p2p_batch_size must be positive if specified")
        if self.ws not in (1, 2):
            raise ValueError("ws must be 1 or 2")
        if self.m2l_precision not in M2L_PRECISIONS:
            raise ValueError(f"m2l_precision must be one of {sorted(M2L_PRECISIONS)}")

    @property
    def native_order(self) -> int:
        """Smallest supported native order >= expansion_order."""
        for p in SUPPORTED_ORDERS:
            if p >= self.expansion_order:
                return p
        raise ValueError(f"expansion_order {self.expansion_order} > {SUPPORTED_ORDERS[-1]}")


__all__ = ["FMM_L_MAX_HARD_LIMIT", "FMM_M2L_LMAX_FACTOR", "BackendKind", "PrecisionKind",
           "FmmConfig", "SUPPORTED_ORDERS", "M2L_PRECISIONS"]
