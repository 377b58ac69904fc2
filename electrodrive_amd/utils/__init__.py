"""Configuration and logging of the solve path (electrodrive/utils)."""
from .config import BEMConfig, K_E  # noqa: F401
from .logging import JsonlLogger  # noqa: F401
from .telemetry import TelemetryWriter  # noqa: F401

__all__ = ["BEMConfig", "K_E", "JsonlLogger", "TelemetryWriter"]
