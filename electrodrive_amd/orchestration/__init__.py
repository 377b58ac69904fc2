"""Spec ingest (electrodrive/orchestration/parser.py)."""
from .parser import CanonicalSpec, parse_spec  # noqa: F401

__all__ = ["CanonicalSpec", "parse_spec"]
