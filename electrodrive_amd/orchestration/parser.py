"""Problem specifications: the reference's ``CanonicalSpec`` and ``parse_spec``
(electrodrive/orchestration/parser.py:31-182), same fields, coercions and
errors, so spec files and dicts written for the reference load unchanged.

A spec is a dict (or a .json / .yaml / .yml file holding one) with
``domain`` (required), ``conductors`` (type plane / sphere / torus / toroid /
cylinder with their geometry, optional ``id`` and ``potential``), ``charges``
(``{"type": "point", "q": ..., "pos": [x, y, z]}``), ``dielectrics``, ``BCs``
(default "Dirichlet"), ``symmetry``, ``queries`` and ``symbols``.
"""
from __future__ import annotations

import copy
import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Dict, List, Mapping, Tuple, Union

import yaml

SpecLike = Union[str, Path, Mapping[str, Any]]


def _as_float(x) -> float:
    """Numbers and numeric strings become floats; a symbolic value that does
    not parse stands in as 1.0 with a warning (parser.py:18-27)."""
    if isinstance(x, (int, float)):
        return float(x)
    try:
        return float(x)
    except (TypeError, ValueError):
        print(f"WARNING: symbolic value '{x}' – substituting 1.0 for numeric path.")
        return 1.0


_ROUND_TYPES = ("sphere", "cylinder", "torus", "toroid")


@dataclass
class CanonicalSpec:
    domain: Any
    conductors: List[Dict[str, Any]]
    dielectrics: List[Dict[str, Any]]
    charges: List[Dict[str, Any]]
    BCs: str
    symmetry: List[str] = field(default_factory=list)
    queries: List[str] = field(default_factory=list)
    symbols: Dict[str, Any] = field(default_factory=dict)

    @staticmethod
    def from_json(d: Dict[str, Any]) -> "CanonicalSpec":
        """Raw dict -> spec, normalising numbers in place (parser.py:49-95):
        charge q and pos, conductor potential, plane z, radius / centre of
        round conductors, major / minor radius of tori."""
        if "domain" not in d:
            raise ValueError("Spec missing required key 'domain'.")
        d.setdefault("BCs", "Dirichlet")
        for ch in d.get("charges", []):
            if "q" in ch:
                ch["q"] = _as_float(ch["q"])
            if "pos" in ch:
                ch["pos"] = [_as_float(v) for v in ch["pos"]]
        for c in d.get("conductors", []):
            if "potential" in c:
                c["potential"] = _as_float(c["potential"])
            kind = c.get("type")
            if kind == "plane" and "z" in c:
                c["z"] = _as_float(c["z"])
            if kind in _ROUND_TYPES:
                if "radius" in c:
                    c["radius"] = _as_float(c["radius"])
                if "center" in c:
                    c["center"] = [_as_float(v) for v in c["center"]]
            if kind in ("torus", "toroid"):
                for key in ("major_radius", "minor_radius"):
                    if key in c:
                        c[key] = _as_float(c[key])
        return CanonicalSpec(domain=d["domain"], conductors=d.get("conductors", []),
                             dielectrics=d.get("dielectrics", []), charges=d.get("charges", []),
                             BCs=d["BCs"], symmetry=d.get("symmetry", []),
                             queries=d.get("queries", []), symbols=d.get("symbols", {}))

    def summary(self) -> Dict[str, Any]:
        return {"conductors_count": len(self.conductors), "charges_count": len(self.charges),
                "queries": self.queries}

    def get_charge_locations(self) -> List[Tuple[float, float, float]]:
        return [tuple(c["pos"]) for c in self.charges
                if c.get("type") == "point" and c.get("pos") is not None]

    def to_json(self) -> Dict[str, Any]:
        return {"domain": copy.deepcopy(self.domain), "conductors": copy.deepcopy(self.conductors),
                "dielectrics": copy.deepcopy(self.dielectrics),
                "charges": copy.deepcopy(self.charges), "BCs": self.BCs,
                "symmetry": copy.deepcopy(self.symmetry), "queries": copy.deepcopy(self.queries),
                "symbols": copy.deepcopy(self.symbols)}

    def point_charges(self) -> List[Tuple[float, Tuple[float, float, float]]]:
        """(q, (x, y, z)) of every point charge, the form the operator takes."""
        out = []
        for c in self.charges:
            if c.get("type") == "point" and c.get("pos") is not None:
                out.append((float(c.get("q", 0.0)), tuple(float(v) for v in c["pos"])))
        return out


def _load_raw(spec: SpecLike) -> Dict[str, Any]:
    """.json / .yaml / .yml path, other paths JSON then YAML, or a mapping
    (shallow-copied) (parser.py:136-167)."""
    if isinstance(spec, (str, Path)):
        path = Path(spec)
        if not path.exists():
            raise FileNotFoundError(f"Spec path does not exist: {path}")
        text = path.read_text(encoding="utf-8")
        suffix = path.suffix.lower()
        if suffix == ".json":
            return json.loads(text)
        if suffix in (".yaml", ".yml"):
            return yaml.safe_load(text)
        try:
            return json.loads(text)
        except ValueError:
            return yaml.safe_load(text)
    return dict(spec)


def parse_spec(spec: SpecLike) -> CanonicalSpec:
    """parser.py:170-182."""
    return CanonicalSpec.from_json(_load_raw(spec))
