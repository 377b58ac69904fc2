"""Spec ingest and meshing (CPU): CanonicalSpec / parse_spec
(orchestration/parser.py:31-182), generate_mesh (core/bem_mesh.py:27-177).

Oracles: oracle.mesh.uv_sphere / spec_mesh_sizes, pinned here by closed forms
(vertices on the sphere, degenerate pole count, area -> 4 pi r^2 at O(h^2)).
The two specs are the reference's specs/sphere_point.yaml and
specs/plane_point.yaml, written out as dicts.
"""
import json
import math

import numpy as np
import pytest

from oracle import mesh as omesh

SPHERE_POINT = {"charges": [{"type": "point", "q": 1.0, "pos": [0.0, 0.0, 2.0]}],
                "conductors": [{"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 1.0}],
                "domain": {"bbox": [[-2, -2, -2], [2, 2, 4]]}}
PLANE_POINT = {"charges": [{"type": "point", "q": 1.0, "pos": [0.0, 0.0, 0.5]}],
               "conductors": [{"type": "plane", "z": 0.0}],
               "domain": {"bbox": [[-1, -1, 0], [1, 1, 2]]}}


def test_parse_spec_normalises_like_the_reference(tmp_path, capsys):
    from electrodrive_amd.orchestration import CanonicalSpec, parse_spec
    raw = {"domain": "R3",
           "conductors": [{"type": "sphere", "center": ["0", 0, 1], "radius": "2.5",
                           "potential": "3"},
                          {"type": "torus", "major_radius": "1", "minor_radius": "0.25"},
                          {"type": "plane", "z": "-1"}],
           "charges": [{"type": "point", "q": "a/b", "pos": [0, "1", 2]}]}
    s = parse_spec(raw)
    assert s.BCs == "Dirichlet"
    assert s.conductors[0]["center"] == [0.0, 0.0, 1.0] and s.conductors[0]["radius"] == 2.5
    assert s.conductors[0]["potential"] == 3.0 and s.conductors[2]["z"] == -1.0
    assert s.conductors[1]["minor_radius"] == 0.25
    assert s.charges[0]["q"] == 1.0  # symbolic value -> 1.0 with a warning
    assert "symbolic value" in capsys.readouterr().out
    assert s.get_charge_locations() == [(0.0, 1.0, 2.0)]
    assert s.summary() == {"conductors_count": 3, "charges_count": 1, "queries": []}
    j = s.to_json()
    j["conductors"][0]["radius"] = 99.0
    assert s.conductors[0]["radius"] == 2.5  # deep copy
    with pytest.raises(ValueError):
        CanonicalSpec.from_json({"conductors": []})
    p = tmp_path / "s.yaml"
    p.write_text("﻿charges:\n  - type: point\n    q: 1.0\n    pos: [0.0, 0.0, 2.0]\n"
                 "conductors:\n  - type: sphere\n    center: [0.0, 0.0, 0.0]\n    radius: 1.0\n"
                 "domain:\n  bbox: [[-2,-2,-2],[2,2,4]]\n", encoding="utf-8")
    assert parse_spec(p).to_json() == parse_spec(dict(SPHERE_POINT)).to_json()
    q = tmp_path / "s.json"
    q.write_text(json.dumps(PLANE_POINT))
    assert parse_spec(q).conductors == [{"type": "plane", "z": 0.0}]
    with pytest.raises(FileNotFoundError):
        parse_spec(tmp_path / "missing.json")


def test_uv_sphere_oracle_pins():
    for nt, nphi in ((10, 20), (24, 48)):
        v, t = omesh.uv_sphere((0.5, -1.0, 2.0), 1.5, nt, nphi)
        assert np.allclose(np.linalg.norm(v - [0.5, -1.0, 2.0], axis=1), 1.5, atol=1e-14)
        tq, c, nrm, a = omesh.panel_quantities(v, t)
        assert t.shape[0] == 2 * (nt - 1) * nphi
        assert tq.shape[0] == t.shape[0] - 2 * nphi  # one degenerate triangle per pole cell
        assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0)
    # area -> 4 pi r^2 from below, second order in h
    errs = []
    for k in (1, 2, 4):
        v, t = omesh.uv_sphere((0, 0, 0), 1.0, 12 * k, 24 * k)
        a = omesh.panel_quantities(v, t)[3].sum()
        errs.append(4 * math.pi - a)
    assert all(e > 0 for e in errs)
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5


def test_generate_mesh_matches_reference_rules():
    from electrodrive_amd.core.bem_mesh import generate_mesh, uv_sphere
    from electrodrive_amd.orchestration import parse_spec
    for raw, h in ((SPHERE_POINT, 0.2), (SPHERE_POINT, 0.35), (PLANE_POINT, 0.2),
                   (PLANE_POINT, 0.5)):
        s = parse_spec(json.loads(json.dumps(raw)))
        m = generate_mesh(s, target_h=h)
        (kind, a, b), = omesh.spec_mesh_sizes(s.conductors, s.charges, h)
        if kind == "sphere":
            v, t = omesh.uv_sphere((0.0, 0.0, 0.0), 1.0, a, b)
            assert m.n_panels == 2 * (a - 1) * b - 2 * b
        else:
            v, t = omesh.plane_patch(0.0, a, b)
            assert m.n_panels == 2 * b * b
            assert abs(m.areas.sum() - a * a) < 1e-9 * a * a
        tq, c, nrm, ar = omesh.panel_quantities(v, t)
        assert np.array_equal(m.triangles, tq)
        assert np.max(np.abs(m.vertices - v)) < 1e-14
        assert np.max(np.abs(m.centroids - c)) < 1e-14
        assert np.max(np.abs(m.areas - ar)) < 1e-14
        assert np.array_equal(m.conductor_ids, np.zeros(m.n_panels, dtype=np.int32))
    # 2112 panels for the sphere spec at h = 0.2 (24 x 48 grid), 3200 for the
    # plane spec (L = 8, 40 cells a side)
    assert generate_mesh(parse_spec(json.loads(json.dumps(SPHERE_POINT))), 0.2).n_panels == 2112
    assert generate_mesh(parse_spec(json.loads(json.dumps(PLANE_POINT))), 0.2).n_panels == 3200
    m = uv_sphere((0, 0, 0), 1.0, 10, 20)
    assert m.n_panels == 2 * 9 * 20 - 40


def test_generate_mesh_multi_conductor_ids_and_errors(monkeypatch):
    from electrodrive_amd.core.bem_mesh import generate_mesh
    from electrodrive_amd.orchestration import parse_spec
    s = parse_spec({"domain": "R3", "charges": [],
                    "conductors": [{"type": "sphere", "radius": 0.5, "center": [2, 0, 0],
                                    "id": 7},
                                   {"type": "torus", "major_radius": 1.0, "minor_radius": 0.3}]})
    m = generate_mesh(s, 0.3)
    sizes = omesh.spec_mesh_sizes(s.conductors, s.charges, 0.3)
    ns = 2 * (sizes[0][1] - 1) * sizes[0][2] - 2 * sizes[0][2]
    nt = 2 * sizes[1][1] * sizes[1][2]
    assert m.n_panels == ns + nt
    assert set(np.unique(m.conductor_ids[:ns])) == {7} and set(np.unique(m.conductor_ids[ns:])) == {1}
    assert abs(m.areas[ns:].sum() - 4 * math.pi ** 2 * 0.3) < 0.05 * 4 * math.pi ** 2 * 0.3
    monkeypatch.setenv("EDE_BEM_SPHERE_DENSITY_SCALE", "3.0")
    m3 = generate_mesh(parse_spec(json.loads(json.dumps(SPHERE_POINT))), 0.2)
    (_, a, b), = omesh.spec_mesh_sizes([SPHERE_POINT["conductors"][0]], [], 0.2, sphere_density=3.0)
    assert m3.n_panels == 2 * (a - 1) * b - 2 * b
    with pytest.raises(NotImplementedError):
        generate_mesh(parse_spec({"domain": "R3", "conductors": [{"type": "cylinder"}]}), 0.2)
    empty = generate_mesh(parse_spec({"domain": "R3"}), 0.2)
    assert empty.n_panels == 0


def test_bem_config_and_logger(tmp_path):
    from electrodrive_amd.utils import BEMConfig, JsonlLogger
    cfg = BEMConfig()
    assert (cfg.initial_h, cfg.max_refine_passes, cfg.gmres_tol, cfg.gmres_restart,
            cfg.target_bc_inf_norm) == (0.2, 3, 5e-8, 256, 1e-7)
    lg = JsonlLogger(tmp_path / "logs")
    lg.info("hello", x=float("nan"), n=3)
    lg.error("bad", exc_info=True)
    lg.close()
    lines = [json.loads(x) for x in (tmp_path / "logs" / "events.jsonl").read_text().splitlines()]
    assert lines[0]["msg"] == "hello" and lines[0]["x"] == "nan" and lines[0]["n"] == 3
    assert lines[1]["level"] == "ERROR"
    JsonlLogger(None).info("nowhere")
