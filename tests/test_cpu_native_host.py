"""CPU-only checks of the native library's host side (no device needed).

- the library loads and exports every symbol declared in include/*.h;
- the host lattice octree and U/V lists equal the oracle restatement
  (oracle/lattice_tree.py) and satisfy the partition property;
- operator tables equal oracle/fmm_math.py (pinned there against direct sums);
- host Morton keys equal oracle/morton.py bit for bit.
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from oracle import fmm_math as fm
from oracle import lattice_tree as lt
from oracle import morton as om

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from electrodrive_amd import _native as nat
    return nat.lib()


def test_exports_every_declared_symbol():
    lib = _lib()
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        names |= set(re.findall(r"^(?:int|void|const char\*|int32_t)\s+(ed_\w+)\s*\(", txt, re.M))
    assert len(names) >= 20
    for nme in sorted(names):
        assert hasattr(lib, nme), nme
    from electrodrive_amd import _native as nat
    assert set(nat.exported_symbols()) >= names


def _host_tree(C, leaf, max_level, ws):
    lib = _lib()
    C = np.ascontiguousarray(C, dtype=np.float64)
    n = C.shape[0]
    cnt = np.zeros(4, dtype=np.int64)
    vp = ctypes.c_void_p
    rc = lib.ed_host_tree_build(C.ctypes.data_as(vp), n, leaf, max_level, ws,
                                cnt.ctypes.data_as(vp), None, None, None, None, None, None, None)
    assert rc == 0
    M, nl, nv, nu = (int(v) for v in cnt)
    perm = np.zeros(n, dtype=np.int64)
    lev = np.zeros(M, dtype=np.int32)
    beg = np.zeros(M, dtype=np.int64)
    end = np.zeros(M, dtype=np.int64)
    par = np.zeros(M, dtype=np.int64)
    V = np.zeros((nv, 2), dtype=np.int64)
    U = np.zeros((nu, 2), dtype=np.int64)
    rc = lib.ed_host_tree_build(C.ctypes.data_as(vp), n, leaf, max_level, ws,
                                cnt.ctypes.data_as(vp), perm.ctypes.data_as(vp),
                                lev.ctypes.data_as(vp), beg.ctypes.data_as(vp),
                                end.ctypes.data_as(vp), par.ctypes.data_as(vp),
                                V.ctypes.data_as(vp), U.ctypes.data_as(vp))
    assert rc == 0
    return dict(perm=perm, level=lev, begin=beg, end=end, parent=par, V=V, U=U, n_leaves=nl)


@pytest.mark.parametrize("case", ["uniform", "flat", "clusters", "sphere"])
@pytest.mark.parametrize("ws", [1, 2])
def test_host_tree_equals_oracle(case, ws):
    rng = np.random.default_rng(11)
    if case == "uniform":
        P = rng.uniform(-1, 1, (900, 3))
    elif case == "flat":
        P = rng.normal(size=(800, 3)) * np.array([1, 1, 1e-3])
    elif case == "clusters":
        P = np.concatenate([rng.normal(size=(400, 3)) * 0.03, rng.normal(size=(300, 3)) * 0.03 + 2])
    else:
        from electrodrive_amd.core import bem_mesh
        P = bem_mesh.icosphere(3).centroids
    leaf = 16
    h = _host_tree(P, leaf, 12, ws)
    t = lt.build(P, leaf_size=leaf, max_level=12, ws=ws)
    assert np.array_equal(h["perm"], t.perm)
    assert np.array_equal(h["level"], t.level)
    assert np.array_equal(h["begin"], t.begin)
    assert np.array_equal(h["end"], t.end)
    assert np.array_equal(h["parent"], t.parent)
    Vo = np.array([(tt, s) for tt in range(t.n_nodes) for s in t.V[tt]], dtype=np.int64).reshape(-1, 2)
    Uo = np.array([(tt, s) for tt in range(t.n_nodes) for s in t.U[tt]], dtype=np.int64).reshape(-1, 2)
    assert np.array_equal(h["V"], Vo)
    assert np.array_equal(h["U"], Uo)
    cov = lt.coverage_counts(t)
    assert cov.min() == 1 and cov.max() == 1


@pytest.mark.parametrize("p", [4, 8])
def test_operator_tables_equal_oracle(p):
    lib = _lib()
    nc = (p + 1) ** 2
    out = np.zeros((nc, nc))
    vp = ctypes.c_void_p
    for off in [(2, 0, 0), (3, -2, 1), (-3, 3, 3), (0, -2, 1)]:
        assert lib.ed_host_operator(p, 0, *off, 0, out.ctypes.data_as(vp)) == 0
        ref = fm.m2l_matrix(off, p)
        assert np.allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    for kind, fn in ((1, fm.m2m_matrix), (2, fm.l2l_matrix)):
        for o in (0, 3, 5, 7):
            assert lib.ed_host_operator(p, kind, 0, 0, 0, o, out.ctypes.data_as(vp)) == 0
            ref = fn(fm.child_offset(o), p)
            assert np.allclose(out, ref, rtol=1e-12, atol=1e-14)
    assert lib.ed_host_operator(p, 0, 1, 0, 0, 0, out.ctypes.data_as(vp)) != 0


def _reflection_signs(p, s):
    """D(s) on the packed real coefficients: index n^2 (m = 0), n^2 + 2m - 1
    (Re), n^2 + 2m (Im); the expansions are polynomials in (x + iy)^m z^(n-m)
    times r^2 powers, so x -> -x conjugates and multiplies by (-1)^m,
    y -> -y conjugates, z -> -z multiplies by (-1)^(n+m); multipoles store
    conj(R), whose packed (Re, -Im) pattern changes the same way."""
    d = np.ones((p + 1) ** 2)
    for n in range(p + 1):
        for m in range(n + 1):
            for part, k in ((0, n * n + (2 * m - 1 if m else 0)), (1, n * n + 2 * m)):
                if m == 0 and part == 1:
                    continue
                f = 1.0
                if s & 1:
                    f *= (-1) ** m * (1 if part == 0 else -1)
                if s & 2 and part == 1:
                    f *= -1
                if s & 4:
                    f *= (-1) ** (n + m)
                d[k] = f
    return d


def test_m2l_reflection_symmetry():
    """The GEMM M2L stores tables for offsets with non-negative components
    only and uses T(R_s o) = D(s) T(o) D(s) for the reflected ones: checked
    for every lattice offset of ws = 1 and a sample of ws = 2 on the host
    operators (pinned above against the oracle's direct sums)."""
    lib = _lib()
    vp = ctypes.c_void_p
    for p in (6, 10):
        nc = (p + 1) ** 2
        T = np.zeros((nc, nc))
        Tc = np.zeros((nc, nc))
        offs = [(x, y, z) for x in range(-3, 4) for y in range(-3, 4) for z in range(-3, 4)
                if max(abs(x), abs(y), abs(z)) >= 2]
        offs += [(-5, 2, -4), (4, -5, 0), (-2, -3, 5)]
        for o in offs:
            s = (1 if o[0] < 0 else 0) | (2 if o[1] < 0 else 0) | (4 if o[2] < 0 else 0)
            assert lib.ed_host_operator(p, 0, *o, 0, T.ctypes.data_as(vp)) == 0
            assert lib.ed_host_operator(p, 0, *(abs(c) for c in o), 0,
                                        Tc.ctypes.data_as(vp)) == 0
            d = _reflection_signs(p, s)
            assert np.allclose(T, d[:, None] * Tc * d[None, :], rtol=0,
                               atol=1e-13 * np.abs(Tc).max()), (p, o)
        # a wrong pattern (x reflection ignored) fails
        o = (-3, 2, 1)
        assert lib.ed_host_operator(p, 0, *o, 0, T.ctypes.data_as(vp)) == 0
        assert lib.ed_host_operator(p, 0, 3, 2, 1, 0, Tc.ctypes.data_as(vp)) == 0
        assert not np.allclose(T, Tc, rtol=0, atol=1e-6 * np.abs(Tc).max())


def test_host_morton_keys_bit_exact():
    lib = _lib()
    rng = np.random.default_rng(3)
    vp = ctypes.c_void_p
    for P in (rng.uniform(-2, 7, (3000, 3)), rng.normal(size=(500, 3)) * [1, 0, 1]):
        P = np.ascontiguousarray(P)
        for level in (1, 7, 16, 21):
            k = np.zeros(P.shape[0], dtype=np.int64)
            assert lib.ed_host_morton_keys_f64(P.ctypes.data_as(vp), P.shape[0], level, None,
                                               None, k.ctypes.data_as(vp)) == 0
            assert np.array_equal(k, om.morton_keys(P, level))


def test_autotune_api_without_device(tmp_path, monkeypatch):
    """autotune_for_problem keeps the reference's dry-run and argument rules
    (fmm3d/autotune.py:100-160); the JSON cache round-trips under a file lock."""
    import pytest
    from electrodrive_amd.fmm3d import FmmConfig, autotune_for_problem
    from electrodrive_amd.fmm3d import autotune as at
    cfg = FmmConfig(expansion_order=8, leaf_size=64)
    r = autotune_for_problem(cfg, 1000)
    assert r.config is cfg and r.metrics["mode"] == 0.0
    with pytest.raises(ValueError):
        autotune_for_problem(cfg, 0)
    assert at._bucket_problem_size(1000) == 1024 and at._bucket_problem_size(1024) == 1024
    monkeypatch.setenv(at.AUTOTUNE_CACHE_ENV_VAR, str(tmp_path / "c.json"))
    at._append_record_to_cache("k1", {"expansion_order": 10, "leaf_size": 128, "metrics": {"ms": 1.0}})
    at._append_record_to_cache("k2", {"expansion_order": 8, "leaf_size": 64, "metrics": {"ms": 2.0}})
    assert at._load_best_from_cache("k1")["leaf_size"] == 128
    assert at._load_best_from_cache("k2")["expansion_order"] == 8
    assert at._load_best_from_cache("missing") is None
