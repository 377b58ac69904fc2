"""CPU: the restated reference FMM (oracle/ref_fmm.py), the second leg of
the CPU baseline.  Pins: the proxy surface (fmm3d/multipole_operators.py:
369-394), the least-squares translation builders against what the
mathematics fixes (M2M of a multipole and L2L of a local expansion are exact
re-expansions, M2L reproduces 1/r between separated clusters to the
truncation level), and the whole matvec against the FP64 dense oracle."""
import numpy as np

from oracle import dense as od
from oracle import ref_fmm as rf


def test_proxy_surface():
    P = rf.proxy_surface(0.7)
    assert P.shape == (16 * 32, 3)
    assert np.allclose(np.linalg.norm(P, axis=1), 0.7)
    assert np.allclose(P[:32], [0.0, 0.0, 0.7])      # theta = 0 row: the pole, 32 times
    assert np.allclose(P[-32:, 2], -0.7)


def test_translation_builders():
    rng = np.random.default_rng(2)
    p = 6
    src = rng.uniform(-0.05, 0.05, (40, 3))
    q = rng.normal(size=40)
    # M2M: the multipole about the parent equals P2M about the parent
    c = np.array([0.05, -0.04, 0.03])
    Mc = np.conj(rf.regular_full(src + c - c, p)).T @ q
    Mp = rf.m2m_matrix(p, -c) @ Mc
    Mp_direct = np.conj(rf.regular_full(src + c, p)).T @ q
    assert np.abs(Mp - Mp_direct).max() < 1e-10 * np.abs(Mp_direct).max()
    # L2L: a local expansion re-expanded about another centre is the same
    # polynomial
    L = rng.normal(size=(p + 1) ** 2) + 1j * rng.normal(size=(p + 1) ** 2)
    t = np.array([0.06, 0.02, -0.05])
    Lc = rf.l2l_matrix(p, t) @ L
    x = t + rng.uniform(-0.1, 0.1, (30, 3))
    assert np.abs(rf.regular_full(x - t, p) @ Lc - rf.regular_full(x, p) @ L).max() \
        < 1e-9 * np.abs(rf.regular_full(x, p) @ L).max()
    # M2L between separated clusters against the direct sum (p = 6: ~1e-3)
    tv = np.array([0.5, 0.12, -0.07])
    tgt = tv + rng.uniform(-0.05, 0.05, (30, 3))
    V = (rf.regular_full(tgt - tv, p) @ (rf.m2l_matrix(p, tv) @ (np.conj(rf.regular_full(src, p)).T @ q))).real
    D = np.array([np.sum(q / np.linalg.norm(x - src, axis=1)) for x in tgt])
    err = np.abs(V - D).max() / np.abs(D).max()
    print("m2l err", err)
    assert err < 1e-2


def test_fmm_matvec_vs_dense():
    """Two clusters 3 apart (60 points each, leaf 16): the cross-cluster
    interactions go through M2L, the rest through P2P and the local
    translations; the matvec matches the FP64 dense oracle to 1e-3, and a
    timing run with an M2L sample reports its per-pair cost."""
    rng = np.random.default_rng(4)
    X = np.concatenate([rng.uniform(-0.3, 0.3, (60, 3)),
                        rng.uniform(-0.3, 0.3, (60, 3)) + [3.0, 0.5, 0.2]])
    A = rng.uniform(0.5, 1.5, 120) * 1e-2
    s = rng.normal(size=120)
    st = {}
    v = rf.fmm_matvec(X, A, s, p=6, leaf_size=16, theta=0.5, stats=st)
    ref = od.dense_matvec(X, A, s, None)
    err = np.linalg.norm(v - ref) / np.linalg.norm(ref)
    print("ref fmm err", err, st["n_m2l_pairs"], st["n_p2p_pairs"])
    assert st["n_m2l_pairs"] > 0 and err < 1e-3
    st2 = {}
    rf.fmm_matvec(X, A, s, p=4, leaf_size=16, stats=st2, m2l_sample=2)
    assert st2["m2l"] == 2 and st2["m2l_s_per_pair"] > 0
