"""Restatement of the reference's CPU FMM matvec, the second leg of the CPU
baseline (the reference FMM algorithm on the host's cores).

Test / benchmark infrastructure only (see oracle/__init__.py): nothing in the
package imports it.  It follows the reference's CPU path step by step
(fmm3d/bem_fmm.py:460-560, fmm3d/kernels_cpu.py:401-640,
fmm3d/multipole_operators.py:369-981 and :1102-1400, :1470-1770):

- tree and lists: the reference's adaptive tight-box octree and dual-tree MAC
  lists (oracle.ref_tree, bit-exact with fmm3d/tree.py:801-960 and
  interaction_lists.py:273-577);
- one global expansion scale S = the largest node radius
  (multipole_operators.py:1470-1483);
- complex solid harmonics in full (p+1)^2 packing, P2M M = sum q conj(R(y)),
  L2P V = Re sum L R(y) / S, y = (x - c) / S (multipole_operators.py:
  1102-1165, :1336-1400);
- every translation matrix built by least squares for its own translation
  vector, at every matvec, as the reference does (no operator cache): M2M
  and M2L from 16 x 32 proxy points with row-normalised lstsq, the
  charge-to-local map by matching 1/r on four check shells, L2L by
  collocation on three shells (multipole_operators.py:369-981);
- P2P by direct summation over the P2P leaf pairs, self excluded.

The harmonics are r^n Y_n^m with orthonormal Y (oracle.fmm_math's recurrence,
rescaled); the least-squares operators are defined relative to P2M and L2P,
so the result does not depend on the normalisation as long as the systems
stay well conditioned.  It is pinned against the FP64 dense matvec in
tests/test_cpu_ref_fmm.py (the reference's own docs report 1e-3 to 1e-2
error at p = 8).
"""
import math

import numpy as np

from . import fmm_math as fm
from . import ref_tree as rt

K_E = 1.0 / (4.0 * math.pi * 8.854187817e-12)


def proxy_surface(radius, n_theta=16, n_phi=32):
    """multipole_operators.py:369-394: theta on [0, pi] inclusive, phi on [0, 2 pi)."""
    th = np.linspace(0.0, math.pi, n_theta)
    ph = np.linspace(0.0, 2.0 * math.pi, n_phi + 1)[:-1]
    T, F = np.meshgrid(th, ph, indexing="ij")
    return np.stack([radius * np.sin(T) * np.cos(F), radius * np.sin(T) * np.sin(F),
                     radius * np.cos(T)], axis=-1).reshape(-1, 3)


def _full(Xh, p):
    """(..., p+1, p+1) half table -> (..., (p+1)^2) complex, m = -n..n per n."""
    out = np.zeros(Xh.shape[:-2] + ((p + 1) ** 2,), dtype=complex)
    for n in range(p + 1):
        for m in range(-n, n + 1):
            out[..., n * n + n + m] = fm.full(Xh, n, m)
    return out


def _ynorm(p):
    """(n+m)! sqrt((2n+1)/(4 pi) (n-m)!/(n+m)!) for 0 <= m <= n <= p: turns
    oracle.fmm_math's R_n^m = r^n P_n^m e^{i m phi} / (n+m)! into r^n Y_n^m
    with orthonormal Y, the scale-balanced form the reference's normalised
    harmonics have (multipole_operators.py:172-266); the least-squares
    translation builders need it to stay well conditioned."""
    f = np.zeros((p + 1, p + 1))
    for n in range(p + 1):
        for m in range(n + 1):
            f[n, m] = math.factorial(n + m) * math.sqrt(
                (2 * n + 1) / (4 * math.pi) * math.factorial(n - m) / math.factorial(n + m))
    return f


def regular_full(y, p):
    return _full(fm.regular(y, p) * _ynorm(p), p)


def _lstsq_rows(A, B):
    """T^T of A T^T = B, rows of both divided by |A row| (the reference's
    conditioning step), solved by least squares."""
    rn = np.maximum(np.linalg.norm(A, axis=1), 1e-15)[:, None]
    return np.linalg.lstsq(A / rn, B / rn, rcond=None)[0]


def charge_to_multipole(p, pts, center):
    """P_M (P2, N_p): M = P_M q (multipole_operators.py:397-452)."""
    return np.conj(regular_full(pts - center, p)).T


def charge_to_local(p, proxy, center, radii):
    """P_L (P2, N_p): local coefficients matching sum q_j / |y - y_j| on check
    shells around `center` in the least-squares sense (:454-577)."""
    Rb, Gb = [], []
    for r in radii:
        chk = proxy_surface(r)
        d = np.linalg.norm((center + chk)[:, None, :] - proxy[None, :, :], axis=-1)
        Gb.append(1.0 / np.maximum(d, 1e-15))
        Rb.append(regular_full(chk, p))
    return np.linalg.lstsq(np.concatenate(Rb), np.concatenate(Gb).astype(complex),
                           rcond=None)[0]


def m2m_matrix(p, t):
    """M_parent = T M_child, t = (parent - child) / S (:665-757)."""
    proxy = proxy_surface(0.5)
    A = charge_to_multipole(p, proxy, np.zeros(3)).T
    B = charge_to_multipole(p, proxy, np.asarray(t)).T
    return _lstsq_rows(A, B).T


def m2l_matrix(p, t):
    """L_target = T M_source, t = (target - source) / S (:758-897)."""
    nt = float(np.linalg.norm(t))
    r_proxy = min(0.5, 0.5 * nt)
    proxy = proxy_surface(r_proxy)
    sp = r_proxy / 0.5
    base = [0.2 * sp, 0.4 * sp, 0.6 * sp, 0.8 * sp]
    r_conv = max(nt - r_proxy, 1e-3)
    sr = min(1.0, r_conv / max(base))
    radii = [max(r * sr, 1e-3) for r in base]
    A = charge_to_multipole(p, proxy, np.zeros(3)).T
    B = charge_to_local(p, proxy, np.asarray(t), radii).T
    return _lstsq_rows(A, B).T


def l2l_matrix(p, t):
    """L_child = T L_parent, t = (child - parent) / S (:898-981)."""
    if np.linalg.norm(t) < 1e-12:
        return np.eye((p + 1) ** 2, dtype=complex)
    chk = np.concatenate([proxy_surface(r) for r in (0.25, 0.5, 0.8)])
    Bc = regular_full(chk, p)
    Bp = regular_full(chk + np.asarray(t), p)
    return _lstsq_rows(Bc, Bp)


def fmm_matvec(centroids, areas, sigma, self_integrals=None, *, p=8, leaf_size=64,
               theta=0.5, stats=None, m2l_sample=None):
    """V_i = K_E sum_{j != i} A_j sigma_j / |c_i - c_j| (+ self_i sigma_i) by
    the reference's FMM: tree, lists, P2M, M2M, M2L, L2L, L2P and P2P, every
    translation matrix built for its own vector by least squares.

    ``m2l_sample = k`` (timing only) builds and applies the first k M2L pairs
    and records in ``stats["m2l_s_per_pair"]`` their mean time; the returned
    vector then lacks the other pairs' contributions.  Per-stage wall times go
    to ``stats["stage_s"]``."""
    import time
    tick = time.perf_counter
    tm = {}
    t0 = tick()
    C = np.ascontiguousarray(centroids, dtype=float)
    A = np.asarray(areas, dtype=float)
    q = np.asarray(sigma, dtype=float) * A
    n = C.shape[0]
    T = rt.build_tree(C, leaf_size, 32)
    p2p, m2l, _ = rt.build_lists(T, None, theta)
    tm["tree_lists"] = tick() - t0
    t0 = tick()
    idx = T["index_map"]
    pts, qt = C[idx], q[idx]
    ctr, rng, leaf = T["centers"], T["ranges"], T["is_leaf"]
    children, parent = T["children"], T["parents"]
    S = float(T["radii"].max()) if T["radii"].size and T["radii"].max() > 0 else 1.0
    nn, P2 = ctr.shape[0], (p + 1) ** 2
    M = np.zeros((nn, P2), dtype=complex)
    L = np.zeros((nn, P2), dtype=complex)
    counts = dict(m2m=0, m2l=0, l2l=0)
    for i in range(nn):  # P2M at leaves
        if leaf[i] and rng[i, 1] > rng[i, 0]:
            b, e = rng[i]
            M[i] = np.conj(regular_full((pts[b:e] - ctr[i]) / S, p)).T @ qt[b:e]
    tm["p2m"] = tick() - t0
    t0 = tick()
    order = np.argsort(-T["levels"], kind="stable")  # M2M, deepest parents first
    for i in order:
        if not leaf[i]:
            for c in children[i]:
                if c >= 0:
                    M[i] += m2m_matrix(p, (ctr[i] - ctr[c]) / S) @ M[c]
                    counts["m2m"] += 1
    tm["m2m"] = tick() - t0
    t0 = tick()
    pairs = m2l if m2l_sample is None else m2l[:int(m2l_sample)]
    for s, t in pairs:  # M2L
        L[t] += m2l_matrix(p, (ctr[t] - ctr[s]) / S) @ M[s]
        counts["m2l"] += 1
    tm["m2l"] = tick() - t0
    t0 = tick()
    for i in np.argsort(T["levels"], kind="stable"):  # L2L, shallowest first
        par = parent[i]
        if par >= 0:
            L[i] += l2l_matrix(p, (ctr[i] - ctr[par]) / S) @ L[par]
            counts["l2l"] += 1
    tm["l2l"] = tick() - t0
    t0 = tick()
    V = np.zeros(n)
    for i in range(nn):  # L2P
        if leaf[i] and rng[i, 1] > rng[i, 0]:
            b, e = rng[i]
            V[b:e] += (regular_full((pts[b:e] - ctr[i]) / S, p) @ L[i]).real / S
    for s, t in p2p:  # P2P, self excluded
        bs, es = rng[s]
        bt, et = rng[t]
        d = np.linalg.norm(pts[bt:et, None, :] - pts[None, bs:es, :], axis=-1)
        with np.errstate(divide="ignore"):
            k = np.where(d > 0.0, 1.0 / d, 0.0)
        V[bt:et] += k @ qt[bs:es]
    tm["l2p_p2p"] = tick() - t0
    out = np.empty(n)
    out[idx] = K_E * V
    if self_integrals is not None:
        out += np.asarray(self_integrals, dtype=float) * np.asarray(sigma, dtype=float)
    if stats is not None:
        stats.update(counts, n_nodes=nn, n_p2p_pairs=len(p2p), n_m2l_pairs=len(m2l), scale=S,
                     stage_s=tm, m2l_s_per_pair=tm["m2l"] / max(1, counts["m2l"]))
    return out
