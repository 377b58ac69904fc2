"""Panel quadrature and near-field correction restated from the reference.

Test infrastructure only (see oracle/__init__.py).  numpy, FP64.

- ``standard_triangle_quadrature``   core/bem_quadrature.py:79-150
- ``reference_triangle_quadrature``  core/bem_quadrature.py:158-209
- ``near_singular_quadrature``       core/bem_quadrature.py:217-308
- ``self_integral_correction``       core/bem_quadrature.py:316-344
- ``self_integrals``                 core/bem.py:1764-1816 (refined rule when
                                     near quadrature is enabled, else the disk)
- ``near_pairs``                     core/bem.py:475-555
- ``near_weights``                   core/bem.py:558-667
- ``near_apply``                     core/bem.py:675-734

Parity: the order-2 rule's nodes by sum(w) = area, degree-2 exactness and its
closed-form degree-3 error +A/540 (which the edge-midpoint rule, also exact to
degree 2, misses by -A/60); the refined self rule by the closed form of its
self-similar recursion at the centroid; ``near_weights`` by scipy's adaptive
2-D integral of panel j seen from C_i on panels 22x apart in area (a swap of
i and j fails by ~99 %); the near-pair set by a brute-force double loop; the
apply by an explicit scatter (tests/test_cpu_oracle.py).
"""
import math

import numpy as np

from .constants import K_E

_A2 = 2.0 / 3.0
_B6 = 1.0 / 6.0


def _tri_area(v):
    return 0.5 * float(np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0])))


def standard_triangle_quadrature(vertices, order=2):
    """core/bem_quadrature.py:79-150."""
    v = np.asarray(vertices, dtype=float).reshape(3, 3)
    if order <= 1:
        bary = np.array([[1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0]])
        w_ref = np.array([1.0])
    elif order == 2:
        bary = np.array([[_A2, _B6, _B6], [_B6, _A2, _B6], [_B6, _B6, _A2]])
        w_ref = np.full(3, 1.0 / 3.0)
    else:
        raise NotImplementedError(order)
    area = _tri_area(v)
    if area <= 0.0:
        return np.mean(v, axis=0, keepdims=True), np.zeros_like(w_ref)
    pts = bary[:, 0:1] * v[0:1] + bary[:, 1:2] * v[1:2] + bary[:, 2:3] * v[2:3]
    return pts, w_ref * area


def reference_triangle_quadrature(order=2):
    """core/bem_quadrature.py:158-209."""
    if order <= 1:
        return np.array([[1.0 / 3.0, 1.0 / 3.0]]), np.array([0.5])
    if order == 2:
        return (np.array([[_A2, _B6], [_B6, _A2], [_B6, _B6]]),
                np.full(3, 1.0 / 6.0))
    raise NotImplementedError(order)


def _split4(tri):
    v0, v1, v2 = tri
    m01 = 0.5 * (v0 + v1)
    m12 = 0.5 * (v1 + v2)
    m20 = 0.5 * (v2 + v0)
    return [np.stack([v0, m01, m20]), np.stack([m01, v1, m12]),
            np.stack([m20, m12, v2]), np.stack([m01, m12, m20])]


def near_singular_quadrature(target, panel_vertices, method="telles", order=2,
                             max_depth=None):
    """core/bem_quadrature.py:217-308 (refine the sub-triangle whose centroid
    is closest to the target, first index on ties, 4-way split appended)."""
    verts = np.asarray(panel_vertices, dtype=float).reshape(3, 3)
    tgt = np.asarray(target, dtype=float).reshape(3)
    depth = max_depth
    if depth is None:
        depth = 4 if method.lower() == "telles" else 1
    if depth <= 0:
        return standard_triangle_quadrature(verts, order)
    tris = [verts]
    for _ in range(depth):
        cents = np.array([t.mean(axis=0) for t in tris])
        d = np.linalg.norm(cents - tgt[None, :], axis=1)
        idx = int(np.argmin(d))
        tri = tris.pop(idx)
        tris.extend(_split4(tri))
    pts, ws = [], []
    for t in tris:
        p, w = standard_triangle_quadrature(t, order)
        if p.size == 0:
            continue
        pts.append(p)
        ws.append(w)
    if not pts:
        return standard_triangle_quadrature(verts, order)
    return np.vstack(pts), np.concatenate(ws)


def self_integral_correction(area):
    """core/bem_quadrature.py:316-344: 2*pi*K_E*sqrt(A/pi)."""
    a = float(area)
    if a <= 0.0:
        return 0.0
    return 2.0 * math.pi * K_E * math.sqrt(a / math.pi)


def self_integrals(centroids, panel_vertices, areas, depth=3, order=2,
                   refined=True):
    """core/bem.py:1777-1816.  Returns (I_self, self_corr = I_self / A)."""
    C = np.asarray(centroids, dtype=float)
    A = np.asarray(areas, dtype=float)
    N = C.shape[0]
    vals = np.empty(N)
    for i in range(N):
        if refined:
            pts, w = near_singular_quadrature(C[i], panel_vertices[i], "telles",
                                              order, max_depth=depth)
            r = np.maximum(np.linalg.norm(pts - C[i][None, :], axis=1), 1e-12)
            vals[i] = float(np.sum((K_E / r) * w))
        else:
            vals[i] = self_integral_correction(A[i])
    A_safe = np.maximum(A, np.finfo(np.float64).tiny)
    return vals, vals / A_safe


def near_pairs(centroids, areas, distance_factor=1.5):
    """core/bem.py:475-555: (i, j), i != j, 0 < |C_i - C_j| <= f (R_i + R_j),
    R = sqrt(A/pi); i-major, j ascending."""
    C = np.asarray(centroids, dtype=float)
    A = np.asarray(areas, dtype=float).reshape(-1)
    N = min(C.shape[0], A.shape[0])
    if N <= 1:
        return np.zeros((0, 2), dtype=np.int64)
    C, A = C[:N], A[:N]
    radii = np.sqrt(np.clip(A, 0.0, np.inf) / math.pi)
    df = float(distance_factor)
    pi_list, pj_list = [], []
    for i in range(N):
        diff = C[i] - C
        d = np.linalg.norm(diff, axis=1)
        th = df * (radii[i] + radii)
        m = np.isfinite(d) & np.isfinite(th) & (d > 0.0) & (d <= th)
        js = np.nonzero(m)[0]
        if js.size:
            pi_list.append(np.full(js.shape, i, dtype=np.int64))
            pj_list.append(js.astype(np.int64))
    if not pi_list:
        return np.zeros((0, 2), dtype=np.int64)
    return np.stack([np.concatenate(pi_list), np.concatenate(pj_list)], axis=1)


def uniform_panel_rule(verts, depth, order=2):
    """core/bem.py:590-631: uniform 4^depth refinement + base rule."""
    v = np.asarray(verts, dtype=float).reshape(3, 3)
    if depth <= 0:
        return standard_triangle_quadrature(v, order)
    tris = [v]
    for _ in range(depth):
        nxt = []
        for t in tris:
            nxt.extend(_split4(t))
        tris = nxt
    pts, ws = [], []
    for t in tris:
        p, w = standard_triangle_quadrature(t, order)
        pts.append(p)
        ws.append(w)
    return np.vstack(pts), np.concatenate(ws)


def near_weights(centroids, areas, panel_vertices, pairs, order=2, depth=None):
    """core/bem.py:558-667: delta_ij = sum_k K_E w_k / max(|C_i - y_k|, 1e-12)
    - K_E A_j / max(|C_i - C_j|, 1e-12); ``depth`` None means 0."""
    pairs = np.asarray(pairs, dtype=np.int64)
    if pairs.size == 0:
        return np.zeros((0,))
    C = np.asarray(centroids, dtype=float)
    A = np.asarray(areas, dtype=float).reshape(-1)
    Vt = np.asarray(panel_vertices, dtype=float)
    d = 0 if depth is None else max(0, int(depth))
    out = np.zeros(pairs.shape[0])
    order_j = np.argsort(pairs[:, 1], kind="stable")
    js = pairs[order_j, 1]
    bounds = np.flatnonzero(np.diff(js)) + 1
    for grp in np.split(order_j, bounds):
        if grp.size == 0:
            continue
        j = int(pairs[grp[0], 1])
        if A[j] <= 0.0:
            continue
        pts, w = uniform_panel_rule(Vt[j], d, order)
        tg = C[pairs[grp, 0]]
        r = np.maximum(np.linalg.norm(tg[:, None, :] - pts[None, :, :], axis=2), 1e-12)
        I_near = np.sum((K_E / r) * w[None, :], axis=1)
        rf = np.maximum(np.linalg.norm(tg - C[j][None, :], axis=1), 1e-12)
        out[grp] = I_near - K_E * A[j] / rf
    return out


def near_apply(V_far, sigma, pairs, weights):
    """core/bem.py:675-734: V_far + index_add(i, sigma_j * delta_ij)."""
    V = np.array(V_far, dtype=float, copy=True)
    pairs = np.asarray(pairs, dtype=np.int64)
    if pairs.size == 0:
        return V
    s = np.asarray(sigma, dtype=float)
    np.add.at(V, pairs[:, 0], s[pairs[:, 1]] * np.asarray(weights, dtype=float))
    return V


def near_target_potentials(V_far, targets, centroids, areas, sigma, panel_vertices,
                           distance_factor, quad_order=2, max_depth=None):
    """core/bem.py:940-1047: near-singular correction of target potentials.
    For every target and every panel with |c_j - x| <= f R_j (R = sqrt(A/pi)),
    A_j > 0 and sigma_j != 0, replace the lumped term K_E A_j / max(r, 1e-12)
    by the refined panel integral toward the target."""
    V = np.array(V_far, dtype=float, copy=True)
    T = np.asarray(targets, dtype=float).reshape(-1, 3)
    C = np.asarray(centroids, dtype=float)
    A = np.asarray(areas, dtype=float).reshape(-1)
    S = np.asarray(sigma, dtype=float).reshape(-1)
    PV = np.asarray(panel_vertices, dtype=float).reshape(-1, 3, 3)
    radii = np.sqrt(np.clip(A, 0.0, np.inf) / math.pi)
    for i in range(T.shape[0]):
        x = T[i]
        d = np.linalg.norm(C - x[None, :], axis=1)
        for j in np.nonzero(d <= distance_factor * radii)[0]:
            if A[j] <= 0.0 or S[j] == 0.0:
                continue
            pts, w = near_singular_quadrature(x, PV[j], "telles", quad_order, max_depth)
            r = np.maximum(np.linalg.norm(pts - x[None, :], axis=1), 1e-12)
            i_near = float(np.sum((K_E / r) * w))
            i_far = K_E * A[j] / max(d[j], 1e-12)
            V[i] += S[j] * (i_near - i_far)
    return V
