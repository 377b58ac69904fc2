"""Pins of the CPU oracle (oracle/) against mathematics and the reference's own
known-answer tests.  CPU only; every check is chosen so that a plausible
mistake (dropped term, wrong sign/index, transposed operand) fails it.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import analytic, dense, fmm_math as fm, gmres as og, mesh as om, morton as omo
from oracle import quadrature as oq
from oracle.constants import K_E


# --------------------------------------------------------------------- mesh
def test_panel_quantities_and_generators():
    from electrodrive_amd.core import bem_mesh
    for s in (0, 1, 3):
        m = bem_mesh.icosphere(s)
        assert m.n_panels == 20 * 4 ** s
        t, c, n, a = om.panel_quantities(m.vertices, m.triangles)
        assert np.allclose(c, m.centroids) and np.allclose(a, m.areas) and np.allclose(n, m.normals)
        assert np.allclose(np.linalg.norm(n, axis=1), 1.0)
        assert np.all(np.einsum("ij,ij->i", n, c) > 0)  # outward
    m = bem_mesh.icosphere(5)
    assert abs(m.areas.sum() - 4 * np.pi) / (4 * np.pi) < 2e-3
    v, t = om.torus((0, 0, 0), 1.0, 0.3, 40, 16)
    mt = bem_mesh.torus(1.0, 0.3, 40, 16)
    assert np.array_equal(v, mt.vertices) and np.array_equal(t, mt.triangles)
    mt = bem_mesh.torus(1.0, 0.3, 400, 120)
    assert abs(mt.areas.sum() - 4 * np.pi ** 2 * 0.3) / (4 * np.pi ** 2 * 0.3) < 1e-3
    # the reference torus triangles are wound so that normals point into the
    # tube (core/bem_mesh.py:205-211); the orientation is consistent
    tube = mt.centroids - np.c_[
        mt.centroids[:, :2] / np.linalg.norm(mt.centroids[:, :2], axis=1, keepdims=True),
        np.zeros(mt.n_panels)]
    assert np.all(np.einsum("ij,ij->i", mt.normals, tube) < 0)
    v, t = om.plane_patch(0.0, 2.0, 7)
    mp = bem_mesh.plane_patch(0.0, 2.0, 7)
    assert np.array_equal(t, mp.triangles)
    assert abs(mp.areas.sum() - 4.0) < 1e-12


# ------------------------------------------------------------------- morton
def test_morton_hand_keys():
    # level 2 in the unit cube [0,1]^3 (bbox given): cell (ix,iy,iz) -> bits
    # x -> 3b, y -> 3b+1, z -> 3b+2 (fmm3d/geometry.py:722-731)
    pts = np.array([[0.1, 0.1, 0.1], [0.9, 0.1, 0.1], [0.1, 0.9, 0.1], [0.1, 0.1, 0.9],
                    [0.6, 0.3, 0.8], [0.99, 0.99, 0.99]])
    keys = omo.morton_keys(pts, 2, bbox_center=[0.5, 0.5, 0.5], bbox_half=[0.5, 0.5, 0.5])
    # (3,0,0) -> bits x0=1 (1), x1=1 (8) -> 9 ; (0,3,0) -> 2+16 ; (0,0,3) -> 4+32
    # (2,1,3): x=10b, y=01b, z=11b -> b0: x0=0,y0=1,z0=1 -> 6 ; b1: x1=1,y1=0,z1=1 -> 8+32
    assert keys.tolist() == [0, 9, 18, 36, 46, 63]
    # keys sort points into nested octants: prefix at level 1 = octant
    rng = np.random.default_rng(0)
    p = rng.uniform(0, 1, (200, 3))
    k5 = omo.morton_keys(p, 5, [0.5] * 3, [0.5] * 3)
    k1 = omo.morton_keys(p, 1, [0.5] * 3, [0.5] * 3)
    assert np.array_equal(k5 >> 12, k1)
    oct_ = (p[:, 0] >= 0.5) * 1 + (p[:, 1] >= 0.5) * 2 + (p[:, 2] >= 0.5) * 4
    assert np.array_equal(k1, oct_)
    # float32 path: 1 - 1e-9 rounds to 1.0f and the final clamp catches it
    k = omo.morton_keys(np.array([[1.0, 1.0, 1.0], [0.0, 0.0, 0.0]], dtype=np.float32), 3)
    assert k.tolist() == [511, 0]


# -------------------------------------------------------------------- dense
def test_dense_matvec_brute_force_and_sphere():
    rng = np.random.default_rng(1)
    n = 37
    C = rng.normal(size=(n, 3))
    A = rng.uniform(0.1, 1.0, n)
    s = rng.normal(size=n)
    si = rng.uniform(1.0, 2.0, n)
    ref = np.zeros(n)
    for i in range(n):
        for j in range(n):
            if i == j:
                ref[i] += si[i] * s[i] * A[i]
            else:
                ref[i] += K_E * A[j] * s[j] / math.sqrt(sum((C[i] - C[j]) ** 2))
    for tile in (1, 8, 1024):
        assert np.allclose(dense.dense_matvec(C, A, s, si, tile_size=tile), ref, rtol=1e-13)
    # uniformly charged unit sphere: potential on the surface 4 pi K_E R sigma
    from electrodrive_amd.core import bem_mesh
    m = bem_mesh.icosphere(4)
    I, corr = oq.self_integrals(m.centroids, m.panel_vertices, m.areas, depth=3)
    V = dense.dense_matvec(m.centroids, m.areas, np.ones(m.n_panels), corr)
    assert abs(V.mean() / (4 * np.pi * K_E) - 1.0) < 2e-2
    b = dense.dense_matvec_batched(C, A, np.stack([s, 2 * s]), si)
    assert np.allclose(b[1], 2 * ref, rtol=1e-13)


# --------------------------------------------------------------- quadrature
def _tri_moments(v):
    """exact int over the triangle of 1, x, x^2, xy (x,y first two coords)."""
    a = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
    x = v[:, 0]
    y = v[:, 1]
    Ix = a * x.mean()
    Ixx = a / 6.0 * (x @ x + x.sum() ** 2) / 2.0
    Ixy = a / 12.0 * (x @ y + x.sum() * y.sum())
    return a, Ix, Ixx, Ixy


def test_standard_rule_exactness():
    rng = np.random.default_rng(2)
    for _ in range(5):
        v = rng.normal(size=(3, 3))
        a, Ix, Ixx, Ixy = _tri_moments(v)
        for order in (1, 2):
            p, w = oq.standard_triangle_quadrature(v, order)
            assert abs(w.sum() - a) < 1e-13 * a
            assert abs(w @ p[:, 0] - Ix) < 1e-12
        p, w = oq.standard_triangle_quadrature(v, 2)
        assert abs(w @ (p[:, 0] ** 2) - Ixx) < 1e-12
        assert abs(w @ (p[:, 0] * p[:, 1]) - Ixy) < 1e-12
    pr, wr = oq.reference_triangle_quadrature(2)
    assert abs(wr.sum() - 0.5) < 1e-15 and pr.shape == (3, 2)


def test_order2_rule_nodes_by_degree3_moments():
    """Symmetric 3-point rules (a, b, b) exact to degree 2 are exactly two:
    a = 2/3 (interior, the reference's rule, core/bem_quadrature.py:111-150)
    and a = 0 (edge midpoints).  Degree-3 moments tell them apart: from the
    barycentric integral int l1^i l2^j l3^k = 2A i! j! k! / (i+j+k+2)!,
    int l1^3 = A/10 and int l1 l2 l3 = A/60; the (2/3, 1/6, 1/6) rule gives
    11A/108 and A/54, i.e. an error of +A/540 on both, where the midpoint rule
    is off by -A/60 on both."""
    rng = np.random.default_rng(31)
    for _ in range(6):
        v = rng.normal(size=(3, 3))
        p, w = oq.standard_triangle_quadrature(v, 2)
        a = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
        # barycentric coordinates of the nodes (least squares on the 3x3 system
        # [v0 v1 v2]^T lam = p with sum(lam) = 1)
        G = np.vstack([v.T, np.ones(3)])
        lam = np.linalg.lstsq(G, np.vstack([p.T, np.ones(p.shape[0])]), rcond=None)[0].T
        assert np.allclose(np.sort(lam, axis=1), [[1 / 6, 1 / 6, 2 / 3]] * 3, atol=1e-12)
        for k in range(3):
            e3 = w @ lam[:, k] ** 3 - a / 10.0
            assert abs(e3 - a / 540.0) < 1e-12 * a, (e3, a / 540.0)
        e111 = w @ (lam[:, 0] * lam[:, 1] * lam[:, 2]) - a / 60.0
        assert abs(e111 - a / 540.0) < 1e-12 * a
    # the edge-midpoint rule passes the degree-2 checks but fails these
    mid = np.array([[0.5, 0.5, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5]])
    wm = np.full(3, 1.0 / 3.0)
    assert abs(wm @ mid[:, 0] ** 2 - 2 * 2 / 24) < 1e-15  # degree 2 exact (A = 1)
    assert abs((wm @ mid[:, 0] ** 3 - 0.1) - 1.0 / 540.0) > 1e-2


def test_near_singular_rule_self_similarity():
    """At the centroid the reference refinement always splits the central
    sub-triangle, which is the panel scaled by -1/2 about the centroid; the
    3-point rule of 1/r scales linearly, so I_d = c0 (2 - 2^(1-d)) + 2^-d I_0
    with c0 = I_1 - I_0 / 2 (closed form of the reference algorithm)."""
    rng = np.random.default_rng(3)
    for _ in range(4):
        v = rng.normal(size=(3, 3))
        P = v.mean(0)

        def I(d):
            p, w = oq.near_singular_quadrature(P, v, "telles", 2, max_depth=d)
            assert abs(w.sum() - 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))) < 1e-12
            return np.sum(K_E / np.maximum(np.linalg.norm(p - P, axis=1), 1e-12) * w)

        I0, I1 = I(0), I(1)
        c0 = I1 - 0.5 * I0
        for d in range(2, 7):
            closed = c0 * (2.0 - 2.0 ** (1 - d)) + 2.0 ** (-d) * I0
            assert abs(I(d) - closed) < 1e-12 * abs(closed)
    # off-centre target: refining toward it converges to the exact integral
    # (in-plane point, polar form int_0^2pi rho(theta) d theta)
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.2, 0.9, 0.0]])
    P = np.array([0.4, 0.3, 0.5])  # off-plane: exact by adaptive 2-D quadrature
    e1, e2 = v[1] - v[0], v[2] - v[0]
    area = 0.5 * np.linalg.norm(np.cross(e1, e2))

    def inner(u):
        return quad(lambda t: 1.0 / np.linalg.norm(v[0] + u * e1 + t * e2 - P), 0, 1 - u,
                    epsabs=1e-13)[0]
    exact = K_E * 2 * area * quad(inner, 0, 1, epsabs=1e-12)[0]
    p, w = oq.uniform_panel_rule(v, 5, 2)
    approx = np.sum(K_E / np.linalg.norm(p - P, axis=1) * w)
    assert abs(approx - exact) / exact < 1e-4


def test_self_integral_disk_formula():
    assert abs(oq.self_integral_correction(math.pi) - 2 * math.pi * K_E) < 1e-6
    assert oq.self_integral_correction(0.0) == 0.0
    from electrodrive_amd.core import bem_mesh
    m = bem_mesh.icosphere(2)
    I, corr = oq.self_integrals(m.centroids, m.panel_vertices, m.areas, refined=False)
    assert np.allclose(I, 2 * np.pi * K_E * np.sqrt(m.areas / np.pi))
    assert np.allclose(corr * m.areas, I)


def test_near_pairs_brute_force():
    from electrodrive_amd.core import bem_mesh
    for m in (bem_mesh.icosphere(2), bem_mesh.square_plate(1.0, 6)):
        C, A = m.centroids, m.areas
        R = np.sqrt(A / math.pi)
        ref = [(i, j) for i in range(len(A)) for j in range(len(A))
               if i != j and 0 < math.sqrt(((C[i] - C[j]) ** 2).sum()) <= 1.5 * (R[i] + R[j])]
        got = oq.near_pairs(C, A, 1.5)
        assert [tuple(x) for x in got.tolist()] == ref
        # symmetric relation
        s = set(map(tuple, got.tolist()))
        assert all((j, i) in s for i, j in s)


def test_near_weights_limits():
    from electrodrive_amd.core import bem_mesh
    m = bem_mesh.icosphere(2)
    pairs = oq.near_pairs(m.centroids, m.areas, 1.5)
    w0 = oq.near_weights(m.centroids, m.areas, m.panel_vertices, pairs, 2, None)
    # depth 0 with the 1-point rule is exactly the centroid rule: delta = 0
    w1 = oq.near_weights(m.centroids, m.areas, m.panel_vertices, pairs, 1, None)
    assert np.max(np.abs(w1)) < 1e-6 * K_E * m.areas.max() / 0.1
    # refined rules converge (deltas of depth 4 and 5 agree to 1e-3 relative)
    w4 = oq.near_weights(m.centroids, m.areas, m.panel_vertices, pairs[:50], 2, 4)
    w5 = oq.near_weights(m.centroids, m.areas, m.panel_vertices, pairs[:50], 2, 5)
    assert np.max(np.abs(w4 - w5)) < 1e-3 * np.max(np.abs(w5))
    assert w0.shape == (pairs.shape[0],)
    V = oq.near_apply(np.zeros(m.n_panels), np.ones(m.n_panels), pairs, w0)
    ref = np.zeros(m.n_panels)
    for (i, j), w in zip(pairs, w0):
        ref[i] += w
    assert np.allclose(V, ref)


def test_near_weights_orientation_vs_exact_integral():
    """delta_ij is panel j's integral seen from the centroid of panel i minus
    its lumped term (core/bem.py:558-667), pinned against scipy's adaptive
    2-D quadrature of that integral at depth 5.  The panels differ 22x in
    area, so swapping the roles of i and j (panel i seen from C_j) moves the
    value by ~99 % and fails."""
    from scipy.integrate import dblquad
    vj = np.array([[0.0, 0, 0], [2.0, 0, 0], [0.0, 2.0, 0]])
    vi = np.array([[0.55, 0.45, 0.35], [0.95, 0.55, 0.35], [0.6, 0.9, 0.45]])

    def area(v):
        return 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))

    def panel_integral(v, P):
        e1, e2 = v[1] - v[0], v[2] - v[0]
        val = dblquad(lambda t, u: 1.0 / np.linalg.norm(v[0] + u * e1 + t * e2 - P),
                      0, 1, lambda u: 0, lambda u: 1 - u, epsabs=1e-13, epsrel=1e-12)[0]
        return 2 * area(v) * val

    Ai, Aj = area(vi), area(vj)
    assert Aj / Ai > 4
    Ci, Cj = vi.mean(0), vj.mean(0)
    d = np.linalg.norm(Ci - Cj)
    exact_ij = K_E * (panel_integral(vj, Ci) - Aj / d)
    exact_ji = K_E * (panel_integral(vi, Cj) - Ai / d)
    w = oq.near_weights(np.stack([Ci, Cj]), np.array([Ai, Aj]), np.stack([vi, vj]),
                        np.array([[0, 1], [1, 0]]), 2, 5)
    assert abs(w[0] - exact_ij) <= 1e-3 * abs(exact_ij)
    assert abs(w[1] - exact_ji) <= 1e-3 * abs(exact_ji)
    assert abs(exact_ji - exact_ij) > 0.1 * abs(exact_ij)  # a swap would fail


# -------------------------------------------------------------------- gmres
def test_gmres_known_answer_spd():
    """tests/test_gmres.py:12-76 of the reference: SPD n=8 systems, with and
    without the diagonal preconditioner, solution error < 1e-5."""
    rng = np.random.default_rng(0)
    for pre in (None, "jacobi"):
        A = rng.normal(size=(8, 8))
        A = A.T @ A + 0.1 * np.eye(8)
        xt = rng.normal(size=8)
        b = A @ xt
        x, info = og.gmres_restart(lambda v: A @ v, b, restart=8, tol=1e-10, maxiter=200,
                                   precond=pre, A_diag=np.diag(A))
        assert info["success"] and info["code"] == 0
        assert np.linalg.norm(x - xt) < 1e-5
        assert np.linalg.norm(b - A @ x) <= info["tol_abs"]
    # larger system vs numpy solve; iteration count is a whole number of cycles
    n = 60
    A = np.eye(n) * 3 + rng.normal(size=(n, n)) * 0.2
    b = rng.normal(size=n)
    x, info = og.gmres_restart(lambda v: A @ v, b, restart=10, tol=1e-12, maxiter=500)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-11)
    assert info["iters"] % 10 == 0


def test_gmres_failure_codes():
    A = np.eye(4)
    x, info = og.gmres_restart(lambda v: A @ v, np.array([1.0, np.nan, 0, 0]))
    assert info["code"] == 3
    x, info = og.gmres_restart(lambda v: np.diag([1, 1e-6, 1e6, 1]) @ v, np.ones(4), restart=1,
                               maxiter=2, tol=1e-14)
    assert info["code"] == 1 and not info["success"]
    x, info = og.gmres_restart(lambda v: A @ v, np.zeros(4))
    assert info["code"] == 0


# ----------------------------------------------------------------- analytic
def test_kelvin_closed_forms():
    R, q = 1.0, 1.0
    r0 = np.array([0.3, -1.1, 0.9])
    d = np.linalg.norm(r0)
    qi, ri = analytic.kelvin_image(R, r0, q)
    rng = np.random.default_rng(0)
    u = rng.normal(size=(100, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    phi = analytic.point_charge_potential(u, r0, q) + analytic.point_charge_potential(u, ri, qi)
    assert np.max(np.abs(phi)) < 1e-6 * K_E
    from electrodrive_amd.core import bem_mesh
    m = bem_mesh.icosphere(6)
    Q = np.sum(analytic.kelvin_sigma(m.centroids, R, r0, q) * m.areas)
    assert abs(Q - (-q * R / d)) < 1e-3
    assert abs(analytic.sphere_capacitance(2.0) * K_E - 2.0) < 1e-12


# ------------------------------------------------------------------ fmm math
def test_expansion_identity_and_vacuum_golden_path():
    rng = np.random.default_rng(0)
    y = rng.uniform(-0.3, 0.3, (4, 3))
    x = np.array([1.2, -0.7, 0.9])
    p = 14
    R = fm.regular(y, p)
    S = fm.irregular(x[None], p)[0]
    for i in range(4):
        acc = sum(np.conj(fm.full(R[i], n, m)) * fm.full(S, n, m)
                  for n in range(p + 1) for m in range(-n, n + 1))
        assert abs(acc.real - 1 / np.linalg.norm(x - y[i])) < 1e-7
        assert abs(acc.imag) < 1e-12
    # fmm3d/tests/test_multipole_vacuum.py:30-69: source at (0.1,0.2,-0.1) in a
    # box at 0, target (3,0,0)+(-0.1,0.1,0.2), p = 8, unit scale: rel err < 1e-5
    src = np.array([[0.1, 0.2, -0.1]])
    M = fm.p2m(src, [1.0], np.zeros(3), 1.0, 8)
    L = fm.m2l_matrix((3.0, 0.0, 0.0), 8) @ M
    tgt = np.array([[2.9, 0.1, 0.2]])
    phi = fm.l2p(L, tgt, np.array([3.0, 0, 0]), 1.0, 8)
    ref = 1.0 / np.linalg.norm(tgt - src)
    assert abs(phi[0] - ref) / ref < 1e-5
    # a sign flip in M2L must fail (fmm3d/tests/test_multipole_vacuum.py:77-114)
    Lbad = fm.m2l_matrix((-3.0, 0.0, 0.0), 8) @ M
    assert abs(fm.l2p(Lbad, tgt, np.array([3.0, 0, 0]), 1.0, 8)[0] - ref) / ref > 1e-2


@pytest.mark.parametrize("off", [(2, 0, 0), (2, 2, 2), (3, -1, 2), (-2, 3, 0)])
def test_m2l_truncation_bound(off):
    rng = np.random.default_rng(5)
    src = rng.uniform(-0.5, 0.5, (60, 3))
    q = rng.normal(size=60)
    o = np.asarray(off, dtype=float)
    tgt = o + rng.uniform(-0.5, 0.5, (40, 3))
    ref = np.array([np.sum(q / np.linalg.norm(t - src, axis=1)) for t in tgt])
    prev = None
    for p in (4, 8, 12):
        L = fm.m2l_matrix(o, p) @ fm.p2m(src, q, np.zeros(3), 1.0, p)
        err = np.max(np.abs(fm.l2p(L, tgt, o, 1.0, p) - ref)) / np.sum(np.abs(q))
        rho = math.sqrt(3) / np.linalg.norm(o)  # (r_s + r_t) / |o|
        assert err < 2 * rho ** (p + 1) / (1 - rho) + 1e-12
        if prev is not None:
            assert err < prev
        prev = err


def test_m2m_l2l_exact():
    rng = np.random.default_rng(9)
    p = 8
    for oc in range(8):
        src = rng.uniform(-0.25, 0.25, (30, 3))
        q = rng.normal(size=30)
        cc = fm.child_offset(oc) * 2.0
        Mc = fm.p2m(cc + src, q, cc, 1.0, p)
        Mp = fm.m2m_matrix(fm.child_offset(oc), p) @ Mc
        assert np.allclose(Mp, fm.p2m(cc + src, q, np.zeros(3), 2.0, p), atol=1e-12)
        L = rng.normal(size=(p + 1) ** 2)
        Lc = fm.l2l_matrix(fm.child_offset(oc), p) @ L
        t = cc + rng.uniform(-0.5, 0.5, (10, 3))
        assert np.allclose(fm.l2p(Lc, t, cc, 1.0, p), fm.l2p(L, t, np.zeros(3), 2.0, p),
                           rtol=1e-11, atol=1e-11)


def test_sai_oracle_exact_inverse_and_improves_on_jacobi():
    """Pins oracle.sai: with J = I = every panel (20-panel icosahedron, huge
    factors) each row solves m^T A = e_i^T exactly, so M A = I; on a larger
    mesh |I - M A|_F is far below Jacobi's |I - D^-1 A|_F."""
    import numpy as np
    from oracle import sai, quadrature as oq
    from electrodrive_amd.core import bem_mesh
    m = bem_mesh.icosphere(0)
    I_self, corr = oq.self_integrals(m.centroids, m.panel_vertices, m.areas, depth=2)
    A = sai.operator_entries(m.centroids, m.areas, corr)
    M = sai.sai_matrix(m.centroids, m.areas, corr, factor_pattern=1e3, factor_rows=1e3)
    assert np.abs(M @ A - np.eye(A.shape[0])).max() < 1e-8
    m = bem_mesh.icosphere(2)
    I_self, corr = oq.self_integrals(m.centroids, m.panel_vertices, m.areas, depth=2)
    A = sai.operator_entries(m.centroids, m.areas, corr)
    M = sai.sai_matrix(m.centroids, m.areas, corr)
    eye = np.eye(A.shape[0])
    e_sai = np.linalg.norm(eye - M @ A)
    e_jac = np.linalg.norm(eye - A / np.diag(A)[:, None])
    print("SAI", e_sai, "Jacobi", e_jac)
    assert e_sai < 0.5 * e_jac


def test_direct_E_field_oracle_is_minus_gradient():
    """Pins oracle.dense.direct_E_field: E = -grad V of direct_potential
    (central differences), and the exterior field of a uniformly charged
    sphere is K_E Q / r^2 radially."""
    import numpy as np
    from oracle import dense as od
    from oracle.constants import K_E
    from electrodrive_amd.core import bem_mesh
    rng = np.random.default_rng(4)
    m = bem_mesh.icosphere(2)
    sig = rng.normal(size=m.n_panels)
    X = rng.uniform(-2, 2, (5, 3)) * 1.5
    E = od.direct_E_field(X, m.centroids, m.areas, sig)
    h = 1e-5
    for k in range(3):
        dx = np.zeros(3); dx[k] = h
        g = (od.direct_potential(X + dx, m.centroids, m.areas, sig)
             - od.direct_potential(X - dx, m.centroids, m.areas, sig)) / (2 * h)
        assert np.allclose(E[:, k], -g, rtol=1e-6, atol=1e-9 * np.abs(E).max())
    m = bem_mesh.icosphere(4)
    x = np.array([[0.0, 0.0, 3.0], [2.0, 2.0, 1.0]])
    E = od.direct_E_field(x, m.centroids, m.areas, np.ones(m.n_panels))
    Q = m.areas.sum()
    r = np.linalg.norm(x, axis=1)
    ref = (K_E * Q / r ** 3)[:, None] * x
    assert np.allclose(E, ref, rtol=2e-3, atol=1e-6 * np.abs(ref).max())


def test_near_target_oracle_pins():
    """Pins oracle.quadrature.near_target_potentials: no near panel -> V
    unchanged; a target on a panel centroid swaps the lumped K_E A / 1e-12
    term for that panel's refined self integral (the self-integral oracle);
    an off-plane target's refined integral approaches the exact panel
    integral (adaptive 2-D quadrature) as the depth grows."""
    import numpy as np
    from scipy import integrate
    from oracle import quadrature as oq
    from oracle.constants import K_E
    from electrodrive_amd.core import bem_mesh
    m = bem_mesh.icosphere(2)
    rng = np.random.default_rng(9)
    sig = rng.normal(size=m.n_panels)
    far = np.array([[5.0, 0.0, 0.0]])
    assert oq.near_target_potentials(np.array([1.0]), far, m.centroids, m.areas, sig,
                                     m.panel_vertices, 1.5)[0] == 1.0
    j = 17
    x = m.centroids[j:j + 1]
    V = oq.near_target_potentials(np.array([0.0]), x, m.centroids, m.areas, sig,
                                  m.panel_vertices, 1.5, max_depth=3)
    I_self, _ = oq.self_integrals(m.centroids[j:j + 1], m.panel_vertices[j:j + 1],
                                  m.areas[j:j + 1], depth=3)
    d = np.linalg.norm(m.centroids - x, axis=1)
    rad = np.sqrt(m.areas / np.pi)
    expect = sig[j] * (I_self[0] - K_E * m.areas[j] / 1e-12)
    for k in np.nonzero((d <= 1.5 * rad) & (np.arange(m.n_panels) != j))[0]:
        pts, w = oq.near_singular_quadrature(x[0], m.panel_vertices[k], "telles", 2, 3)
        r = np.maximum(np.linalg.norm(pts - x, axis=1), 1e-12)
        expect += sig[k] * (np.sum(K_E / r * w) - K_E * m.areas[k] / d[k])
    assert abs(V[0] - expect) <= 1e-9 * abs(expect)
    # flat right triangle: a distant target is integrated to quadrature accuracy
    # at every depth; a target 0.05 above the panel stays within the few-percent
    # plateau of the reference rule (it refines only toward the nearest sub-panel)
    tri = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    for tgt, tol in ((np.array([0.3, 0.3, 3.0]), 1e-4), (np.array([0.3, 0.3, 0.05]), 0.05)):
        exact, _ = integrate.dblquad(
            lambda y, xx: K_E / np.sqrt((xx - tgt[0]) ** 2 + (y - tgt[1]) ** 2 + tgt[2] ** 2),
            0.0, 1.0, lambda xx: 0.0, lambda xx: 1.0 - xx, epsabs=1e-9 * K_E)
        errs = []
        for depth in (0, 2, 4, 6):
            pts, w = oq.near_singular_quadrature(tgt, tri, "telles", 2, depth)
            val = np.sum(K_E / np.linalg.norm(pts - tgt, axis=1) * w)
            errs.append(abs(val - exact) / exact)
        assert max(errs[1:]) < tol
        if tol > 1e-3:
            assert errs[0] > 0.1  # unrefined: ~20 % off, the reason for the refinement
