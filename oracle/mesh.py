"""Panel geometry restated from the reference mesher (numpy, FP64).

Test infrastructure only (see oracle/__init__.py).

- ``panel_quantities``  core/bem_mesh.py:103-126 (centroids, areas, degenerate
  drop at area <= 1e-16, unit normals cross/(2A + 1e-30)).
- ``torus``             core/bem_mesh.py:180-212 (tensor-product torus).
- ``plane_patch``       core/bem_mesh.py:139-151.
- ``uv_sphere``         core/bem_mesh.py:154-177 (latitude-longitude grid,
  pole rings included), point by point.
- ``spec_mesh_sizes``   the sizing rules of generate_mesh,
  core/bem_mesh.py:27-99 (grid counts per conductor).

Parity: pinned by closed forms in tests (total area of a flat patch, torus area
4*pi^2*R*a and sphere area 4*pi*r^2 in the fine-mesh limit with O(h^2)
convergence, every sphere vertex at distance r, 2 n_phi degenerate pole
triangles dropped, unit normals, outward orientation).
"""
import math

import numpy as np


def panel_quantities(vertices, triangles):
    """core/bem_mesh.py:103-126."""
    v = np.asarray(vertices, dtype=float)
    t = np.asarray(triangles, dtype=np.int64)
    v0, v1, v2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    centroids = (v0 + v1 + v2) / 3.0
    cross = np.cross(v1 - v0, v2 - v0)
    areas = 0.5 * np.linalg.norm(cross, axis=1)
    mask = areas > 1e-16
    t, centroids, cross, areas = t[mask], centroids[mask], cross[mask], areas[mask]
    normals = cross / (2.0 * areas[:, None] + 1e-30)
    return t, centroids, normals, areas


def torus(center, R, a, n_u, n_v):
    """core/bem_mesh.py:180-212 (vectorised; same vertex and triangle order)."""
    u = np.linspace(0.0, 2.0 * np.pi, n_u, endpoint=False)
    w = np.linspace(0.0, 2.0 * np.pi, n_v, endpoint=False)
    uu, ww = np.meshgrid(u, w, indexing="ij")
    x = (R + a * np.cos(ww)) * np.cos(uu) + center[0]
    y = (R + a * np.cos(ww)) * np.sin(uu) + center[1]
    z = a * np.sin(ww) + center[2]
    verts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    i = np.arange(n_u)[:, None]
    j = np.arange(n_v)[None, :]
    inext = (i + 1) % n_u
    jnext = (j + 1) % n_v
    a_idx = (i * n_v + j).ravel()
    b_idx = (i * n_v + jnext).ravel()
    c_idx = (inext * n_v + j).ravel()
    d_idx = (inext * n_v + jnext).ravel()
    tris = np.empty((n_u * n_v * 2, 3), dtype=np.int64)
    tris[0::2] = np.stack([a_idx, b_idx, c_idx], axis=1)
    tris[1::2] = np.stack([b_idx, d_idx, c_idx], axis=1)
    return verts, tris


def plane_patch(z, L, n):
    """core/bem_mesh.py:139-151."""
    x = np.linspace(-L / 2.0, L / 2.0, n + 1)
    y = np.linspace(-L / 2.0, L / 2.0, n + 1)
    xx, yy = np.meshgrid(x, y, indexing="ij")
    verts = np.stack([xx.ravel(), yy.ravel(), np.full(xx.size, z)], axis=1)
    tris = []
    for i in range(n):
        for j in range(n):
            idx = i * (n + 1) + j
            tris.append([idx, idx + 1, idx + (n + 1)])
            tris.append([idx + 1, idx + (n + 2), idx + (n + 1)])
    return verts, np.asarray(tris, dtype=np.int64)


def uv_sphere(center, radius, n_theta, n_phi):
    """core/bem_mesh.py:154-177: vertex (i, j) at polar angle theta_i =
    pi i / (n_theta - 1), azimuth phi_j = 2 pi j / n_phi; cell (i, j) gives
    triangles (a, b, c) and (b, d, c) with a = (i, j), b = (i, j+1 mod n_phi),
    c = (i+1, j), d = (i+1, j+1 mod n_phi)."""
    theta = np.linspace(0.0, np.pi, n_theta)
    phi = np.linspace(0.0, 2.0 * np.pi, n_phi, endpoint=False)
    verts = []
    for t in theta:
        st, ct = math.sin(t), math.cos(t)
        for p in phi:
            verts.append([center[0] + radius * st * math.cos(p),
                          center[1] + radius * st * math.sin(p),
                          center[2] + radius * ct])
    tris = []
    for i in range(n_theta - 1):
        for j in range(n_phi):
            a = i * n_phi + j
            b = i * n_phi + (j + 1) % n_phi
            c = (i + 1) * n_phi + j
            d = (i + 1) * n_phi + (j + 1) % n_phi
            tris += [[a, b, c], [b, d, c]]
    return np.array(verts, dtype=float), np.array(tris, dtype=np.int64)


def spec_mesh_sizes(conductors, charges, target_h, sphere_density=1.5, plane_scale=1.0):
    """Grid counts generate_mesh uses per conductor (core/bem_mesh.py:27-99):
    plane -> (L, n), sphere -> (n_theta, n_phi), torus -> (n_u, n_v)."""
    h = max(1e-3, target_h)
    out = []
    for c in conductors:
        kind = c.get("type")
        if kind == "plane":
            z = float(c.get("z", 0.0))
            d = max([0.3] + [abs(float(q["pos"][2]) - z) for q in charges
                             if q.get("type") == "point"])
            L = plane_scale * max(8.0, 16.0 * d)
            out.append(("plane", L, max(8, int(L / h))))
        elif kind == "sphere":
            r = float(c.get("radius", 1.0))
            k = max(1.0, sphere_density)
            out.append(("sphere", max(10, math.ceil(k * math.pi * r / h)),
                        max(20, math.ceil(2.0 * k * math.pi * r / h))))
        elif kind in ("torus", "toroid"):
            R = float(c.get("major_radius", c.get("radius", 1.0)))
            a = float(c.get("minor_radius", 0.25 * R))
            out.append(("torus", max(24, math.ceil(2.0 * math.pi * R / h)),
                        max(16, math.ceil(2.0 * math.pi * a / h))))
        else:
            raise NotImplementedError(kind)
    return out
