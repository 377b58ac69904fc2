"""Panel geometry restated from the reference mesher (numpy, FP64).

Test infrastructure only (see oracle/__init__.py).

- ``panel_quantities``  core/bem_mesh.py:103-126 (centroids, areas, degenerate
  drop at area <= 1e-16, unit normals cross/(2A + 1e-30)).
- ``torus``             core/bem_mesh.py:180-212 (tensor-product torus).
- ``plane_patch``       core/bem_mesh.py:139-151.

Parity: pinned by closed forms in tests (total area of a flat patch, torus area
4*pi^2*R*a in the fine-mesh limit, unit normals, outward orientation).
"""
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
