"""Triangle meshes for the BEM solve path (host numpy; input generation only).

``TriMesh`` and the panel quantities follow core/bem_mesh.py:13-25 and :103-126
(centroid = vertex mean, area = |e1 x e2|/2, panels with area <= 1e-16
dropped, unit normal = cross / (2A + 1e-30)).  ``torus``, ``plane_patch`` and
``uv_sphere`` follow core/bem_mesh.py:180-212, :139-151 and :154-177, and
``generate_mesh`` (meshing a parsed spec at a target edge length) follows
core/bem_mesh.py:27-136.  ``icosphere``, ``sphere_dimer`` and ``square_plate``
have no reference counterpart.
"""
import math
import os
from dataclasses import dataclass

import numpy as np


@dataclass
class TriMesh:
    vertices: np.ndarray       # (Nv, 3)
    triangles: np.ndarray      # (Nt, 3)
    centroids: np.ndarray      # (Nt, 3)
    normals: np.ndarray        # (Nt, 3)
    areas: np.ndarray          # (Nt,)
    conductor_ids: np.ndarray  # (Nt,)

    @property
    def n_panels(self) -> int:
        return int(self.triangles.shape[0])

    @property
    def panel_vertices(self) -> np.ndarray:
        """(Nt, 3, 3) vertex coordinates of each panel."""
        return self.vertices[self.triangles]


def make_trimesh(vertices, triangles, conductor_ids=None) -> TriMesh:
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    t = np.ascontiguousarray(triangles, dtype=np.int64)
    cid = (np.zeros(t.shape[0], dtype=np.int32) if conductor_ids is None
           else np.asarray(conductor_ids, dtype=np.int32))
    v0, v1, v2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    centroids = (v0 + v1 + v2) / 3.0
    cross = np.cross(v1 - v0, v2 - v0)
    areas = 0.5 * np.linalg.norm(cross, axis=1)
    mask = areas > 1e-16
    t, centroids, cross, areas, cid = t[mask], centroids[mask], cross[mask], areas[mask], cid[mask]
    normals = cross / (2.0 * areas[:, None] + 1e-30)
    return TriMesh(v, t, centroids, normals, areas, cid)


def merge(meshes) -> TriMesh:
    """Concatenate meshes (conductor ids kept per input mesh)."""
    verts, tris, cids = [], [], []
    off = 0
    for k, m in enumerate(meshes):
        verts.append(m.vertices)
        tris.append(m.triangles + off)
        cids.append(np.full(m.n_panels, k, dtype=np.int32))
        off += m.vertices.shape[0]
    return make_trimesh(np.vstack(verts), np.vstack(tris), np.concatenate(cids))


_PHI = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_V = np.array([[-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
                   [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
                   [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]], dtype=float)
_ICO_F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                   [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                   [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                   [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)


def icosphere(subdivisions: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)) -> TriMesh:
    """Geodesic sphere: icosahedron, each face split 4-way ``subdivisions``
    times, new vertices projected to the sphere; 20*4^s outward-oriented
    panels (s = 3: 1,280; s = 6: 81,920)."""
    v = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    f = _ICO_F.copy()
    for _ in range(int(subdivisions)):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e_sorted = np.sort(e, axis=1)
        uniq, inv = np.unique(e_sorted, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.vstack([v, mid])
        nf = f.shape[0]
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], axis=1),
            np.stack([m01, f[:, 1], m12], axis=1),
            np.stack([m20, m12, f[:, 2]], axis=1),
            np.stack([m01, m12, m20], axis=1)], axis=0)
    verts = v * float(radius) + np.asarray(center, dtype=float)[None, :]
    return make_trimesh(verts, f)


def torus(R: float, a: float, n_u: int, n_v: int, center=(0.0, 0.0, 0.0)) -> TriMesh:
    """Tensor-product torus (core/bem_mesh.py:180-212), 2*n_u*n_v panels."""
    u = np.linspace(0.0, 2.0 * np.pi, n_u, endpoint=False)
    w = np.linspace(0.0, 2.0 * np.pi, n_v, endpoint=False)
    uu, ww = np.meshgrid(u, w, indexing="ij")
    c = np.asarray(center, dtype=float)
    x = (R + a * np.cos(ww)) * np.cos(uu) + c[0]
    y = (R + a * np.cos(ww)) * np.sin(uu) + c[1]
    z = a * np.sin(ww) + c[2]
    verts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    i = np.arange(n_u)[:, None]
    j = np.arange(n_v)[None, :]
    a_idx = (i * n_v + j).ravel()
    b_idx = (i * n_v + (j + 1) % n_v).ravel()
    c_idx = (((i + 1) % n_u) * n_v + j).ravel()
    d_idx = (((i + 1) % n_u) * n_v + (j + 1) % n_v).ravel()
    tris = np.empty((n_u * n_v * 2, 3), dtype=np.int64)
    tris[0::2] = np.stack([a_idx, b_idx, c_idx], axis=1)
    tris[1::2] = np.stack([b_idx, d_idx, c_idx], axis=1)
    return make_trimesh(verts, tris)


def plane_patch(z: float, L: float, n: int) -> TriMesh:
    """Square patch (core/bem_mesh.py:139-151), 2 n^2 panels."""
    x = np.linspace(-L / 2.0, L / 2.0, n + 1)
    xx, yy = np.meshgrid(x, x, indexing="ij")
    verts = np.stack([xx.ravel(), yy.ravel(), np.full(xx.size, z)], axis=1)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    idx = (i * (n + 1) + j).ravel()
    tris = np.empty((n * n * 2, 3), dtype=np.int64)
    tris[0::2] = np.stack([idx, idx + 1, idx + (n + 1)], axis=1)
    tris[1::2] = np.stack([idx + 1, idx + (n + 2), idx + (n + 1)], axis=1)
    return make_trimesh(verts, tris)


def square_plate(L: float, n: int, center=(0.0, 0.0, 0.0)) -> TriMesh:
    m = plane_patch(0.0, L, n)
    return make_trimesh(m.vertices + np.asarray(center, dtype=float)[None, :], m.triangles)


def sphere_dimer(radius: float, gap: float, subdivisions: int) -> TriMesh:
    """Two icospheres of equal radius separated by ``gap`` along x."""
    d = radius + 0.5 * gap
    return merge([icosphere(subdivisions, radius, (-d, 0.0, 0.0)),
                  icosphere(subdivisions, radius, (d, 0.0, 0.0))])


def torus_for_panels(R: float, a: float, target_panels: int) -> TriMesh:
    """Torus whose n_u : n_v ratio follows R : a with about target panels."""
    ratio = R / a
    n_v = max(3, int(round((target_panels / (2.0 * ratio)) ** 0.5)))
    n_u = max(3, int(round(target_panels / (2.0 * n_v))))
    return torus(R, a, n_u, n_v)


def uv_sphere(center, radius: float, n_theta: int, n_phi: int) -> TriMesh:
    """Latitude-longitude sphere (core/bem_mesh.py:154-177): n_theta rings from
    pole to pole (both poles included, n_phi copies each), n_phi meridians;
    two triangles per cell, the degenerate ones at the poles dropped."""
    c = np.asarray(center, dtype=float)
    theta = np.linspace(0.0, np.pi, n_theta)
    phi = np.linspace(0.0, 2.0 * np.pi, n_phi, endpoint=False)
    tt, pp = np.meshgrid(theta, phi, indexing="ij")
    rs = radius * np.sin(tt)
    verts = np.stack([(c[0] + rs * np.cos(pp)).ravel(), (c[1] + rs * np.sin(pp)).ravel(),
                      (c[2] + radius * np.cos(tt)).ravel()], axis=1)
    i = np.arange(n_theta - 1)[:, None]
    j = np.arange(n_phi)[None, :]
    jn = (j + 1) % n_phi
    a_idx = (i * n_phi + j).ravel()
    b_idx = (i * n_phi + jn).ravel()
    c_idx = ((i + 1) * n_phi + j).ravel()
    d_idx = ((i + 1) * n_phi + jn).ravel()
    tris = np.empty(((n_theta - 1) * n_phi * 2, 3), dtype=np.int64)
    tris[0::2] = np.stack([a_idx, b_idx, c_idx], axis=1)
    tris[1::2] = np.stack([b_idx, d_idx, c_idx], axis=1)
    return make_trimesh(verts, tris)


def _env_scale(name: str, default: float) -> float:
    try:
        return float(os.getenv(name, str(default)))
    except ValueError:
        return default


def generate_mesh(spec, target_h: float = 0.1, logger=None) -> TriMesh:
    """Mesh every conductor of a spec at edge length ~target_h
    (core/bem_mesh.py:27-136), with the reference's sizing rules:

    - plane at z: square patch of side L = s * max(8, 16 d), d = max(0.3,
      largest |z_q - z| of the point charges), n = max(8, int(L / h)) cells a
      side (s = EDE_BEM_PLANE_PATCH_SCALE, default 1);
    - sphere: n_theta = max(10, ceil(k pi r / h)), n_phi = max(20,
      ceil(2 k pi r / h)) with k = max(1, EDE_BEM_SPHERE_DENSITY_SCALE),
      default 1.5;
    - torus / toroid: n_u = max(24, ceil(s_u 2 pi R / h)), n_v = max(16,
      ceil(s_v 2 pi a / h)), R = major_radius (or radius, default 1),
      a = minor_radius (default R / 4).

    h = max(1e-3, target_h).  Panels carry the conductor's ``id`` (default:
    its index); other conductor types raise NotImplementedError."""
    if logger:
        logger.info("Mesh generation started.", target_h=target_h)
    h = max(1e-3, float(target_h))
    verts, tris, cids = [], [], []
    off = 0
    for idx, c in enumerate(spec.conductors):
        kind = c.get("type")
        cid = int(c.get("id", idx))
        if kind == "plane":
            z = float(c.get("z", 0.0))
            dmin = 0.3
            for ch in spec.charges:
                if ch.get("type") == "point":
                    dmin = max(dmin, abs(float(ch["pos"][2] - z)))
            L = _env_scale("EDE_BEM_PLANE_PATCH_SCALE", 1.0) * max(8.0, 16.0 * dmin)
            n = max(8, int(L / h))
            part = plane_patch(z, L, n)
        elif kind == "sphere":
            ctr = np.array(c.get("center", [0.0, 0.0, 0.0]), dtype=float)
            r = float(c.get("radius", 1.0))
            k = max(1.0, _env_scale("EDE_BEM_SPHERE_DENSITY_SCALE", 1.5))
            n_theta = max(10, int(math.ceil(k * math.pi * r / h)))
            n_phi = max(20, int(math.ceil(k * 2.0 * math.pi * r / h)))
            part = uv_sphere(ctr, r, n_theta, n_phi)
        elif kind in ("torus", "toroid"):
            ctr = np.array(c.get("center", [0.0, 0.0, 0.0]), dtype=float)
            R = float(c.get("major_radius", c.get("radius", 1.0)))
            a = float(c.get("minor_radius", 0.25 * R))
            n_u = max(24, int(math.ceil(_env_scale("EDE_BEM_TORUS_MAJOR_SCALE", 1.0)
                                        * (2.0 * math.pi * R) / h)))
            n_v = max(16, int(math.ceil(_env_scale("EDE_BEM_TORUS_MINOR_SCALE", 1.0)
                                        * (2.0 * math.pi * a) / h)))
            part = torus(R, a, n_u, n_v, center=ctr)
        else:
            raise NotImplementedError(f"Unsupported conductor type: {kind}")
        # the parts keep every vertex; only their triangle lists were filtered
        verts.append(part.vertices)
        tris.append(part.triangles + off)
        cids.append(np.full(part.n_panels, cid, dtype=np.int32))
        off += part.vertices.shape[0]
    if not verts:
        return TriMesh(np.empty((0, 3)), np.empty((0, 3), dtype=np.int64), np.empty((0, 3)),
                       np.empty((0, 3)), np.empty(0), np.empty(0, dtype=np.int32))
    mesh = make_trimesh(np.vstack(verts), np.vstack(tris), np.concatenate(cids))
    if logger:
        logger.info("Mesh generation complete.", n_panels=mesh.n_panels,
                    total_area=float(mesh.areas.sum()))
    return mesh
