"""Dense single-layer collocation matvec restated from the reference (numpy).

Test infrastructure only (see oracle/__init__.py).

- ``dense_matvec``  core/bem_kernel.py:157-230 (``_bem_matvec_core_torch``):
      V_i = sum_j K_ij * sigma_j * A_j,   K_ij = K_E * rsqrt(max(r_ij^2, 1e-24)),
      K_ii = self_integrals[i] (per-area self term) or 0.
  Row tiles of ``tile_size`` as in the reference (the tile size only changes
  summation order).
- ``dense_matvec_batched``  core/bem_kernel.py:715-810 (B right-hand sides).
- ``direct_potential``  core/bem_kernel.py:233-276 (potential at targets).

Parity: pinned by brute-force double loops on tiny inputs and by the
closed-form potential of a uniformly charged sphere in tests.
"""
import numpy as np

from .constants import K_E


def dense_matvec(centroids, areas, sigma, self_integrals=None, tile_size=1024,
                 dtype=np.float64):
    C = np.asarray(centroids, dtype=dtype)
    A = np.asarray(areas, dtype=dtype)
    s = np.asarray(sigma, dtype=dtype)
    N = s.shape[0]
    V = np.zeros(N, dtype=dtype)
    if N == 0:
        return V
    sA = s * A
    kE = dtype(K_E) if dtype is np.float32 else K_E
    for i0 in range(0, N, tile_size):
        i1 = min(i0 + tile_size, N)
        diff = C[i0:i1, None, :] - C[None, :, :]
        r2 = np.maximum((diff * diff).sum(axis=-1), dtype(1e-24))
        K = kE / np.sqrt(r2)
        rows = np.arange(i1 - i0)
        cols = np.arange(i0, i1)
        if self_integrals is not None:
            K[rows, cols] = np.asarray(self_integrals, dtype=dtype)[cols]
        else:
            K[rows, cols] = 0.0
        V[i0:i1] = K @ sA
    return V


def dense_matvec_batched(centroids, areas, sigma_b, self_integrals=None,
                         tile_size=1024):
    sig = np.asarray(sigma_b, dtype=float)
    return np.stack([dense_matvec(centroids, areas, sig[b], self_integrals,
                                  tile_size) for b in range(sig.shape[0])])


def direct_potential(targets, centroids, areas, sigma):
    """core/bem_kernel.py:233-276: V(x_i) = sum_j K_E sigma_j A_j / |x_i - c_j|."""
    T = np.asarray(targets, dtype=float)
    C = np.asarray(centroids, dtype=float)
    sA = np.asarray(sigma, dtype=float) * np.asarray(areas, dtype=float)
    out = np.zeros(T.shape[0])
    for i0 in range(0, T.shape[0], 1024):
        d = T[i0:i0 + 1024, None, :] - C[None, :, :]
        r2 = np.maximum((d * d).sum(-1), 1e-24)
        out[i0:i0 + 1024] = (K_E / np.sqrt(r2)) @ sA
    return out


def direct_E_field(targets, centroids, areas, sigma):
    """core/bem_kernel.py:280-335:
    E(x_i) = sum_j K_E sigma_j A_j (x_i - c_j) / max(|x_i - c_j|^2, 1e-24)^(3/2)."""
    T = np.asarray(targets, dtype=float)
    C = np.asarray(centroids, dtype=float)
    sA = np.asarray(sigma, dtype=float) * np.asarray(areas, dtype=float)
    out = np.zeros((T.shape[0], 3))
    for i0 in range(0, T.shape[0], 512):
        d = T[i0:i0 + 512, None, :] - C[None, :, :]
        r2 = np.maximum((d * d).sum(-1), 1e-24)
        w = K_E * sA[None, :] / (r2 * np.sqrt(r2))
        out[i0:i0 + 512] = np.einsum("tn,tnk->tk", w, d)
    return out
