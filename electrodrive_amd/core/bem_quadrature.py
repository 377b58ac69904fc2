"""Panel quadrature and near-field correction on the GPU.

Native replacements of the reference's set-up loops and near-field apply:

- ``self_integrals``           core/bem.py:1777-1816 (+ bem_quadrature.py:217-308)
- ``build_near_pairs``         core/bem.py:475-555 (bit-exact index set and order)
- ``near_correction_weights``  core/bem.py:558-667
- ``near_quadrature_matvec``   core/bem_near_cuda.cpp:61-233 (pybind signature
                               V_far, sigma, centroids, areas, panel_vertices,
                               near_pairs, K_E, quad_order) and, with weights,
                               _apply_near_quadrature_matvec_torch (core/bem.py:675-734)
- ``self_integral_correction`` core/bem_quadrature.py:316-344 (closed form)
"""
import ctypes
import math

import torch

from .. import _native as nat
from .kernels import K_E


def _dev():
    nat.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def _f64(x, dev):
    return torch.as_tensor(x, dtype=torch.float64).to(dev).contiguous()


def self_integral_correction(area) -> float:
    a = float(area)
    if a <= 0.0:
        return 0.0
    return 2.0 * math.pi * K_E * math.sqrt(a / math.pi)


def self_integrals(centroids, panel_vertices, areas=None, *, depth: int = 3, order: int = 2):
    """Returns (I_self, self_corr = I_self / max(A, tiny)) as FP64 device tensors."""
    dev = _dev()
    C = _f64(centroids, dev).reshape(-1, 3)
    PV = _f64(panel_vertices, dev).reshape(-1, 3, 3)
    n = C.shape[0]
    out = torch.empty(n, dtype=torch.float64, device=dev)
    nat.check(nat.lib().ed_self_integrals_f64(nat.ptr(PV), nat.ptr(C), n, int(depth), int(order),
                                              nat.ptr(out), nat.stream_ptr(dev)),
              "ed_self_integrals_f64")
    if areas is None:
        return out, None
    A = _f64(areas, dev).reshape(-1)
    corr = out / A.clamp_min(torch.finfo(torch.float64).tiny)
    return out, corr


def build_near_pairs(centroids, areas, distance_factor: float = 1.5):
    """(P, 2) int64 device tensor of near pairs (i-major, j ascending)."""
    dev = _dev()
    C = _f64(centroids, dev).reshape(-1, 3)
    A = _f64(areas, dev).reshape(-1)
    n = min(C.shape[0], A.shape[0])
    cnt = ctypes.c_int64(0)
    st = nat.stream_ptr(dev)
    L = nat.lib()
    # one search fills a buffer sized for ~16 partners a row (a 1.5 factor on
    # a surface mesh gives ~10); a fuller result reports its size and the
    # search runs again into an exact buffer
    cap = 16 * max(n, 0) + 1024
    buf = torch.empty((cap, 2), dtype=torch.int64, device=dev)
    nat.check(L.ed_near_pairs_f64(nat.ptr(C), nat.ptr(A), n, float(distance_factor),
                                  nat.ptr(buf), cap, ctypes.byref(cnt), st), "ed_near_pairs_f64")
    P = max(int(cnt.value), 0)
    if P <= cap:
        return buf[:P].clone()
    pairs = torch.empty((P, 2), dtype=torch.int64, device=dev)
    nat.check(L.ed_near_pairs_f64(nat.ptr(C), nat.ptr(A), n, float(distance_factor),
                                  nat.ptr(pairs), P, ctypes.byref(cnt), st), "ed_near_pairs_f64")
    return pairs


def near_correction_weights(centroids, areas, panel_vertices, near_pairs, *, quad_order: int = 2,
                            max_depth=None):
    dev = _dev()
    C = _f64(centroids, dev).reshape(-1, 3)
    A = _f64(areas, dev).reshape(-1)
    PV = _f64(panel_vertices, dev).reshape(-1, 3, 3)
    pairs = torch.as_tensor(near_pairs, dtype=torch.int64).to(dev).contiguous().reshape(-1, 2)
    out = torch.zeros(pairs.shape[0], dtype=torch.float64, device=dev)
    depth = 0 if max_depth is None else max(0, int(max_depth))
    if pairs.shape[0]:
        nat.check(nat.lib().ed_near_weights_f64(nat.ptr(C), nat.ptr(A), nat.ptr(PV),
                                                nat.ptr(pairs), pairs.shape[0], int(quad_order),
                                                depth, nat.ptr(out), nat.stream_ptr(dev)),
                  "ed_near_weights_f64")
    return out


def near_quadrature_matvec(V_far, sigma, centroids, areas, panel_vertices, near_pairs,
                           K_E_value: float = K_E, quad_order: int = 2, weights=None):
    """V_far + sum_j sigma_j delta_ij.  Without ``weights`` delta is the
    unrefined order-``quad_order`` rule evaluated on the fly (CUDA-extension
    semantics); with weights it is the precomputed correction."""
    if abs(float(K_E_value) - K_E) > 1e-6 * K_E:
        raise ValueError("only the reference K_E is supported")
    if quad_order not in (1, 2):
        raise ValueError("only quad_order 1 or 2 are supported")
    dev = V_far.device if V_far.is_cuda else _dev()
    vf = V_far.to(dev, torch.float64).contiguous()
    s = sigma.to(dev, torch.float64).contiguous()
    n = vf.shape[0]
    pairs = torch.as_tensor(near_pairs, dtype=torch.int64).to(dev).contiguous().reshape(-1, 2)
    if pairs.shape[0] > 1 and bool((pairs[1:, 0] < pairs[:-1, 0]).any()):
        order = torch.argsort(pairs[:, 0], stable=True)
        pairs = pairs[order]
        if weights is not None:
            weights = torch.as_tensor(weights).to(dev)[order]
    C = _f64(centroids, dev).reshape(-1, 3)
    A = _f64(areas, dev).reshape(-1)
    PV = None if weights is not None else _f64(panel_vertices, dev).reshape(-1, 3, 3)
    W = None if weights is None else torch.as_tensor(weights, dtype=torch.float64).to(dev).contiguous()
    out = torch.empty_like(vf)
    nat.check(nat.lib().ed_near_apply_f64(nat.ptr(vf), nat.ptr(s), nat.ptr(C), nat.ptr(A),
                                          nat.ptr(PV), nat.ptr(pairs), nat.ptr(W), pairs.shape[0],
                                          n, int(quad_order), nat.ptr(out), nat.stream_ptr(dev)),
              "ed_near_apply_f64")
    return out.to(V_far.device, V_far.dtype)


def near_quadrature_potentials(V_far, targets, centroids, areas, sigma, panel_vertices, *,
                               distance_factor: float, quad_order: int = 2, cache=None,
                               max_depth=None):
    """Near-field correction of target potentials (core/bem.py:940-1047): for
    panels j with |c_j - x_i| <= distance_factor R_j, A_j > 0, sigma_j != 0,
    V_i += sigma_j (I_near - K_E A_j / max(|c_j - x_i|, 1e-12)), I_near refined
    ``max_depth`` times toward x_i (None: 4, the reference default).  ``cache``
    is accepted for signature parity (the native path needs none)."""
    del cache
    if quad_order not in (1, 2):
        raise ValueError("only quad_order 1 or 2 are supported")
    depth = 4 if max_depth is None else int(max_depth)
    dev = V_far.device if V_far.is_cuda else _dev()
    V = V_far.to(dev, torch.float64).contiguous().clone()
    X = _f64(targets, dev).reshape(-1, 3)
    C = _f64(centroids, dev).reshape(-1, 3)
    A = _f64(areas, dev).reshape(-1)
    PV = _f64(panel_vertices, dev).reshape(-1, 3, 3)
    s = _f64(sigma, dev).reshape(-1)
    if X.shape[0] == 0 or s.shape[0] == 0:
        return V_far
    nat.check(nat.lib().ed_near_targets_f64(nat.ptr(X), X.shape[0], nat.ptr(C), nat.ptr(A),
                                            nat.ptr(PV), nat.ptr(s), s.shape[0],
                                            float(distance_factor), depth, int(quad_order),
                                            nat.ptr(V), nat.stream_ptr(dev)),
              "ed_near_targets_f64")
    return V.to(V_far.device, V_far.dtype)
