"""BEM matvec dispatcher and dense single-layer matvec (mirror of core/bem_kernel.py).

``bem_matvec_gpu`` keeps the reference signature and dispatch rules
(core/bem_kernel.py:558-707): backends {"auto", "torch_tiled", "keops",
"external"}; unknown -> ValueError; "external" needs ``matvec_impl``;
``self_correction`` is an alias of ``self_integrals``; ``use_near_quad`` is
accepted and ignored.  "auto" / "torch_tiled" / "keops" all run the native
dense kernel (``ed_dense_matvec_f64/f32``), which computes exactly what
``_bem_matvec_core_torch`` does (core/bem_kernel.py:157-230):
    V_i = sum_{j != i} K_E sigma_j A_j rsqrt(max(r_ij^2, 1e-24)) + K_ii sigma_i A_i,
    K_ii = self_integrals[i] or 0, non-finite outputs clamped to 0.
Inputs may be CPU tensors; they are moved to the current HIP device and the
result is returned on sigma's device.
"""
from typing import Callable, Optional

import torch

from .. import _native as nat
from .kernels import DEFAULT_SINGLE_LAYER_KERNEL, K_E, LaplaceSingleLayerKernel

__all__ = ["bem_matvec_gpu", "bem_matvec_batched", "dense_matvec", "K_E",
           "bem_potential_targets", "bem_E_field_targets",
           "_bem_matvec_core_torch", "_bem_potential_targets_core_torch",
           "_bem_E_field_targets_core_torch", "bem_matvec_diff", "bem_potential_targets_diff",
           "bem_E_field_targets_diff", "LaplaceSingleLayerKernel", "DEFAULT_SINGLE_LAYER_KERNEL"]


def _ensure_tile_size(tile_size: Optional[int]) -> int:
    """core/bem_kernel.py:31-52 (kept for signature parity; the native kernel
    tiles internally)."""
    import os
    if tile_size is not None and tile_size > 0:
        return int(tile_size)
    env = os.getenv("EDE_BEM_TILE_SIZE")
    if env is not None:
        try:
            v = int(env)
            if v > 0:
                return v
        except ValueError:
            pass
    return 1024


def dense_matvec(sigma, centroids, areas, self_integrals=None):
    """Native dense matvec; sigma (N,) or (B, N).  Returns a new tensor on
    sigma's device with sigma's dtype."""
    nat.require_gpu()
    dt = sigma.dtype
    if dt not in (torch.float32, torch.float64):
        raise ValueError("sigma must be float32 or float64")
    dev = torch.device("cuda", torch.cuda.current_device()) if not sigma.is_cuda else sigma.device
    s = sigma.to(dev).contiguous()
    squeeze = s.dim() == 1
    if squeeze:
        s = s.unsqueeze(0)
    if s.dim() != 2:
        raise ValueError("sigma must have shape (N,) or (B, N).")
    B, N = s.shape
    C = centroids.to(dev, dt).contiguous()
    A = areas.to(dev, dt).contiguous()
    if C.shape != (N, 3) or A.shape != (N,):
        raise ValueError("centroids/areas shape mismatch with sigma")
    si = None if self_integrals is None else self_integrals.to(dev, dt).contiguous()
    out = torch.empty_like(s)
    fn = nat.lib().ed_dense_matvec_f64 if dt == torch.float64 else nat.lib().ed_dense_matvec_f32
    with torch.cuda.device(dev):
        nat.check(fn(nat.ptr(C), nat.ptr(A), nat.ptr(s), nat.ptr(si), nat.ptr(out), N, B,
                     nat.stream_ptr(dev)), "ed_dense_matvec")
    out = out[0] if squeeze else out
    return out.to(sigma.device)


def bem_matvec_gpu(sigma, src_centroids, areas, *, tile_size: int = 1024,
                   self_integrals=None, logger=None, use_keops: bool = False, kernel=None,
                   backend: str = "auto", matvec_impl: Optional[Callable] = None, **kwargs):
    if "self_correction" in kwargs and self_integrals is None:
        self_integrals = kwargs.get("self_correction")
    if backend not in {"auto", "torch_tiled", "keops", "external"}:
        raise ValueError(f"Unknown BEM backend: {backend!r}")
    ker = kernel or DEFAULT_SINGLE_LAYER_KERNEL
    tile_size = _ensure_tile_size(tile_size)
    if backend == "external":
        if matvec_impl is None:
            raise ValueError("backend='external' requires matvec_impl.")
        return matvec_impl(sigma=sigma, src_centroids=src_centroids, areas=areas,
                           tile_size=tile_size, self_integrals=self_integrals, logger=logger,
                           kernel=ker, **kwargs)
    if not isinstance(ker, LaplaceSingleLayerKernel):
        raise ValueError("only the Laplace single-layer kernel is implemented natively")
    del use_keops
    return dense_matvec(sigma, src_centroids, areas, self_integrals)


def bem_matvec_batched(sigma, src_centroids, areas, *, tile_size: int = 1024,
                       self_integrals=None, kernel=None, logger=None):
    """core/bem_kernel.py:715-810: sigma (N,) or (B, N) -> V of the same shape."""
    if sigma.ndim not in (1, 2):
        raise ValueError("sigma must have shape (N,) or (B, N).")
    ker = kernel or DEFAULT_SINGLE_LAYER_KERNEL
    if not isinstance(ker, LaplaceSingleLayerKernel):
        raise ValueError("only the Laplace single-layer kernel is implemented natively")
    return dense_matvec(sigma, src_centroids, areas, self_integrals)


def _targets(targets, src_centroids, areas, sigma, want_v, want_e):
    nat.require_gpu()
    dt = sigma.dtype if sigma.dtype in (torch.float32, torch.float64) else torch.float64
    dev = sigma.device if sigma.is_cuda else torch.device("cuda", torch.cuda.current_device())
    X = torch.as_tensor(targets).to(dev, dt).reshape(-1, 3).contiguous()
    C = torch.as_tensor(src_centroids).to(dev, dt).reshape(-1, 3).contiguous()
    A = torch.as_tensor(areas).to(dev, dt).reshape(-1).contiguous()
    s = sigma.to(dev, dt).reshape(-1).contiguous()
    M, N = X.shape[0], s.shape[0]
    if C.shape[0] != N or A.shape[0] != N:
        raise ValueError("src_centroids, areas and sigma must have the same length")
    V = torch.empty(M, dtype=dt, device=dev) if want_v else None
    E = torch.empty((M, 3), dtype=dt, device=dev) if want_e else None
    fn = nat.lib().ed_targets_f64 if dt == torch.float64 else nat.lib().ed_targets_f32
    with torch.cuda.device(dev):
        nat.check(fn(nat.ptr(X), M, nat.ptr(C), nat.ptr(A), nat.ptr(s), N, nat.ptr(V), nat.ptr(E),
                     nat.stream_ptr(dev)), "ed_targets")
    out_dev = sigma.device
    return (None if V is None else V.to(out_dev)), (None if E is None else E.to(out_dev))


def bem_potential_targets(targets, src_centroids, areas, sigma, tile_size: int = 1024, *,
                          logger=None, use_keops: bool = False, kernel=None,
                          backend: str = "auto"):
    """Induced potential at arbitrary targets (core/bem_kernel.py:818-918):
    V_i = sum_j K_E sigma_j A_j rsqrt(max(|x_i - c_j|^2, 1e-24)) on the native
    target kernel; same backends and errors as the reference."""
    if backend not in {"auto", "torch_tiled", "keops"}:
        raise ValueError(f"Unknown backend for bem_potential_targets: {backend!r}")
    ker = kernel or DEFAULT_SINGLE_LAYER_KERNEL
    if not isinstance(ker, LaplaceSingleLayerKernel):
        raise ValueError("only the Laplace single-layer kernel is implemented natively")
    _ensure_tile_size(tile_size)
    del use_keops, logger
    return _targets(targets, src_centroids, areas, sigma, True, False)[0]


def bem_E_field_targets(targets, src_centroids, areas, sigma, tile_size: int = 1024, *,
                        logger=None, use_keops: bool = False, kernel=None,
                        backend: str = "auto"):
    """Induced E-field at arbitrary targets (core/bem_kernel.py:921-1008):
    E_i = sum_j K_E sigma_j A_j (x_i - c_j) rsqrt(max(r^2, 1e-24))^3, shape (M, 3)."""
    if backend not in {"auto", "torch_tiled", "keops"}:
        raise ValueError(f"Unknown backend for bem_E_field_targets: {backend!r}")
    ker = kernel or DEFAULT_SINGLE_LAYER_KERNEL
    if not isinstance(ker, LaplaceSingleLayerKernel):
        raise ValueError("only the Laplace single-layer kernel is implemented natively")
    _ensure_tile_size(tile_size)
    del use_keops, logger
    return _targets(targets, src_centroids, areas, sigma, False, True)[1]


# -- differentiable kernels (core/bem_kernel.py:157-335, :1016-1080) ------------
# Forward and backward both run on the native kernels.  With q = sigma * A and
# rho_ij = rsqrt(max(|x_i - c_j|^2, 1e-24)), upstream gradient g:
#   dense matvec  V_i = sum_{j != i} K_E q_j rho_ij + s_i q_i
#     dL/dq_j = sum_{i != j} K_E g_i rho_ij + s_j g_j   (the dense kernel on g, unit areas)
#     dL/dc_k = -g_k E_q(c_k) - q_k E_g(c_k)            (E_w(x) = sum K_E w (x - c) rho^3)
#     dL/ds_i = g_i q_i
#   target potential  V_i = sum_j K_E q_j rho_ij at x_i
#     dL/dq_j = sum_i K_E g_i rho_ij                     (targets and sources swapped)
#     dL/dx_i = -g_i E_q(x_i),  dL/dc_j = -q_j E_g(c_j) (g as charges at the targets)
#   target field  E_i = sum_j K_E q_j (x_i - c_j) rho_ij^3
#     dL/dq_j = -sum_k [field at c_j of charges G_.k at the targets]_k
# dL/dsigma = A dL/dq and dL/dA = sigma dL/dq.  The r^2 clamp is constant
# below 1e-24 (zero derivative there), as in the reference's torch cores.
def _unit(n, like):
    return torch.ones(n, dtype=like.dtype, device=like.device)


class _DenseMatvecFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigma, centroids, areas, self_integrals):
        ctx.save_for_backward(sigma, centroids, areas, self_integrals)
        ctx.has_self = self_integrals is not None
        return dense_matvec(sigma.detach(), centroids.detach(), areas.detach(),
                            None if self_integrals is None else self_integrals.detach())

    @staticmethod
    def backward(ctx, g):
        sigma, C, A, s = ctx.saved_tensors
        g = g.detach().contiguous()
        C, A, sig = C.detach(), A.detach(), sigma.detach()
        n = sig.shape[0]
        need_s, need_c, need_a, need_si = ctx.needs_input_grad[:4]
        d_sig = d_c = d_a = d_si = None
        q = sig * A
        if need_s or need_a:
            dq = dense_matvec(g, C, _unit(n, g))
            if ctx.has_self:
                dq = dq + s.detach() * g
            d_sig = dq * A if need_s else None
            d_a = dq * sig if need_a else None
        if need_c:
            Eq = _targets(C, C, A, sig, False, True)[1]
            Eg = _targets(C, C, _unit(n, g), g, False, True)[1]
            d_c = -(g[:, None] * Eq) - q[:, None] * Eg
        if need_si and ctx.has_self:
            d_si = g * q
        return d_sig, d_c, d_a, d_si


class _TargetPotentialFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, targets, centroids, areas, sigma):
        ctx.save_for_backward(targets, centroids, areas, sigma)
        return _targets(targets.detach(), centroids.detach(), areas.detach(), sigma.detach(),
                        True, False)[0]

    @staticmethod
    def backward(ctx, g):
        X, C, A, sig = (t.detach() for t in ctx.saved_tensors)
        g = g.detach().contiguous()
        need_x, need_c, need_a, need_s = ctx.needs_input_grad[:4]
        d_x = d_c = d_a = d_s = None
        q = sig * A
        if need_s or need_a or need_c:
            ones = _unit(X.shape[0], g)
            if need_s or need_a:
                dq = _targets(C, X, ones, g, True, False)[0]
                d_s = dq * A if need_s else None
                d_a = dq * sig if need_a else None
            if need_c:
                d_c = -q[:, None] * _targets(C, X, ones, g, False, True)[1]
        if need_x:
            d_x = -g[:, None] * _targets(X, C, A, sig, False, True)[1]
        return d_x, d_c, d_a, d_s


class _TargetFieldFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, targets, centroids, areas, sigma):
        ctx.save_for_backward(targets, centroids, areas, sigma)
        return _targets(targets.detach(), centroids.detach(), areas.detach(), sigma.detach(),
                        False, True)[1]

    @staticmethod
    def backward(ctx, G):
        X, C, A, sig = (t.detach() for t in ctx.saved_tensors)
        G = G.detach().contiguous()
        need_x, need_c, need_a, need_s = ctx.needs_input_grad[:4]
        if need_x or need_c:
            raise NotImplementedError("gradients of the target field with respect to target or "
                                      "source positions are not provided")
        d_a = d_s = None
        if need_s or need_a:
            ones = _unit(X.shape[0], G)
            dq = torch.zeros_like(sig)
            for k in range(3):
                dq = dq - _targets(C, X, ones, G[:, k].contiguous(), False, True)[1][:, k]
            d_s = dq * A if need_s else None
            d_a = dq * sig if need_a else None
        return None, None, d_a, d_s


def _as_dev(t, like):
    return torch.as_tensor(t).to(like.device, like.dtype)


def _bem_matvec_core_torch(centroids, areas, sigma, self_integrals=None, tile_size: int = 1024,
                           kernel=None):
    """Differentiable dense single-layer matvec (core/bem_kernel.py:157-230):
    same values as ``bem_matvec_gpu``, gradients to sigma, centroids, areas and
    self_integrals (native kernels both ways)."""
    ker = kernel or DEFAULT_SINGLE_LAYER_KERNEL
    if not isinstance(ker, LaplaceSingleLayerKernel):
        raise ValueError("only the Laplace single-layer kernel is implemented natively")
    _ensure_tile_size(tile_size)
    nat.require_gpu()
    if sigma.dim() != 1:
        raise ValueError("sigma must have shape (N,)")
    dev = sigma.device if sigma.is_cuda else torch.device("cuda", torch.cuda.current_device())
    s = sigma.to(dev)
    C = _as_dev(centroids, s).reshape(-1, 3)
    A = _as_dev(areas, s).reshape(-1)
    si = None if self_integrals is None else _as_dev(self_integrals, s).reshape(-1)
    return _DenseMatvecFn.apply(s, C, A, si).to(sigma.device)


def _bem_potential_targets_core_torch(targets, src_centroids, areas, sigma, tile_size: int = 1024,
                                      kernel=None):
    """Differentiable induced potential at targets (core/bem_kernel.py:233-276)."""
    ker = kernel or DEFAULT_SINGLE_LAYER_KERNEL
    if not isinstance(ker, LaplaceSingleLayerKernel):
        raise ValueError("only the Laplace single-layer kernel is implemented natively")
    nat.require_gpu()
    dev = sigma.device if sigma.is_cuda else torch.device("cuda", torch.cuda.current_device())
    s = sigma.to(dev).reshape(-1)
    return _TargetPotentialFn.apply(_as_dev(targets, s).reshape(-1, 3),
                                    _as_dev(src_centroids, s).reshape(-1, 3),
                                    _as_dev(areas, s).reshape(-1), s).to(sigma.device)


def _bem_E_field_targets_core_torch(targets, src_centroids, areas, sigma, tile_size: int = 1024,
                                    kernel=None):
    """Differentiable induced E-field at targets (core/bem_kernel.py:279-335);
    gradients to sigma and areas."""
    ker = kernel or DEFAULT_SINGLE_LAYER_KERNEL
    if not isinstance(ker, LaplaceSingleLayerKernel):
        raise ValueError("only the Laplace single-layer kernel is implemented natively")
    nat.require_gpu()
    dev = sigma.device if sigma.is_cuda else torch.device("cuda", torch.cuda.current_device())
    s = sigma.to(dev).reshape(-1)
    return _TargetFieldFn.apply(_as_dev(targets, s).reshape(-1, 3),
                                _as_dev(src_centroids, s).reshape(-1, 3),
                                _as_dev(areas, s).reshape(-1), s).to(sigma.device)


def bem_matvec_diff(sigma, src_centroids, areas, *, tile_size: int = 1024, self_integrals=None,
                    kernel=None):
    """core/bem_kernel.py:1016-1037."""
    return _bem_matvec_core_torch(src_centroids, areas, sigma, self_integrals, tile_size, kernel)


def bem_potential_targets_diff(targets, src_centroids, areas, sigma, tile_size: int = 1024,
                               kernel=None):
    """core/bem_kernel.py:1040-1058."""
    return _bem_potential_targets_core_torch(targets, src_centroids, areas, sigma, tile_size,
                                             kernel)


def bem_E_field_targets_diff(targets, src_centroids, areas, sigma, tile_size: int = 1024,
                             kernel=None):
    """core/bem_kernel.py:1061-1079."""
    return _bem_E_field_targets_core_torch(targets, src_centroids, areas, sigma, tile_size,
                                           kernel)
