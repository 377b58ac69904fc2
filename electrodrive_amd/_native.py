"""ctypes binding of ``libednative.so`` (the C ABI declared in include/ednative.h).

The product path fails loudly when the library is missing: there is no CPU
fallback for any compute entry point.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libednative.so")

ED_OK = 0
ED_ERR_ARG = 1
ED_ERR_DEVICE = 2
ED_ERR_ALLOC = 3
ED_ERR_LAUNCH = 4
ED_ERR_NONFINITE = 5
ED_ERR_INTERNAL = 6

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_i64p = ctypes.POINTER(ctypes.c_int64)


class PlanOpts(ctypes.Structure):
    _fields_ = [("leaf_size", c_i32), ("order", c_i32), ("ws", c_i32), ("max_level", c_i32),
                ("m2l_precision", c_i32), ("reserved", c_i32 * 3)]


class GmresOpts(ctypes.Structure):
    _fields_ = [("restart", c_i32), ("maxiter", c_i32), ("tol", c_dbl), ("precond", c_i32),
                ("refine", c_i32), ("max_refine", c_i32), ("inner_tol", c_dbl),
                ("use_graph", c_i32), ("check_every", c_i32), ("residual_update", c_i32)]


class GmresInfo(ctypes.Structure):
    _fields_ = [("code", c_i32), ("iters", c_i32), ("cycles", c_i32), ("refine_steps", c_i32),
                ("first_iter_below_tol", c_i32), ("resid", c_dbl), ("resid0", c_dbl),
                ("rel_resid", c_dbl), ("tol_abs", c_dbl), ("n_accurate", c_i32),
                ("n_updated", c_i32), ("resid_bound", c_dbl), ("update_err", c_dbl)]


_SIGS = {
    "ed_last_error": (ctypes.c_char_p, []),
    "ed_abi_version": (c_i32, []),
    "ed_plan_opts_default": (None, [ctypes.POINTER(PlanOpts)]),
    "ed_plan_create": (c_i32, [c_vp, c_vp, c_i64, ctypes.POINTER(PlanOpts), ctypes.POINTER(c_vp)]),
    "ed_plan_destroy": (c_i32, [c_vp]),
    "ed_plan_create_batched": (c_i32, [c_vp, c_vp, c_i64, c_vp, c_i64, ctypes.POINTER(PlanOpts),
                                       ctypes.POINTER(c_vp)]),
    "ed_plan_geometries": (c_i32, [c_vp, c_i64p, c_vp, c_i64]),
    "ed_plan_stats": (c_i32, [c_vp, c_vp, c_i32]),
    "ed_plan_permutation": (c_i32, [c_vp, c_vp]),
    "ed_fmm_matvec_f32": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "ed_fmm_matvec_f64": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "ed_fmm_matvec_multi_f64": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i32, c_vp]),
    "ed_plan_attach_near": (c_i32, [c_vp, c_vp, c_vp, c_i64]),
    "ed_fmm_profile": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_i32, c_vp]),
    "ed_fmm_profile_multi": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_vp]),
    "ed_dense_matvec_f64": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp]),
    "ed_dense_matvec_f32": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp]),
    "ed_dense_matvec_f64_noalloc": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp]),
    "ed_self_integrals_f64": (c_i32, [c_vp, c_vp, c_i64, c_i32, c_i32, c_vp, c_vp]),
    "ed_near_pairs_f64": (c_i32, [c_vp, c_vp, c_i64, c_dbl, c_vp, c_i64, c_i64p, c_vp]),
    "ed_near_weights_f64": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i32, c_i32, c_vp, c_vp]),
    "ed_near_apply_f64": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64,
                                  c_i32, c_vp, c_vp]),
    "ed_morton_keys_f64": (c_i32, [c_vp, c_i64, c_i32, c_vp, c_vp, c_vp, c_vp]),
    "ed_morton_keys_f32": (c_i32, [c_vp, c_i64, c_i32, c_vp, c_vp, c_vp, c_vp]),
    "ed_gmres_opts_default": (None, [ctypes.POINTER(GmresOpts)]),
    "ed_gmres_solve": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(GmresOpts),
                               ctypes.POINTER(GmresInfo), c_vp]),
    "ed_gmres_solve_multi": (c_i32, [c_vp, c_vp, c_vp, c_i32, c_vp, c_vp,
                                     ctypes.POINTER(GmresOpts), ctypes.POINTER(GmresInfo), c_vp]),
    "ed_gmres_block_resid": (c_i32, [c_vp, c_vp, c_vp, c_i32, ctypes.POINTER(c_i32),
                                     ctypes.POINTER(c_i32)]),
    "ed_gmres_trace": (c_i32, [c_vp, c_vp, c_i32, ctypes.POINTER(c_i32)]),
    "ed_plan_set_coarse": (c_i32, [c_vp, c_vp, c_i32, c_vp, c_vp]),
    "ed_plan_coarse_solve": (c_i32, [c_vp, c_vp, c_vp, c_vp]),
    "ed_plan_build_sai": (c_i32, [c_vp, c_vp, ctypes.c_double, ctypes.c_double, c_vp]),
    "ed_plan_attach_near_device": (c_i32, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "ed_targets_f64": (c_i32, [c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp]),
    "ed_targets_f32": (c_i32, [c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp]),
    "ed_near_targets_f64": (c_i32, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, ctypes.c_double,
                                    c_i32, c_i32, c_vp, c_vp]),
    "ed_plan_sai_apply": (c_i32, [c_vp, c_vp, c_vp, c_vp]),
    "ed_vec_dot_f64": (c_i32, [c_vp, c_vp, c_i64, c_vp, ctypes.POINTER(c_dbl), c_vp]),
    "ed_vec_axpy_f64": (c_i32, [c_dbl, c_vp, c_vp, c_i64, c_vp]),
    "ed_vec_scale_f64": (c_i32, [c_dbl, c_vp, c_vp, c_i64, c_vp]),
    "ed_vec_mul_f64": (c_i32, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "ed_gemm_f32": (c_i32, [c_i32, c_i32, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp,
                            c_i64, c_vp, c_i32, c_vp, ctypes.c_float, c_vp]),
    "ed_act_backward_f32": (c_i32, [c_vp, c_vp, c_vp, c_i64, c_i32, c_vp]),
    "ed_colsum_f32": (c_i32, [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp]),
    "ed_adam_f32": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i64, ctypes.c_float, ctypes.c_float,
                            ctypes.c_float, ctypes.c_float, c_i32, ctypes.c_float, c_vp]),
    "ed_wmse_f32": (c_i32, [c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, ctypes.POINTER(ctypes.c_float),
                            c_vp]),
    "ed_panel_features": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "ed_panel_feature_count": (c_i32, []),
    "ed_proj_loss_f32": (c_i32, [c_vp, c_i32, c_vp, c_i32, c_vp, c_i64, c_i32, ctypes.c_float,
                                 c_vp, c_vp, c_vp]),
    "ed_host_tree_build": (c_i32, [c_vp, c_i64, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp,
                                   c_vp, c_vp, c_vp, c_vp]),
    "ed_host_operator": (c_i32, [c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_vp]),
    "ed_host_morton_keys_f64": (c_i32, [c_vp, c_i64, c_i32, c_vp, c_vp, c_vp]),
    "ed_ref_tree_create": (c_i32, [c_vp, c_i32, c_i64, c_i32, c_i32, ctypes.POINTER(c_vp)]),
    "ed_ref_tree_sizes": (c_i32, [c_vp, c_vp]),
    "ed_ref_tree_export": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "ed_ref_tree_destroy": (c_i32, [c_vp]),
    "ed_ref_lists_create": (c_i32, [c_vp, c_vp, c_dbl, c_i64, c_i32, ctypes.POINTER(c_vp)]),
    "ed_ref_lists_sizes": (c_i32, [c_vp, c_vp]),
    "ed_ref_lists_export": (c_i32, [c_vp, c_vp, c_vp]),
    "ed_ref_lists_destroy": (c_i32, [c_vp]),
}

_lib = None
_lib_lock = threading.Lock()


def lib():
    """Load the native library (once, thread-safe).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"native library {LIB_PATH} is missing; build it with "
                "`python -m electrodrive_amd.build` (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS.keys())


def check(rc, what=""):
    if rc == ED_OK:
        return
    msg = lib().ed_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == ED_ERR_ARG:
        raise ValueError(text)
    raise RuntimeError(f"{text} (status {rc})")


def ptr(t):
    """Raw data pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_gpu_seen = False


def stream_ptr(device=None):
    """Current HIP stream of torch on ``device`` (torch's stream is plumbing).
    Reads the raw handle without building a torch Stream object where torch
    offers that (about 0.5 instead of 3 us per call)."""
    global _gpu_seen
    if not _gpu_seen:
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device available")
        _gpu_seen = True
    if _raw_stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    if isinstance(device, torch.device) and device.index is not None:
        idx = device.index
    elif isinstance(device, int):
        idx = device
    else:
        idx = torch.cuda.current_device() if device is None else torch.device(device).index
        if idx is None:
            idx = torch.cuda.current_device()
    return ctypes.c_void_p(_raw_stream(idx))


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("electrodrive_amd compute paths require a HIP device "
                           "(no CPU fallback)")
