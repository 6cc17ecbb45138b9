"""ctypes binding of libhtamd.so (the C ABI declared in include/htamd.h).

The library is built in-tree by ``python -m holotomocupy_amd.build``.  There is
no CPU fallback: if the library or a GPU is missing, every operator raises.
"""
import ctypes
import os

import torch

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libhtamd.so")

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_fp = ctypes.POINTER(ctypes.c_float)
c_dp = ctypes.POINTER(ctypes.c_double)

# name -> argtypes; every entry point returns int status (0 = OK)
SIGNATURES = {
    "htamd_prop_plan_create": [c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dp, ctypes.POINTER(c_vp)],
    "htamd_prop_apply": [c_vp, c_vp, c_vp, c_i64, c_i64, c_int, c_vp],
    "htamd_prop_plan_destroy": [c_vp],
    "htamd_prop_apply_strided": [c_vp, c_vp, c_i64, c_vp, c_i64, c_i64, c_i64, c_int, c_vp],
    "htamd_prop_pass": [c_vp, c_vp, c_i64, c_vp, c_i64, c_i64, c_i64, c_int, c_int, c_vp],
    "htamd_cas_probe_combo": [c_vp] + [c_vp] * 8 + [c_i64, c_i64, c_i64, ctypes.c_float, c_vp],
    "htamd_cas_expi": [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_vp],
    "htamd_cas_residual": [c_vp, c_vp, c_vp, c_i64, ctypes.c_float, c_vp],
    "htamd_cas_probe_grad": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_int, c_vp],
    "htamd_cas_reduce_workspace_size": [c_i64, ctypes.POINTER(c_i64)],
    "htamd_cas_reduce": [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp],
    "htamd_bh_plan_create": [c_vp, c_vp, ctypes.POINTER(c_vp)],
    "htamd_bh_plan_destroy": [c_vp],
    "htamd_bh_workspace_size": [c_vp, c_i64, ctypes.POINTER(c_i64)],
    "htamd_bh_set_stage_events": [c_vp, ctypes.POINTER(c_vp), c_int],
    "htamd_bh_gradient": [c_vp, c_i64] + [c_vp] * 8 + [c_int] + [c_vp] * 5 + [c_i64, c_vp],
    "htamd_bh_gradient_tab": [c_vp, c_i64, c_i64] + [c_vp] * 6 + [c_int, c_vp, c_vp, c_vp, c_i64, c_vp],
    "htamd_bh_hessian_workspace_size": [c_vp, c_i64, c_int, ctypes.POINTER(c_i64)],
    "htamd_bh_hessian": [c_vp, c_i64] + [c_vp] * 11 + [c_dbl, c_vp, c_int, c_vp, c_i64, c_vp],
    "htamd_bh_hessian_pair": [c_vp, c_i64] + [c_vp] * 11 + [c_dbl, c_vp, c_int, c_vp, c_i64, c_vp],
    "htamd_bspline_sback": [c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp],
    "htamd_bspline_coeffback": [c_vp, c_vp, c_vp, c_i64, c_dbl, c_vp],
    "htamd_axpby": [c_vp, c_vp, ctypes.c_float, ctypes.c_float, c_vp, c_i64, c_vp],
    "htamd_blas1_workspace_size": [ctypes.POINTER(c_i64)],
    "htamd_redot": [c_vp, c_vp, c_i64, c_vp, c_int, c_vp, c_i64, c_vp],
    "htamd_axpby_redot": [c_vp, c_vp, ctypes.c_float, ctypes.c_float, c_i64, c_vp, c_int, c_vp, c_i64, c_vp],
    "htamd_axpby_dc": [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp],
    "htamd_axpby_redot_dc": [c_vp, c_vp, c_vp, c_i64, c_vp, c_dbl, c_int, c_vp, c_i64, c_vp],
    "htamd_laplacian3d": [c_vp, c_i64, c_i64, c_i64, c_int, c_i64, c_i64, c_vp, ctypes.c_float, c_vp, c_vp],
    "htamd_tomo_plan_create": [c_i64, c_i64, c_vp, c_vp, ctypes.POINTER(c_vp)],
    "htamd_tomo_plan_destroy": [c_vp],
    "htamd_tomo_workspace_size": [c_vp, c_i64, ctypes.POINTER(c_i64)],
    "htamd_tomo_R": [c_vp, c_vp, c_int, c_vp, c_i64, c_vp, c_i64, c_vp],
    "htamd_tomo_RT": [c_vp, c_vp, c_int, c_vp, c_i64, c_vp, c_i64, c_vp],
    "htamd_tomo_filter": [c_vp, c_vp, c_int, c_int, c_vp, c_i64, c_vp],
    "htamd_tomo_fbp": [c_vp, c_vp, c_int, c_int, c_vp, c_i64, c_vp, c_i64, c_vp],
    "htamd_shift_plan_create": [c_i64, c_i64, c_i64, c_i64, c_int, ctypes.POINTER(c_vp)],
    "htamd_shift_plan_destroy": [c_vp],
    "htamd_bspline_prefilter": [c_vp, c_vp, c_vp, c_i64, c_vp],
    "htamd_bspline_interp": [c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp],
    "htamd_bspline_interp_adj": [c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp],
    "htamd_bspline_d1": [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp],
    "htamd_bspline_d2": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64,
                         c_vp, c_i64, c_vp],
    "htamd_bspline_d1_adj_workspace_size": [c_vp, c_i64, ctypes.POINTER(c_i64)],
    "htamd_bspline_d1_adj": [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64,
                             c_i64, c_vp],
    "htamd_bspline_d1_adj_noalloc_max_bytes": [ctypes.POINTER(c_i64)],
    "htamd_bspline_d1_adj_noalloc": [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp],
}

_lib = None


class HtamdError(RuntimeError):
    """Raised when a libhtamd entry point returns a non-zero status."""


def load():
    """Load libhtamd.so once and attach the ctypes signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HtamdError(
            f"{LIB_PATH} is missing: build it with `python -m holotomocupy_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.htamd_last_error.restype = ctypes.c_char_p
    lib.htamd_last_error.argtypes = []
    lib.htamd_abi_version.restype = c_int
    lib.htamd_abi_version.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = c_int
    _lib = lib
    return lib


def check(status, what):
    if status != 0:
        msg = load().htamd_last_error().decode(errors="replace")
        raise HtamdError(f"{what} failed ({status}): {msg}")


def call(name, *args):
    check(getattr(load(), name)(*args), name)


def require_gpu():
    if not torch.cuda.is_available():
        raise HtamdError("holotomocupy_amd needs a ROCm GPU; none is visible (no CPU fallback)")


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return c_vp(s.cuda_stream)


def ptr(t):
    return c_vp(t.data_ptr())


def as_device(x, dtype, device):
    """Return a contiguous tensor of `dtype` on `device` (copies only if needed)."""
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x)
    if x.device != device or x.dtype != dtype:
        x = x.to(device=device, dtype=dtype)
    return x.contiguous()
