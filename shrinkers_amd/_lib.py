"""ctypes binding of libshrinkers_hip.so (the C ABI in include/).

Loading fails loudly: there is no CPU implementation behind this package,
so a missing or unloadable library raises ``ImportError``/``OSError`` at the
first call that needs it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshrinkers_hip.so")

# shk_status codes (include/shrinkers_stieltjes.h)
SHK_OK = 0
SHK_ERR_EMPTY = 1
SHK_ERR_NONFINITE = 2
SHK_ERR_BAD_ETA = 3
SHK_ERR_BAD_METHOD = 4
SHK_ERR_UNSUPPORTED = 5
SHK_ERR_BAD_ARG = 6
SHK_ERR_HIP = 7
SHK_ERR_NO_DEVICE = 8

# shk_precision
PREC_F64 = 0
PREC_F32C = 1
PREC_F32 = 2

# Exported entry points and their C signatures.
_d = ctypes.c_double
_f = ctypes.c_float
_i = ctypes.c_int
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_szp = ctypes.POINTER(ctypes.c_size_t)
_ip = ctypes.POINTER(ctypes.c_int)

SIGNATURES = {
    "shk_last_error": (ctypes.c_char_p, []),
    "shk_release_cached_memory": (_i, []),
    "shk_version": (ctypes.c_char_p, []),
    "shk_method_from_name": (_i, [ctypes.c_char_p, _ip]),
    "shk_method_name": (ctypes.c_char_p, [_i]),
    "shk_default_eta": (_d, [_sz]),
    "shk_resolve_auto_method": (_i, [_i, _i, _sz]),
    "shk_symmetric_range": (_i, [_i, _szp, _szp]),
    "shk_set_symmetric_scratch_limit": (_i, [_sz]),
    "shk_compute_all_stieltjes": (_i, [_dp, _sz, _d, _i, _d, _i, _i, _dp, _dp]),
    "shk_compute_all_stieltjes_f32": (_i, [_fp, _sz, _f, _i, _d, _fp, _fp]),
    "shk_compute_stieltjes_at_points": (_i, [_dp, _sz, _dp, _sz, _d, _i, _d, _i, _dp, _dp]),
    "shk_compute_all_stieltjes_with_deriv": (_i, [_dp, _sz, _d, _dp, _dp, _dp, _dp]),
    "shk_compute_stieltjes_scaled_ray": (_i, [_dp, _sz, _d, _dp, _dp]),
    "shk_all_points_dev": (_i, [_vp, _sz, _d, _i, _d, _i, _i, _vp, _vp, _vp]),
    "shk_all_points_noalloc_dev": (_i, [_vp, _sz, _d, _i, _i, _vp, _vp, _vp]),
    "shk_at_points_dev": (_i, [_vp, _sz, _vp, _sz, _d, _i, _d, _vp, _vp, _vp]),
    "shk_chebcode_workspace_bytes": (_i, [_sz, _i, _szp]),
    "shk_chebcode_all_points_ws_dev": (_i, [_vp, _sz, _d, _i, _i, _vp, _sz, _vp, _vp, _vp]),
    "shk_chebcode_at_points_ws_dev": (_i, [_vp, _sz, _vp, _sz, _d, _i, _vp, _sz, _vp, _vp,
                                           _vp]),
    "shk_chebcode_workspace_status": (_i, [_vp, _vp, _ip]),
    "shk_all_points_many_dev": (_i, [_vp, _i64p, _sz, _dp, _i, _i, _i, _vp, _vp, _i,
                                     ctypes.POINTER(ctypes.c_float), _vp]),
    "shk_chebcode_build": (_i, [_dp, _sz, _d, _i, _sz, _i, ctypes.POINTER(_vp)]),
    "shk_chebcode_build_preset": (_i, [_dp, _sz, _i, ctypes.POINTER(_vp)]),
    "shk_chebcode_info": (_i, [_vp, _szp, _szp]),
    "shk_chebcode_export": (_i, [_vp, _dp, _dp, _i64p, _i64p, _i32p, _dp, _dp]),
    "shk_chebcode_evaluate": (_i, [_vp, _d, _dp, _dp]),
    "shk_chebcode_evaluate_points": (_i, [_vp, _dp, _sz, _d, _dp, _dp]),
    "shk_chebcode_evaluate_many": (_i, [_vp, _dp, _sz, _dp, _dp]),
    "shk_chebcode_destroy": (_i, [_vp]),
    "shk_chebcode_build_dev": (_i, [_vp, _sz, _d, _i, _sz, _i, _vp, ctypes.POINTER(_vp)]),
    "shk_chebcode_evaluate_dev": (_i, [_vp, _d, _d, _vp, _vp, _vp]),
    "shk_chebcode_evaluate_many_dev": (_i, [_vp, _dp, _sz, _vp, _vp, _vp]),
    "shk_chebcode_evaluate_local_dev": (_i, [_vp, _d, _d, _vp, _vp, _vp]),
    "shk_chebcode_count_work": (_i, [_vp, _vp, _sz, _d, _i64p, _i64p, _vp]),
    "shk_batch_create": (_i, [_vp, _i64p, _sz, _dp, _i, ctypes.POINTER(_vp)]),
    "shk_batch_run": (_i, [_vp, _vp, _vp, _vp]),
    "shk_batch_tune": (_i, [_vp, _vp, _vp, _vp, _ip, _dp]),
    "shk_batch_set_config": (_i, [_vp, _i]),
    "shk_batch_spectrum_times": (_i, [_vp, _dp, _dp]),
    "shk_batch_destroy": (_i, [_vp]),
    "shk_exact_at_points_dev": (_i, [_vp, _sz, _vp, _sz, _d, _i, _vp, _vp, _vp]),
    "shk_exact_adjoint_dev": (_i, [_vp, _sz, _vp, _sz, _d, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "shk_exact_adjoint_all_points_dev": (_i, [_vp, _sz, _d, _vp, _vp, _i, _vp, _vp, _vp]),
    "shk_el_karoui_fit_dev": (_i, [_vp, _vp, _vp, _sz, _d, _d, _vp, _vp, _sz, _vp, _vp, _vp,
                                   _vp, _vp, _vp]),
    "shk_el_karoui_fit_step_dev": (_i, [_vp, _sz, _vp, _sz, _d, _d, _vp, _vp, _sz, _i, _vp,
                                        _vp, _vp, _vp, _vp]),
    "shk_profile_enable": (_i, [_i]),
    "shk_profile_read": (_i, [_dp, _i64p]),
    "shk_profile_reset": (_i, []),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the library once; raise if it is absent (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m shrinkers_amd.build` "
            "(hipcc, gfx950). shrinkers_amd has no CPU implementation.")
    # PyTorch-ROCm ships its own HIP runtime, which its libraries load under the
    # name libamdhip64.so while this library asks for libamdhip64.so.7.  If this
    # library were loaded first, a later torch import would bring in a second
    # HIP/HSA runtime and neither would share device memory with the other.
    # Importing torch first makes its runtime the one this library binds to.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class ShkError(RuntimeError):
    """A non-OK shk_status from the library."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[shk status {status}] {message}")
        self.status = status
        self.message = message


def last_error() -> str:
    msg = load().shk_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    if status != SHK_OK:
        raise ShkError(status, last_error())


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)
