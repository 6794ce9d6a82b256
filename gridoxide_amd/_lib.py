"""ctypes binding of ``libgridox_hip.so`` (the C ABI declared in ``include/gridox_hip.h``).

Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GRIDOX_HIP_LIB names another build of the same library (the host-sanitizer build)
LIB_PATH = os.environ.get("GRIDOX_HIP_LIB") or os.path.join(HERE, "libgridox_hip.so")

GX_OK = 0
GX_ERR_ARG = -1
GX_ERR_SINGULAR = -2
GX_ERR_DEVICE = -3
GX_ERR_STATE = -4
GX_ERR_CAPACITY = -5
GX_ERR_UNSUPPORTED = -6
GX_ERR_BUDGET = -7
GX_CONVERGED, GX_MAX_ITER, GX_SINGULAR = 0, 1, 2
GX_MODE_FP64, GX_MODE_MIXED = 0, 1
GX_DEVICE_PTRS = 1
GX_INIT_LINEAR = 2
GX_INIT_WARM = 4
GX_KEEP_WARM = 8
KERNEL_CLASSES = ["mismatch", "fill", "lu", "refine", "solve", "update"]

_i32p = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_longlong)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); the full exported surface of include/gridox_hip.h
SIGNATURES = {
    "gx_version": (ctypes.c_char_p, []),
    "gx_last_error": (ctypes.c_char_p, []),
    "gx_ybus_build": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                     ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, ctypes.c_int, _i32p, _i32p,
                                     _f64p, _f64p]),
    "gx_topo_create": (_vp, [ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, _i32p]),
    "gx_topo_destroy": (None, [_vp]),
    "gx_topo_sizes": (ctypes.c_int, [_vp, _i64p]),
    "gx_topo_jacobian_pattern": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "gx_topo_factor_at": (ctypes.c_int, [_vp, _f64p, _f64p, _f64p, _f64p]),
    "gx_topo_set_pivot_order": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "gx_topo_ordering": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _i32p]),
    "gx_topo_lu_pattern": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p]),
    "gx_batch_solve": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gx_batch_fallbacks": (ctypes.c_int, [_vp]),
    "gx_topo_islands": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "gx_batch_island_status": (ctypes.c_int, [_vp, _i32p]),
    "gx_topo_profile": (ctypes.c_int, [_vp, ctypes.c_int, _f64p, _i32p]),
    "gx_topo_set_launch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "gx_topo_set_refactor_groups": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gx_topo_set_zip": (ctypes.c_int, [_vp, _f64p]),
    "gx_topo_back_levels": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_longlong, _i32p]),
    "gx_topo_forward_levels": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_longlong, _i32p]),
    "gx_topo_refactor_parts": (ctypes.c_int, [_vp, ctypes.c_int, _i64p, _i32p, _i32p, _i32p, _i32p]),
    "gx_topo_set_branches": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                            ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]),
    "gx_topo_branch_is_bridge": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gx_engine_create": (_vp, [ctypes.c_int, ctypes.c_int, ctypes.c_longlong, _vp, _i32p]),
    "gx_engine_destroy": (None, [_vp]),
    "gx_engine_request_bytes": (ctypes.c_longlong, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int]),
    "gx_engine_bytes": (ctypes.c_longlong, [_vp]),
    "gx_engine_submit": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "gx_engine_step": (ctypes.c_int, [_vp]),
    "gx_engine_run": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gx_engine_collect": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "gx_engine_fallbacks": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gx_batch_solve_outages": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                              _vp]),
    "gx_sys_create": (_vp, [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p]),
    "gx_sys_factor_and_solve_values": (ctypes.c_int, [_vp, _f64p, _f64p, _f64p]),
    "gx_sys_destroy": (None, [_vp]),
    "gx_sys_factor_and_solve_values_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "gx_sys_ordering": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _i32p, _i32p]),
}

_LIB = None


class GxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gridox_hip error {code}: {msg}")
        self.code = code


def lib():
    """The loaded library; raises when it is missing (build it with ``gridoxide_amd.build``)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build the HIP library first (python -m gridoxide_amd.build)")
        # One HIP runtime per process: PyTorch ships its own libamdhip64 (same soname, other file
        # name).  Loading torch first lets this library bind to that already-loaded runtime
        # instead of pulling in a second one, which would leave torch without a usable device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int) -> int:
    if rc < 0:
        raise GxError(rc, lib().gx_last_error().decode())
    return rc


def ptr(a, t=_f64p):
    return a.ctypes.data_as(t)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
