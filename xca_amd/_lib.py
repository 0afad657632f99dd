"""ctypes binding of the native library libxca_amd.so (declared in include/xca_amd.h)."""
import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libxca_amd.so"
_lib = None

d_p = C.POINTER(C.c_double)
i_p = C.POINTER(C.c_int)
i32_p = C.POINTER(C.c_int32)
i64_p = C.POINTER(C.c_int64)


class XcaProblem(C.Structure):
    _fields_ = [
        ("beta", C.c_double), ("r", C.c_int), ("it", d_p), ("rf", d_p),
        ("p", C.c_int), ("poles", d_p), ("coeffs", d_p),
        ("n", C.c_int), ("nblocks", C.c_int), ("dims", i_p),
        ("c_conn", i_p), ("cdag_conn", i_p), ("c_mats", d_p), ("cdag_mats", d_p),
        ("coeffs_shape", i64_p),
    ]


EXPORTS = {
    "xca_last_error": (C.c_char_p, []),
    "xca_dlr_build": (C.c_int, [C.c_double, C.c_double, C.c_int, d_p, d_p]),
    "xca_dlr_kernel": (C.c_double, [C.c_double, C.c_double]),
    "xca_ctx_create": (C.c_void_p, [C.POINTER(XcaProblem), C.c_int]),
    "xca_ctx_destroy": (None, [C.c_void_p]),
    "xca_num_backbones": (C.c_int64, [C.c_void_p, C.c_int]),
    "xca_num_sym_sets": (C.c_int, [C.c_void_p]),
    "xca_num_cached_plans": (C.c_int64, [C.c_void_p]),
    "xca_sigma": (C.c_int, [C.c_void_p, d_p, C.c_int, i_p, i32_p, C.c_int64, d_p]),
    "xca_set_G": (C.c_int, [C.c_void_p, d_p]),
    "xca_sigma_zero": (C.c_int, [C.c_void_p]),
    "xca_sigma_accumulate": (C.c_int, [C.c_void_p, C.c_int, i_p, C.c_double]),
    "xca_get_sigma": (C.c_int, [C.c_void_p, d_p]),
    "xca_synchronize": (C.c_int, [C.c_void_p]),
    "xca_sigma_order": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "xca_last_stats": (C.c_int, [C.c_void_p, i64_p, d_p, d_p]),
    "xca_server_create": (C.c_void_p, [C.c_int, C.c_int64, C.c_int, C.c_int]),
    "xca_server_destroy": (None, [C.c_void_p]),
    "xca_server_submit": (C.c_int64, [C.c_void_p, C.POINTER(XcaProblem), d_p, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double]),
    "xca_server_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "xca_server_poll": (C.c_int, [C.c_void_p, i64_p, C.c_int]),
    "xca_server_result": (C.c_int, [C.c_void_p, C.c_int64, d_p, d_p, C.POINTER(C.c_int), d_p, C.POINTER(C.c_int), d_p]),
    "xca_server_request_bytes": (C.c_int64, [C.c_void_p, C.c_int64]),
    "xca_server_request_device_bytes": (C.c_int64, [C.c_void_p, C.c_int64]),
    "xca_server_shared_bytes": (C.c_int64, [C.c_void_p]),
    "xca_server_stats": (C.c_int, [C.c_void_p, i64_p, i64_p, i64_p]),
    "xca_set_tile_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "xca_set_launch_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "xca_last_tile": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "xca_dyson_init": (C.c_int, [C.c_void_p, d_p]),
    "xca_dyson_solve": (C.c_int, [C.c_void_p, d_p, C.c_double, d_p]),
    "xca_dyson_solve_noalloc": (C.c_int, [C.c_void_p, d_p, d_p, C.c_double, d_p]),
    "xca_ppsc_solve": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, d_p, d_p, C.POINTER(C.c_int), d_p]),
    "xca_ppsc_history": (C.c_int, [C.c_void_p, d_p, C.c_int]),
    "xca_num_correlator_backbones": (C.c_int64, [C.c_void_p, C.c_int]),
    "xca_spgf": (C.c_int, [C.c_void_p, d_p, C.c_int, i_p, i32_p, C.c_int64, d_p]),
    "xca_correlator": (C.c_int, [C.c_void_p, d_p, C.c_int, i_p, i32_p, C.c_int64, C.c_int, i_p, d_p, C.c_int, i_p, d_p, d_p]),
}


class XcaError(RuntimeError):
    pass


def lib():
    """Load libxca_amd.so. Raises if it has not been built: there is no fallback path."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise XcaError(f"native library {_LIB_PATH} is missing; build it with `python -m xca_amd.build`")
        L = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in EXPORTS.items():
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().xca_last_error().decode()


def check(status):
    if status < 0:
        raise XcaError(last_error())
    return status


def dptr(a):
    return a.ctypes.data_as(d_p)


def iptr(a):
    return a.ctypes.data_as(i_p)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
