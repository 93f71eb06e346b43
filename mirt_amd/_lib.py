"""ctypes binding of ``libfdkhip.so`` (the C ABI in ``include/fdkhip.h``).

The library is the only compute path: if it is missing or fails to load this
module raises, it never falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

from . import build as _build

FDKHIP_OK = 0
FDKHIP_EARG = -1
FDKHIP_EGEOM = -2
FDKHIP_ETHREADS = -3
FDKHIP_EHIP = -4
FDKHIP_ENOMEM = -5
FDKHIP_EUNSUPPORTED = -6

F32, F16, BF16 = 0, 1, 2
MATH_EXACT, MATH_FAST = 0, 1

FMT_NAMES = {"f32": F32, "fp32": F32, "float32": F32, "f16": F16, "fp16": F16,
             "float16": F16, "bf16": BF16, "bfloat16": BF16}


class FdkhipError(RuntimeError):
    """Failure reported by libfdkhip (code + message, like the reference's Fail)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"fdkhip error {code}: {msg}")
        self.code = code
        self.msg = msg


class Geom(C.Structure):
    """``fdkhip_geom`` (include/fdkhip.h)."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("offset_x", C.c_double), ("offset_y", C.c_double), ("offset_z", C.c_double),
        ("dso", C.c_double), ("dsd", C.c_double), ("dfs", C.c_double),
        ("ns", C.c_int32), ("nt", C.c_int32),
        ("ds", C.c_double), ("dt", C.c_double),
        ("offset_s", C.c_double), ("offset_t", C.c_double),
    ]


# exported symbol -> (restype, argtypes)
_P = C.c_void_p
_SIGS = {
    "fdkhip_last_error": (C.c_char_p, []),
    "fdkhip_version": (C.c_char_p, []),
    "fdkhip_fdk_ts_back": (C.c_int, [_P, _P, _P, _P, C.c_double, C.c_double, C.c_double,
                                     _P, _P, _P, C.c_int32, C.c_int32, C.c_int32, _P,
                                     C.c_int32, _P]),
    "fdkhip_release_host_cache": (C.c_int, []),
    "fdkhip_fdk_st_back": (C.c_int, [_P, _P, _P, _P, C.c_double, C.c_double, C.c_double,
                                     _P, _P, _P, C.c_int32, C.c_int32, C.c_int32, _P,
                                     C.c_int32, _P]),
    "fdkhip_cbct_back": (C.c_int, [C.POINTER(Geom), _P, C.c_int32, _P, C.c_int32, _P,
                                   C.c_int32, C.c_double, _P]),
    "fdkhip_fdk_ts_back_helix": (C.c_int, [_P, _P, _P, _P, C.c_double, C.c_double,
                                           C.c_double, _P, _P, _P, C.c_int32, C.c_int32,
                                           C.c_int32, _P, _P, C.c_int32, C.c_int32,
                                           C.c_int32, C.c_double, C.c_float, C.c_int32,
                                           _P]),
    "fdkhip_fdk_ts_back_helix_direct": (C.c_int, [C.POINTER(Geom), _P, C.c_int32, _P, _P,
                                                  C.c_int32, C.c_int32, C.c_int32,
                                                  C.c_double, C.c_float, _P, _P, _P,
                                                  C.POINTER(C.c_float)]),
    "fdkhip_cbct_back_mat": (C.c_int, [C.POINTER(Geom), _P, C.c_int32, C.c_int32, _P, _P,
                                       C.c_int32, C.c_double, _P, C.c_double, _P]),
    "fdkhip_cbct_back_helix_mat": (C.c_int, [C.POINTER(Geom), _P, C.c_int32, C.c_int32, _P,
                                             _P, C.c_int32, C.c_double, _P, _P, C.c_int32,
                                             C.c_double, _P, _P]),
    "fdkhip_cbct_back_mat_direct": (C.c_int, [C.POINTER(Geom), _P, C.c_int32, C.c_int32, _P,
                                              _P, C.c_int32, C.c_double, _P, C.c_double, _P,
                                              _P, C.POINTER(C.c_float)]),
    "fdkhip_plan_create": (C.c_int, [C.POINTER(_P), C.POINTER(Geom), C.c_int32, _P, _P,
                                     C.c_int32, C.c_int32]),
    "fdkhip_plan_create_helix": (C.c_int, [C.POINTER(_P), C.POINTER(Geom), C.c_int32, _P, _P,
                                           C.c_int32, C.c_float, _P, C.c_int32, C.c_int32]),
    "fdkhip_plan_execute": (C.c_int, [_P, _P, _P, _P, _P]),
    "fdkhip_plan_execute_direct": (C.c_int, [_P, _P, _P, _P, _P]),
    "fdkhip_plan_config": (C.c_int, [_P, _P]),
    "fdkhip_plan_set_config": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "fdkhip_plan_set_occupancy": (C.c_int, [_P, C.c_int32]),
    "fdkhip_plan_occupancy": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "fdkhip_plan_destroy": (C.c_int, [_P]),
    "fdkhip_plan_count_updates": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "fdkhip_quantize_views": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                        _P, _P, _P]),
    "fdkhip_plan_profile": (C.c_int, [_P, _P, _P, _P, _P, C.c_int32, _P]),
    "fdkhip_plan_autotune": (C.c_int, [_P, _P, _P, _P, _P, C.c_int32, C.POINTER(C.c_float)]),
    "fdkhip_batch_create": (C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(Geom), _P, _P, _P,
                                      C.c_int32, C.c_int32]),
    "fdkhip_batch_execute": (C.c_int, [_P, _P, _P, _P, _P]),
    "fdkhip_batch_execute_direct": (C.c_int, [_P, _P, _P, _P, _P]),
    "fdkhip_batch_config": (C.c_int, [_P, _P]),
    "fdkhip_batch_set_config": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32]),
    "fdkhip_batch_set_per_scan": (C.c_int, [_P]),
    "fdkhip_batch_set_zper": (C.c_int, [_P, _P]),
    "fdkhip_fdk_filter_views": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _P, _P,
                                          C.c_int32, _P, _P]),
    "fdkhip_batch_scan_zper": (C.c_int, [_P, _P]),
    "fdkhip_batch_autotune": (C.c_int, [_P, _P, _P, _P, _P, C.c_int32,
                                        C.POINTER(C.c_float)]),
    "fdkhip_batch_profile": (C.c_int, [_P, _P, _P, _P, _P, C.c_int32, _P]),
    "fdkhip_batch_count_updates": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "fdkhip_batch_destroy": (C.c_int, [_P]),
    "fdkhip_ellipsoid_proj": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, C.c_int32,
                                        C.c_double, C.c_double, C.c_double, _P, C.c_int32,
                                        C.c_int32, _P, _P]),
    "fdkhip_back_noalloc": (C.c_int, [C.POINTER(Geom), C.c_int32, _P, _P, C.c_int32, _P, _P,
                                      _P, _P]),
}

_lib = None


def lib_path() -> str:
    """In-tree libfdkhip.so, or the library named by $MIRT_AMD_LIB."""
    return os.environ.get("MIRT_AMD_LIB") or _build.LIB


def load() -> C.CDLL:
    """Load libfdkhip.so (built in-tree by mirt_amd.build); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -m mirt_amd.build` "
            "(the HIP library is the only compute path)")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def check(rc: int) -> None:
    if rc != FDKHIP_OK:
        msg = load().fdkhip_last_error().decode(errors="replace")
        raise FdkhipError(rc, msg)


def version() -> str:
    return load().fdkhip_version().decode()
