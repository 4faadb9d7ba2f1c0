"""Loader for the in-tree HIP library libqotential_hip.so.

The library is the product: there is no Python or CPU fallback for the close
evaluation.  If the shared object is missing this raises immediately.
"""
import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libqotential_hip.so"
_LIB = None

i64 = C.c_int64
dbl = C.c_double
vp = C.c_void_p


class QolError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise QolError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "(the HIP extension is required, there is no fallback)")
        _LIB = C.CDLL(str(LIB_PATH))
        _LIB.qol_last_error.restype = C.c_char_p
        _LIB.qol_geo_count.restype = i64
        _LIB.qol_geo_charts.restype = i64
        _LIB.qol_geo_charts_gpu.restype = i64
        _LIB.qol_geo_curvature_tail.restype = C.c_double
        _LIB.qol_last_status.restype = i64
        _LIB.qol_last_projection_unconverged.restype = i64
    return _LIB


def last_error():
    return lib().qol_last_error().decode()


def check(rc, what):
    if rc != 0:
        raise QolError(f"{what} failed ({rc}): {last_error()}")


def ptr(a):
    return None if a is None else a.ctypes.data_as(vp)
