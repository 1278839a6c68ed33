"""ctypes binding of libbandedhip.so (the C ABI declared in include/bandedhip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object is missing, and every call raises on a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libbandedhip.so")

i64 = C.c_int64
vp = C.c_void_p
dbl = C.c_double
flt = C.c_float
ch = C.c_char
pi64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes); mirrors include/bandedhip.h
SIGNATURES: dict[str, tuple] = {
    "bh_init": (C.c_int, [C.c_int]),
    "bh_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "bh_last_error": (C.c_char_p, []),
    "bh_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "bh_free": (C.c_int, [vp]),
    "bh_memcpy_htod": (C.c_int, [vp, vp, C.c_size_t]),
    "bh_memcpy_dtoh": (C.c_int, [vp, vp, C.c_size_t]),
    "bh_memcpy_dtod": (C.c_int, [vp, vp, C.c_size_t]),
    "bh_memset": (C.c_int, [vp, C.c_int, C.c_size_t]),
    "bh_sync": (C.c_int, []),
    "bh_set_stream": (C.c_int, [vp]),
    "bh_get_stream": (vp, []),
    "bh_event_create": (C.c_int, [C.POINTER(vp)]),
    "bh_event_record": (C.c_int, [vp]),
    "bh_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(flt)]),
    "bh_event_destroy": (C.c_int, [vp]),
}
for _p, _t in (("d", dbl), ("s", flt)):
    SIGNATURES[f"bh_{_p}gbmv"] = (C.c_int, [ch, i64, i64, i64, i64, _t, vp, i64, vp, i64, _t, vp, i64])
    SIGNATURES[f"bh_{_p}gbmv_multi"] = (C.c_int, [ch, i64, i64, i64, i64, i64, _t, vp, i64, vp, i64, _t, vp, i64])
    SIGNATURES[f"bh_{_p}gbmm"] = (C.c_int, [i64, i64, i64, _t, vp, i64, i64, i64, vp, i64, i64, i64, _t, vp, i64,
                                            i64, i64])
    SIGNATURES[f"bh_{_p}gbpad"] = (C.c_int, [i64, i64, i64, i64, vp, i64, vp, i64])
    SIGNATURES[f"bh_{_p}gbtrf"] = (C.c_int, [i64, i64, i64, i64, vp, i64, vp, pi64])
    SIGNATURES[f"bh_{_p}gbtrs"] = (C.c_int, [ch, i64, i64, i64, i64, vp, i64, vp, vp, i64])
    SIGNATURES[f"bh_{_p}gbqrf"] = (C.c_int, [i64, i64, i64, i64, vp, i64, vp])
    SIGNATURES[f"bh_{_p}gbqrm"] = (C.c_int, [ch, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64])
    SIGNATURES[f"bh_{_p}gbqrs"] = (C.c_int, [i64, i64, i64, i64, vp, i64, vp, i64, vp, i64])
    SIGNATURES[f"bh_{_p}gbqrmr"] = (C.c_int, [ch, i64, i64, i64, i64, i64, vp, i64, vp, vp, i64])
    SIGNATURES[f"bh_{_p}tbsv"] = (C.c_int, [ch, ch, ch, i64, i64, vp, i64, vp, i64])
    SIGNATURES[f"bh_{_p}tbmv"] = (C.c_int, [ch, ch, ch, i64, i64, vp, i64, vp, i64])
    SIGNATURES[f"bh_{_p}tbsv_multi"] = (C.c_int, [ch, ch, ch, i64, i64, vp, i64, i64, vp, i64])
    SIGNATURES[f"bh_{_p}gbscal"] = (C.c_int, [i64, i64, i64, i64, _t, vp, i64, i64, i64, i64, i64, C.c_int])
    SIGNATURES[f"bh_{_p}gbaxpy"] = (C.c_int, [i64, i64, _t, vp, i64, i64, i64, vp, i64, i64, i64])
    SIGNATURES[f"bh_{_p}gbcopy"] = (C.c_int, [i64, i64, vp, i64, i64, i64, vp, i64, i64, i64])
    SIGNATURES[f"bh_{_p}pbtrf"] = (C.c_int, [ch, i64, i64, vp, i64, pi64])
    SIGNATURES[f"bh_{_p}pbtrs"] = (C.c_int, [ch, i64, i64, i64, vp, i64, vp, i64])
    SIGNATURES[f"bh_{_p}sbldlt"] = (C.c_int, [ch, i64, i64, vp, i64])
    SIGNATURES[f"bh_{_p}sbldlts"] = (C.c_int, [ch, i64, i64, i64, vp, i64, vp, i64])
    SIGNATURES[f"bh_{_p}sbmv"] = (C.c_int, [ch, i64, i64, _t, vp, i64, vp, i64, _t, vp, i64])

SIGNATURES["bh_batch_plan_create"] = (C.c_int, [i64, vp, vp, vp, C.POINTER(vp)])
SIGNATURES["bh_batch_plan_layout"] = (C.c_int, [vp, vp, vp, pi64, pi64])
SIGNATURES["bh_batch_plan_schedule"] = (C.c_int, [vp, C.c_int, C.c_int, C.c_int])
SIGNATURES["bh_batch_plan_get_schedule"] = (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                      C.POINTER(C.c_int)])
SIGNATURES["bh_batch_plan_destroy"] = (C.c_int, [vp])
for _p in ("d", "s"):
    SIGNATURES[f"bh_{_p}gbsv_batched"] = (C.c_int, [vp, vp, vp, vp, vp])
    SIGNATURES[f"bh_{_p}gbsv_batched_autotune"] = (C.c_int, [vp, vp, vp, C.POINTER(dbl)])
    SIGNATURES[f"bh_{_p}gbsv_batched_graph"] = (C.c_int, [vp, vp, vp, vp, vp, C.POINTER(vp)])
for _p, _t in (("d", dbl), ("s", flt)):
    SIGNATURES[f"bh_{_p}gb_band_grad"] = (C.c_int, [i64, i64, i64, i64, vp, vp, vp, vp, vp, i64])
    SIGNATURES[f"bh_{_p}gb_band_outer"] = (C.c_int, [i64, i64, i64, i64, _t, vp, vp, _t, vp, vp, _t, vp, i64])
SIGNATURES["bh_batch_graph_launch"] = (C.c_int, [vp])
SIGNATURES["bh_batch_graph_destroy"] = (C.c_int, [vp])
SIGNATURES["bh_dsgbpad"] = (C.c_int, [i64, i64, i64, i64, vp, i64, vp, i64])
SIGNATURES["bh_dsgbsv_workspace"] = (C.c_int, [i64, i64, i64, i64, C.POINTER(C.c_size_t)])
SIGNATURES["bh_dsgbsv"] = (C.c_int, [i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, C.c_size_t, vp, pi64, pi64])
SIGNATURES["bh_gbtrs_blocked_workspace"] = (C.c_int, [C.c_char, i64, i64, i64, C.POINTER(C.c_size_t)])
for _p in "ds":
    SIGNATURES[f"bh_{_p}gbtrs_blocked"] = (C.c_int, [i64, i64, i64, i64, vp, i64, vp, vp, i64, vp, C.c_size_t])
SIGNATURES["bh_gbtrs_part_workspace"] = (C.c_int, [C.c_char, i64, i64, i64, i64, C.POINTER(C.c_size_t)])
for _p in "ds":
    SIGNATURES[f"bh_{_p}gbtrs_part"] = (C.c_int, [C.c_char, i64, i64, i64, i64, vp, i64, vp, vp, i64, vp, C.c_size_t])
SIGNATURES["bh_dsgbsv_noalloc"] = (C.c_int, [i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, pi64, pi64])

_lib = None


class BandedHIPError(RuntimeError):
    """A non-zero status from libbandedhip (argument error or runtime failure)."""

    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} returned {status}: {msg}")
        self.fn = fn
        self.status = status


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python bandedmatrices.jl_amd/build.py` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def call(name: str, *args) -> int:
    """Call an entry point and raise BandedHIPError on a non-zero status."""
    L = load()
    st = getattr(L, name)(*args)
    if st != 0:
        raise BandedHIPError(name, st, L.bh_last_error().decode(errors="replace"))
    return st
