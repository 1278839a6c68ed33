"""Device arrays in HBM, owned by the caller, moved with the C ABI's copy calls."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    """A column-major (Fortran order) array in device memory.

    Holds a raw device pointer from ``bh_malloc``; freed on ``free()`` or
    garbage collection.  ``ptr`` is what the C ABI receives.
    """

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        _lib.call("bh_malloc", C.byref(p), self.nbytes)
        self._ptr = p.value
        self._owner = True

    @property
    def ptr(self) -> int:
        return self._ptr or 0

    @property
    def ld(self) -> int:
        """Leading dimension (rows) for 2-d arrays."""
        return max(1, self.shape[0]) if self.shape else 1

    @classmethod
    def from_numpy(cls, a: np.ndarray, dtype=None) -> "DeviceArray":
        h = np.asfortranarray(a, dtype=dtype or a.dtype)
        d = cls(h.shape, h.dtype)
        if d.nbytes:
            _lib.call("bh_memcpy_htod", d.ptr, h.ctypes.data, d.nbytes)
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float64) -> "DeviceArray":
        d = cls(shape, dtype)
        if d.nbytes:
            _lib.call("bh_memset", d.ptr, 0, d.nbytes)
        return d

    def numpy(self) -> np.ndarray:
        h = np.empty(self.shape, dtype=self.dtype, order="F")
        if self.nbytes:
            _lib.call("bh_memcpy_dtoh", h.ctypes.data, self.ptr, self.nbytes)
        return h

    def copy(self) -> "DeviceArray":
        d = DeviceArray(self.shape, self.dtype)
        if self.nbytes:
            _lib.call("bh_memcpy_dtod", d.ptr, self.ptr, self.nbytes)
        return d

    def upload(self, a: np.ndarray) -> None:
        h = np.asfortranarray(a, dtype=self.dtype)
        if h.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {h.shape} vs {self.shape}")
        if self.nbytes:
            _lib.call("bh_memcpy_htod", self.ptr, h.ctypes.data, self.nbytes)

    def free(self) -> None:
        if self._ptr and self._owner:
            _lib.load().bh_free(self._ptr)
        self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def sync() -> None:
    _lib.call("bh_sync")


class Timer:
    """HIP events on the library stream; elapsed() synchronises the stop event."""

    def __init__(self):
        self._a = C.c_void_p()
        self._b = C.c_void_p()
        _lib.call("bh_event_create", C.byref(self._a))
        _lib.call("bh_event_create", C.byref(self._b))

    def start(self):
        _lib.call("bh_event_record", self._a)

    def stop(self):
        _lib.call("bh_event_record", self._b)

    def elapsed_ms(self) -> float:
        ms = C.c_float()
        _lib.call("bh_event_elapsed_ms", self._a, self._b, C.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            L = _lib.load()
            L.bh_event_destroy(self._a)
            L.bh_event_destroy(self._b)
        except Exception:
            pass
