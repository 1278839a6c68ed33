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
        self._ld = None
        self._parent = None

    @property
    def ptr(self) -> int:
        return self._ptr or 0

    @property
    def ld(self) -> int:
        """Leading dimension (rows) for 2-d arrays."""
        if self._ld is not None:
            return self._ld
        return max(1, self.shape[0]) if self.shape else 1

    @property
    def contiguous(self) -> bool:
        return len(self.shape) < 2 or self.ld == max(1, self.shape[0])

    def view(self, row0: int, col0: int, shape) -> "DeviceArray":
        """A non-owning window (rows row0.., columns col0..) of a 2-d array with
        this array's leading dimension, like view(A, rows, cols) of a strided matrix."""
        if len(self.shape) != 2:
            raise ValueError("view needs a 2-d array")
        nr, nc = (int(s) for s in shape)
        if row0 < 0 or col0 < 0 or nr < 0 or nc < 0 or (nr and row0 + nr > self.shape[0]) or \
                (nc and col0 + nc > self.shape[1]):
            raise IndexError(f"view {row0}:{row0 + nr}, {col0}:{col0 + nc} of {self.shape}")
        v = DeviceArray.__new__(DeviceArray)
        v.shape = (nr, nc)
        v.dtype = self.dtype
        v.nbytes = nr * nc * self.dtype.itemsize
        v._ptr = self.ptr + (row0 + col0 * self.ld) * self.dtype.itemsize
        v._owner = False
        v._ld = self.ld
        v._parent = self  # keeps the allocation alive
        return v

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
        if not self.contiguous:
            nr, nc = self.shape
            if self.nbytes == 0:
                return np.empty(self.shape, dtype=self.dtype, order="F")
            span = (nc - 1) * self.ld + nr
            buf = np.empty(nc * self.ld, dtype=self.dtype)
            _lib.call("bh_memcpy_dtoh", buf.ctypes.data, self.ptr, span * self.dtype.itemsize)
            return np.asfortranarray(buf.reshape((self.ld, nc), order="F")[:nr, :])
        h = np.empty(self.shape, dtype=self.dtype, order="F")
        if self.nbytes:
            _lib.call("bh_memcpy_dtoh", h.ctypes.data, self.ptr, self.nbytes)
        return h

    def copy(self) -> "DeviceArray":
        if not self.contiguous:
            return DeviceArray.from_numpy(self.numpy())
        d = DeviceArray(self.shape, self.dtype)
        if self.nbytes:
            _lib.call("bh_memcpy_dtod", d.ptr, self.ptr, self.nbytes)
        return d

    def upload(self, a: np.ndarray) -> None:
        h = np.asfortranarray(a, dtype=self.dtype)
        if h.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {h.shape} vs {self.shape}")
        if not self.contiguous:
            for c in range(self.shape[1]):
                col = np.ascontiguousarray(h[:, c])
                _lib.call("bh_memcpy_htod", self.ptr + c * self.ld * self.dtype.itemsize, col.ctypes.data, col.nbytes)
            return
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
