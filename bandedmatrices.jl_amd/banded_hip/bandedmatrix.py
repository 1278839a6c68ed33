"""Host-side mirror of the reference's BandedMatrix interface on the hot path.

Names, argument meaning and errors follow BandedMatrices.jl:
  BandedMatrix / _BandedMatrix   src/banded/BandedMatrix.jl:16-28
  bandwidths / bandeddata        src/banded/BandedMatrix.jl:408-409
  brand / brandn                 src/banded/BandedMatrix.jl:184-204
  mul! (gbmv, banded*dense)      src/generic/matmul.jl:26-64,243-256
  A*B (gbmm!)                    src/generic/matmul.jl:1-6,177-189; src/banded/gbmm.jl:207-293
  prodbandwidths                 src/generic/utils.jl:19-21
  BandError                      src/generic/Band.jl:75-88
Data lives in HBM (``DeviceArray``); every operation calls libbandedhip.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceArray


class BandError(Exception):
    """Raised when C lacks the bands a product needs (src/generic/Band.jl:75-88)."""


class DimensionMismatch(ValueError):
    """Size mismatch (Julia's DimensionMismatch)."""


def _prefix(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "d"
    if dt == np.float32:
        return "s"
    raise TypeError(f"BLAS element types only (Float64/Float32), got {dt}")


class BandedMatrix:
    """m x n matrix with bandwidths (l, u) whose (l+u+1) x n band array is in HBM.

    ``data[u + i - j, j] == A[i, j]`` for -l <= j-i <= u (0-based), the layout of
    src/banded/BandedMatrix.jl:16-28.
    """

    def __init__(self, data: DeviceArray, m: int, l: int, u: int):
        if data.shape[0] != l + u + 1 and not (data.shape[0] == 0 and -l > u):
            raise ValueError("Data matrix must have number rows equal to number of bands")
        self.data = data
        self.m = int(m)
        self.l = int(l)
        self.u = int(u)

    # --------------------------------------------------------------- interface
    @property
    def shape(self):
        return (self.m, self.data.shape[1])

    def size(self, d=None):
        return self.shape if d is None else self.shape[d - 1]

    @property
    def dtype(self):
        return self.data.dtype

    def bandwidths(self):
        return (self.l, self.u)

    def bandeddata(self) -> DeviceArray:
        return self.data

    @classmethod
    def from_host(cls, data: np.ndarray, m: int, l: int, u: int, dtype=None) -> "BandedMatrix":
        return cls(DeviceArray.from_numpy(np.asarray(data), dtype=dtype), m, l, u)

    @classmethod
    def undef(cls, shape, bands, dtype=np.float64) -> "BandedMatrix":
        (m, n), (l, u) = shape, bands
        return cls(DeviceArray((max(0, l + u + 1), n), dtype), m, l, u)

    @classmethod
    def zeros(cls, shape, bands, dtype=np.float64) -> "BandedMatrix":
        (m, n), (l, u) = shape, bands
        return cls(DeviceArray.zeros((max(0, l + u + 1), n), dtype), m, l, u)

    def host_data(self) -> np.ndarray:
        return self.data.numpy()

    def to_dense(self) -> np.ndarray:
        d = self.host_data()
        m, n = self.shape
        M = np.zeros((m, n), dtype=d.dtype)
        for j in range(n):
            for i in range(max(0, j - self.u), min(m, j + self.l + 1)):
                M[i, j] = d[self.u + i - j, j]
        return M

    def copy(self) -> "BandedMatrix":
        return BandedMatrix(self.data.copy(), self.m, self.l, self.u)

    def __matmul__(self, other):
        if isinstance(other, BandedMatrix):
            return gbmm(self, other)
        if isinstance(other, DeviceArray):
            if len(other.shape) == 1:
                y = DeviceArray((self.m,), self.dtype)
                return mul(y, self, other)
            Y = DeviceArray((self.m, other.shape[1]), self.dtype)
            return mul(Y, self, other)
        raise TypeError("A @ x needs a device vector/matrix or a BandedMatrix")

    def __repr__(self):
        return f"BandedMatrix({self.m}x{self.shape[1]}, bands={self.bandwidths()}, {self.dtype})"


def _BandedMatrix(data: DeviceArray, m: int, l: int, u: int) -> BandedMatrix:
    return BandedMatrix(data, m, l, u)


def brand(dtype, n, m, l, u, *, seed=0, normal=False) -> BandedMatrix:
    """_BandedMatrix(rand(T, l+u+1, m), n, l, u): every stored slot random
    (src/banded/BandedMatrix.jl:184-192), from a seeded numpy generator."""
    rng = np.random.default_rng(seed)
    h = max(0, l + u + 1)
    d = rng.standard_normal((h, m)) if normal else rng.random((h, m))
    return BandedMatrix.from_host(np.asfortranarray(d.astype(dtype)), n, l, u)


def prodbandwidths(*As):
    return tuple(sum(A.bandwidths()[k] for A in As) for k in (0, 1))


# ------------------------------------------------------------------ products
def mul(y: DeviceArray, A: BandedMatrix, x: DeviceArray, alpha=1.0, beta=0.0, trans="N") -> DeviceArray:
    """mul!(y, A, x, alpha, beta): y = alpha*op(A)*x + beta*y on the device.

    Negative bandwidths are handled with shifted views like
    _banded_muladd! (src/generic/matmul.jl:41-59).
    """
    m, n = A.shape
    rows, cols = (m, n) if trans == "N" else (n, m)
    vec = len(x.shape) == 1
    if x.shape[0] != cols or y.shape[0] != rows or (not vec and (len(y.shape) != 2 or y.shape[1] != x.shape[1])):
        raise DimensionMismatch(f"A is {m}x{n}, x is {x.shape}, y is {y.shape}")
    p = _prefix(A.dtype)
    if x.dtype != A.dtype or y.dtype != A.dtype:
        raise TypeError("mixed element types")
    l, u = A.l, A.u
    nrhs = 1 if vec else x.shape[1]
    itemsize = A.dtype.itemsize
    a_ptr, x_ptr, y_ptr = A.data.ptr, x.ptr, y.ptr
    mm, nn = m, n
    if -l > u:
        # no bands: y = beta*y
        l, u, mm, nn = 0, 0, 0, 0
    elif l < 0:
        # all bands above the diagonal: drop the first -l columns (view(A, :, 1-l:n))
        s = -l
        a_ptr += s * A.data.ld * itemsize
        if trans == "N":
            x_ptr += s * itemsize
        else:
            y_ptr += s * itemsize
            rows -= s
        nn = n - s
        l, u = 0, u + (-s)
    elif u < 0:
        # all bands below the diagonal: drop the first -u rows (view(A, 1-u:m, :))
        s = -u
        if trans == "N":
            _scale_rows(y, 0, s, beta)
            y_ptr += s * itemsize
        else:
            x_ptr += s * itemsize
        mm = m - s
        l, u = l - s, 0
    mm, nn = max(mm, 0), max(nn, 0)
    ldx = x.ld if not vec else max(1, x.shape[0])
    ldy = y.ld if not vec else max(1, y.shape[0])
    if vec:
        _lib.call(f"bh_{p}gbmv", trans.encode(), mm, nn, l, u, alpha, a_ptr, A.data.ld, x_ptr, 1, beta, y_ptr, 1)
    else:
        _lib.call(f"bh_{p}gbmv_multi", trans.encode(), mm, nn, l, u, nrhs, alpha, a_ptr, A.data.ld, x_ptr, ldx,
                  beta, y_ptr, ldy)
    return y


def _scale_rows(y: DeviceArray, r0: int, r1: int, beta: float) -> None:
    """y[r0:r1, :] *= beta (NaN-safe zero for beta == 0), via a 1-row gbmv."""
    if r1 <= r0:
        return
    p = _prefix(y.dtype)
    nrhs = 1 if len(y.shape) == 1 else y.shape[1]
    ld = y.shape[0]
    for r in range(nrhs):
        ptr = y.ptr + (r * ld + r0) * y.dtype.itemsize
        # m rows, zero columns: the kernel only applies beta
        _lib.call(f"bh_{p}gbmv", b"N", r1 - r0, 0, 0, 0, 0.0, 0, 1, 0, 1, beta, ptr, 1)


def gbmm(A: BandedMatrix, B: BandedMatrix, C: BandedMatrix | None = None, alpha=1.0, beta=0.0) -> BandedMatrix:
    """C = alpha*A*B + beta*C (gbmm!, src/banded/gbmm.jl:207-293).  When C is
    None it is allocated with bandwidths min.(size.-1, prodbandwidths)
    (src/generic/matmul.jl:1-6)."""
    n, k = A.shape
    k2, m = B.shape
    if k != k2:
        raise DimensionMismatch(f"A is {A.shape}, B is {B.shape}")
    if A.dtype != B.dtype:
        raise TypeError("mixed element types")
    if C is None:
        Cl = min(n - 1, A.l + B.l)
        Cu = min(m - 1, A.u + B.u)
        C = BandedMatrix.undef((n, m), (Cl, Cu), A.dtype)
        beta = 0.0
    if C.shape != (n, m):
        raise DimensionMismatch(f"C is {C.shape}, expected {(n, m)}")
    if min(A.l, A.u, B.l, B.u, C.l, C.u) < 0:
        raise NotImplementedError("negative bandwidths in banded*banded")
    p = _prefix(A.dtype)
    try:
        _lib.call(f"bh_{p}gbmm", n, k, m, alpha, A.data.ptr, A.data.ld, A.l, A.u, B.data.ptr, B.data.ld, B.l, B.u,
                  beta, C.data.ptr, C.data.ld, C.l, C.u)
    except _lib.BandedHIPError as e:
        if e.status == -1003:
            raise BandError(str(e)) from None
        raise
    return C
