"""Host-side mirror of the reference's BandedMatrix interface on the hot path.

Names, argument meaning and errors follow BandedMatrices.jl:
  BandedMatrix / _BandedMatrix   src/banded/BandedMatrix.jl:16-28
  bandwidths / bandeddata        src/banded/BandedMatrix.jl:408-409
  brand / brandn                 src/banded/BandedMatrix.jl:184-204
  mul! (gbmv, banded*dense)      src/generic/matmul.jl:26-64,243-256
  A*B (gbmm!)                    src/generic/matmul.jl:1-6,177-189; src/banded/gbmm.jl:207-293
  prodbandwidths                 src/generic/utils.jl:19-21
  BandError                      src/generic/Band.jl:75-88
  view(A, kr, jr) (banded view)  src/banded/BandedMatrix.jl:918-954 (bandshift, bandwidths, bandeddata)
  lmul!/rmul!, axpy!, copyto!    src/generic/broadcast.jl:942-1020, src/banded/BandedMatrix.jl:124-165
  negative-band products         src/generic/matmul.jl:41-59, src/banded/gbmm.jl:231-279
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

    def view(self, kr: range, jr: range) -> "BandedMatrix":
        """view(A, kr, jr) for unit ranges (0-based, step 1): a banded view sharing
        A's band array (BandedSubBandedMatrix, src/banded/BandedMatrix.jl:918-954).
        bandshift = first(kr) - first(jr); bandwidths = min.(size .- 1, (l - s, u + s));
        the view's band rows are U-u .. U+l of A's (0-based), U = u + s, columns jr."""
        m, n = self.shape
        if len(kr) and (kr.start < 0 or kr.stop > m) or len(jr) and (jr.start < 0 or jr.stop > n):
            raise IndexError(f"view {kr}, {jr} of a {m}x{n} matrix")
        lv, uv, row0, h = view_geometry(self.l, self.u, kr, jr)
        return BandedMatrix(self.data.view(row0, jr.start if len(jr) else 0, (h, len(jr))), len(kr), lv, uv)

    # element-wise (the reference's fast paths on bandeddata)
    def lmul_(self, alpha) -> "BandedMatrix":
        """lmul!(alpha, A): every stored entry times alpha (src/generic/broadcast.jl:942-949)."""
        m, n = self.shape
        _lib.call(f"bh_{_prefix(self.dtype)}gbscal", m, n, self.l, self.u, alpha, self.data.ptr, self.data.ld,
                  0, m, 0, n, 0)
        return self

    rmul_ = lmul_

    def __mul__(self, alpha):
        if not np.isscalar(alpha):
            return NotImplemented
        return self.copy().lmul_(alpha)

    __rmul__ = __mul__

    def __add__(self, other):
        return _plus(self, other, 1.0)

    def __sub__(self, other):
        return _plus(self, other, -1.0)

    def __neg__(self):
        return self * -1.0

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


def view_geometry(l: int, u: int, kr: range, jr: range):
    """Bandwidths (lv, uv) of view(A, kr, jr) and the first row / row count of its
    band inside A's band array (src/banded/BandedMatrix.jl:918-954, 0-based):
    s = first(kr) - first(jr), (lv, uv) = min.(size .- 1, (l - s, u + s)), rows
    U-uv .. U+lv with U = u + s."""
    if kr.step != 1 or jr.step != 1:
        raise ValueError("banded views take unit ranges")
    sft = kr.start - jr.start
    lv = min(len(kr) - 1, l - sft)
    uv = min(len(jr) - 1, u + sft)
    h = max(0, lv + uv + 1)
    return lv, uv, (u + sft - uv) if h else 0, h


def _BandedMatrix(data: DeviceArray, m: int, l: int, u: int) -> BandedMatrix:
    return BandedMatrix(data, m, l, u)


def brand(dtype, n, m, l, u, *, seed=0, normal=False) -> BandedMatrix:
    """_BandedMatrix(rand(T, l+u+1, m), n, l, u): every stored slot random
    (src/banded/BandedMatrix.jl:184-192), from a seeded numpy generator."""
    rng = np.random.default_rng(seed)
    h = max(0, l + u + 1)
    d = rng.standard_normal((h, m)) if normal else rng.random((h, m))
    return BandedMatrix.from_host(np.asfortranarray(d.astype(dtype)), n, l, u)


def axpy_(a, X: BandedMatrix, Y: BandedMatrix) -> BandedMatrix:
    """axpy!(a, X, Y): Y = a*X + Y on the bands X and Y share; BandError if X has
    a nonzero outside Y's bands (banded_generic_axpy!, src/generic/broadcast.jl:978-1020)."""
    if X.shape != Y.shape:
        raise DimensionMismatch(f"X has size {X.shape} but Y has size {Y.shape}")
    if X.dtype != Y.dtype:
        raise TypeError("mixed element types")
    m, n = X.shape
    try:
        _lib.call(f"bh_{_prefix(X.dtype)}gbaxpy", m, n, a, X.data.ptr, X.data.ld, X.l, X.u, Y.data.ptr, Y.data.ld,
                  Y.l, Y.u)
    except _lib.BandedHIPError as e:
        if e.status == -1003:
            raise BandError(str(e)) from None
        raise
    return Y


def copyto_(B: BandedMatrix, A: BandedMatrix) -> BandedMatrix:
    """copyto!(B, A) between banded matrices of the same size: A's entries copied,
    the rest of B's band zeroed; BandError if A has a nonzero outside B's bands."""
    if A.shape != B.shape:
        raise DimensionMismatch(f"A has size {A.shape} but B has size {B.shape}")
    if A.dtype != B.dtype:
        raise TypeError("mixed element types")
    m, n = A.shape
    try:
        _lib.call(f"bh_{_prefix(A.dtype)}gbcopy", m, n, A.data.ptr, A.data.ld, A.l, A.u, B.data.ptr, B.data.ld,
                  B.l, B.u)
    except _lib.BandedHIPError as e:
        if e.status == -1003:
            raise BandError(str(e)) from None
        raise
    return B


def convert(A: BandedMatrix, bands) -> BandedMatrix:
    """BandedMatrix(A, (l, u)): a new matrix with the given bandwidths holding A."""
    B = BandedMatrix.undef(A.shape, bands, A.dtype)
    return copyto_(B, A)


def _plus(A: BandedMatrix, B: BandedMatrix, sign: float) -> BandedMatrix:
    """A + B / A - B with bandwidths max.(bandwidths(A), bandwidths(B))."""
    if not isinstance(B, BandedMatrix):
        return NotImplemented
    C = convert(A, (max(A.l, B.l), max(A.u, B.u)))
    return axpy_(sign, B, C)


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
        # Negative bandwidths: the product runs on views that drop the zero
        # rows/columns (src/banded/gbmm.jl:231-251).  A view's band slots past
        # its own rows are entries of the parent, so beta is applied to C once
        # here (0 when beta == 0) and the pruned product adds alpha*A*B with
        # beta = 1, which leaves every slot it scales unchanged.
        _scale_block(C, 0, n, 0, m, beta)
        _gbmm_views(A, B, C, alpha)
        return C
    p = _prefix(A.dtype)
    try:
        _lib.call(f"bh_{p}gbmm", n, k, m, alpha, A.data.ptr, A.data.ld, A.l, A.u, B.data.ptr, B.data.ld, B.l, B.u,
                  beta, C.data.ptr, C.data.ld, C.l, C.u)
    except _lib.BandedHIPError as e:
        if e.status == -1003:
            raise BandError(str(e)) from None
        raise
    return C


def _rows_from(A: BandedMatrix, s: int) -> BandedMatrix:
    """view(A, s:m, :): the same band array, bands (l - s, u + s)."""
    return BandedMatrix(A.data, A.m - s, A.l - s, A.u + s)


def _cols_from(A: BandedMatrix, s: int) -> BandedMatrix:
    """view(A, :, s:n): the band array from column s, bands (l + s, u - s)."""
    h, n = A.data.shape
    return BandedMatrix(A.data.view(0, s, (h, n - s)), A.m, A.l + s, A.u - s)


def _scale_block(C: BandedMatrix, r0, r1, c0, c1, beta) -> None:
    """C[r0:r1, c0:c1] = beta*C there (0 when beta == 0): _fill_lmul! on a view."""
    m, n = C.shape
    if r1 > r0 and c1 > c0 and C.l + C.u + 1 > 0:
        _lib.call(f"bh_{_prefix(C.dtype)}gbscal", m, n, C.l, C.u, beta, C.data.ptr, C.data.ld, r0, r1, c0, c1, 1)


def _gbmm_views(A: BandedMatrix, B: BandedMatrix, C: BandedMatrix, alpha) -> None:
    """C += alpha*A*B when a bandwidth of A, B or C is negative: zero rows and
    columns are dropped through views until every bandwidth is >= 0, as the
    band-pruning branches of gbmm! do (src/banded/gbmm.jl:231-251)."""
    (Am, An), Bn = A.shape, B.shape[1]
    if -A.l > A.u or -B.l > B.u or C.l + C.u < 0 or Am == 0 or An == 0 or Bn == 0:
        return  # a factor without bands (or nothing to add)
    if A.l < 0:  # A's first -l columns are zero: drop them and B's first -l rows
        s = min(-A.l, An)
        return _gbmm_views(_cols_from(A, s), _rows_from(B, s), C, alpha)
    if A.u < 0:  # A's first -u rows are zero, and so are C's rows there
        s = min(-A.u, Am)
        return _gbmm_views(_rows_from(A, s), B, _rows_from(C, s), alpha)
    if B.l < 0:  # B's first -l columns are zero, and so are C's columns there
        s = min(-B.l, Bn)
        return _gbmm_views(A, _cols_from(B, s), _cols_from(C, s), alpha)
    if B.u < 0:  # B's first -u rows are zero: drop them and A's first -u columns
        s = min(-B.u, An)
        return _gbmm_views(_cols_from(A, s), _rows_from(B, s), C, alpha)
    if C.u < 0:  # C stores nothing in its first -u rows
        s = min(-C.u, Am)
        return _gbmm_views(_rows_from(A, s), B, _rows_from(C, s), alpha)
    if C.l < 0:  # C stores nothing in its first -l columns
        s = min(-C.l, Bn)
        return _gbmm_views(A, _cols_from(B, s), _cols_from(C, s), alpha)
    gbmm(A, B, C, alpha, 1.0)
