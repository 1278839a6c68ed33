"""Triangular views of a BandedMatrix, mirroring the reference's triangular banded layer.

  tbsv! / tbmv!  (checks, argument order)          src/blas.jl:71-147
  bandeddata of a triangle: rows 1:u+1 (upper) or u+1:l+u+1 (lower)
                 of the parent's band array       src/tribanded.jl:19-26
  ldiv!(UpperTriangular(A), x) and friends -> tbsv src/tribanded.jl:75-96
  lmul!(UpperTriangular(A), x) and friends -> tbmv src/tribanded.jl:47-54
The triangle's data is never copied: the library receives a pointer into the
parent's band array with the parent's leading dimension.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .bandedmatrix import BandedMatrix, DimensionMismatch, _prefix
from .device import DeviceArray


@dataclass
class BandView:
    """Rows row0 .. row0+rows-1 of a band array in HBM (a view, like
    view(bandeddata(A), rowinds, :) in src/tribanded.jl:19-26)."""

    data: DeviceArray
    row0: int
    rows: int

    @property
    def ptr(self) -> int:
        return self.data.ptr + self.row0 * self.data.dtype.itemsize

    @property
    def ld(self) -> int:
        return self.data.ld

    @property
    def ncols(self) -> int:
        return self.data.shape[1]

    @property
    def dtype(self):
        return self.data.dtype


def _as_view(A) -> BandView:
    if isinstance(A, BandView):
        return A
    if isinstance(A, DeviceArray):
        return BandView(A, 0, A.shape[0])
    raise TypeError("expected band data (DeviceArray or BandView)")


def _columns(x: DeviceArray):
    """(pointer, stride) of each column of a device vector or matrix."""
    if len(x.shape) == 1:
        return [x.ptr]
    return [x.ptr + c * x.ld * x.dtype.itemsize for c in range(x.shape[1])]


def _tb_call(name: str, uplo: str, trans: str, diag: str, m: int, k: int, A, x: DeviceArray) -> DeviceArray:
    V = _as_view(A)
    n = V.ncols
    if V.rows < k + 1:
        raise ValueError("triangular banded data missing")
    if n != m:
        raise DimensionMismatch(f"matrix is not square: dimensions are {n}, {m}")
    if x.shape[0] != n:
        raise DimensionMismatch(f"size of A is {n} != length(x) = {x.shape[0]}")
    if V.dtype != x.dtype:
        raise TypeError("A and x must have the same element type")
    if n == 0:
        return x
    fn = f"bh_{_prefix(x.dtype)}{name}"
    if name == "tbsv" and len(x.shape) == 2:  # all columns in one launch
        _lib.call(fn + "_multi", uplo.encode(), trans.encode(), diag.encode(), n, k, V.ptr, V.ld, x.shape[1], x.ptr,
                  x.ld)
        return x
    for p in _columns(x):
        _lib.call(fn, uplo.encode(), trans.encode(), diag.encode(), n, k, V.ptr, V.ld, p, 1)
    return x


def tbsv_(uplo: str, trans: str, diag: str, m: int, k: int, A, x: DeviceArray) -> DeviceArray:
    """tbsv!(uplo, trans, diag, m, k, A, x): x <- op(T)^{-1} x in place (src/blas.jl:110-147).
    A is the (>= k+1) x n triangular band data; the columns of a matrix x are
    solved independently (one launch for all of them)."""
    return _tb_call("tbsv", uplo, trans, diag, m, k, A, x)


_IDENTITY = {}


def _identity_pivots(n: int) -> DeviceArray:
    """1..n as device int64 (no interchanges), kept per length."""
    ipiv = _IDENTITY.get(n)
    if ipiv is None:
        ipiv = DeviceArray.from_numpy(np.arange(1, n + 1, dtype=np.int64))
        _IDENTITY.clear()
        _IDENTITY[n] = ipiv
    return ipiv


def _tbsv_partitioned(trans: str, k: int, V: BandView, x: DeviceArray) -> bool:
    """Upper, non-unit op(T) x = b by the partitioned band solve: T's band data
    (diagonal at row k, leading dimension >= k+1) is LAPACK's AB layout of a
    band LU with kl = 0 and ku = k, so bh_?gbtrs_part with identity pivots
    solves it (rows split into blocks solved at once; equal to tbsv to
    rounding).  False when k is outside its range (k <= 63)."""
    from .lu import _part_workspace

    n = V.ncols
    nrhs = 1 if len(x.shape) == 1 else x.shape[1]
    p = _prefix(x.dtype)
    ws = _part_workspace(p, n, 0, k, nrhs)
    if ws is None:
        return False
    _lib.call(f"bh_{p}gbtrs_part", trans.encode(), n, 0, k, nrhs, V.ptr, V.ld, _identity_pivots(n).ptr, x.ptr,
              x.ld, ws.ptr, ws.nbytes)
    return True


def tbmv_(uplo: str, trans: str, diag: str, m: int, k: int, A, x: DeviceArray) -> DeviceArray:
    """tbmv!(uplo, trans, diag, m, k, A, x): x <- op(T) x in place (src/blas.jl:71-107)."""
    return _tb_call("tbmv", uplo, trans, diag, m, k, A, x)


class _Triangular:
    UPLO = "U"
    UNIT = "N"

    def __init__(self, parent: BandedMatrix):
        if not isinstance(parent, BandedMatrix):
            raise TypeError("parent must be a BandedMatrix")
        l, u = parent.bandwidths()
        if l < 0 or u < 0:
            raise ValueError("triangular views need bandwidths l, u >= 0")
        self.parent = parent

    @property
    def shape(self):
        return self.parent.shape

    @property
    def dtype(self):
        return self.parent.dtype

    def bandwidths(self):
        """(min(0, l), u) for upper, (l, min(0, u)) for lower (src/tribanded.jl:4-7)."""
        l, u = self.parent.bandwidths()
        return (min(0, l), u) if self.UPLO == "U" else (l, min(0, u))

    def bandwidth(self) -> int:
        """The triangle's off-diagonal count: bandwidth(A, 2) upper, bandwidth(A, 1) lower."""
        l, u = self.parent.bandwidths()
        return u if self.UPLO == "U" else l

    def bandeddata(self) -> BandView:
        l, u = self.parent.bandwidths()
        D = self.parent.bandeddata()
        return BandView(D, 0, u + 1) if self.UPLO == "U" else BandView(D, u, l + 1)

    def ldiv_(self, x: DeviceArray, trans: str = "N", method: str = "exact") -> DeviceArray:
        """ldiv!(T, x) ('N'), ldiv!(transpose(T), x) ('T'), in place.

        method="exact": tbsv with OpenBLAS's operation order, bit for bit.
        method="partitioned": upper, non-unit triangles with k <= 63
        off-diagonals go through the partitioned band solve (equal to
        rounding); every other case falls back to the exact solve."""
        m, n = self.shape
        if m != n:
            raise DimensionMismatch(f"matrix is not square: dimensions are {m}, {n}")
        if method not in ("exact", "partitioned"):
            raise ValueError(f"unknown solve method {method!r}")
        k, V = self.bandwidth(), self.bandeddata()
        if method == "partitioned" and self.UPLO == "U" and self.UNIT == "N" and m > 0:
            if x.shape[0] != n:
                raise DimensionMismatch(f"size of A is {n} != length(x) = {x.shape[0]}")
            if V.dtype != x.dtype:
                raise TypeError("A and x must have the same element type")
            if _tbsv_partitioned(trans, k, V, x):
                return x
        return tbsv_(self.UPLO, trans, self.UNIT, m, k, V, x)

    def lmul_(self, x: DeviceArray, trans: str = "N") -> DeviceArray:
        """lmul!(T, x) ('N'), lmul!(transpose(T), x) ('T'), in place."""
        m, n = self.shape
        if m != n:
            raise DimensionMismatch(f"matrix is not square: dimensions are {m}, {n}")
        return tbmv_(self.UPLO, trans, self.UNIT, m, self.bandwidth(), self.bandeddata(), x)

    def solve(self, b: DeviceArray) -> DeviceArray:
        """T \\ b (a new vector)."""
        return self.ldiv_(b.copy())

    def __matmul__(self, x: DeviceArray) -> DeviceArray:
        return self.lmul_(x.copy())

    def to_dense(self) -> np.ndarray:
        M = self.parent.to_dense()
        M = np.triu(M) if self.UPLO == "U" else np.tril(M)
        if self.UNIT == "U":
            np.fill_diagonal(M, 1.0)
        return M


class UpperTriangular(_Triangular):
    UPLO, UNIT = "U", "N"


class LowerTriangular(_Triangular):
    UPLO, UNIT = "L", "N"


class UnitUpperTriangular(_Triangular):
    UPLO, UNIT = "U", "U"


class UnitLowerTriangular(_Triangular):
    UPLO, UNIT = "L", "U"
