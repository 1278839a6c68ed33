"""Symmetric views of a BandedMatrix, band Cholesky and the symmetric product,
mirroring the reference's symmetric banded layer.

  Symmetric(A, uplo), bandwidth, symbandeddata   src/symbanded/symbanded.jl:47-69
  Symmetric(A) * x  (BLAS.sbmv!)                 src/symbanded/symbanded.jl:72-91
  cholesky(Symmetric(A)) (pbtrf!)                src/symbanded/BandedCholesky.jl:1-11, src/lapack.jl:268-292
  ldiv!(::Cholesky, B) (pbtrs!)                  src/symbanded/BandedCholesky.jl:66-79, src/lapack.jl:297-333
  ldlt(Symmetric(A)), ldiv!(::LDLt, B)            src/symbanded/ldlt.jl:8-69
"""
from __future__ import annotations

from dataclasses import dataclass

import ctypes as C
import numpy as np

from . import _lib
from .bandedmatrix import BandedMatrix, DimensionMismatch, _prefix, copyto_
from .device import DeviceArray
from .tribanded import BandView


class PosDefException(Exception):
    """The matrix is not positive definite (Julia's PosDefException); info is LAPACK's."""

    def __init__(self, info: int):
        super().__init__(f"matrix is not positive definite; leading minor of order {info} is not positive")
        self.info = info


class Symmetric:
    """Symmetric(A, uplo): the upper ('U') or lower ('L') triangle of a square
    BandedMatrix read as a symmetric matrix of bandwidth u or l."""

    def __init__(self, parent: BandedMatrix, uplo: str = "U"):
        if uplo not in ("U", "L"):
            raise ValueError("uplo is 'U' or 'L'")
        m, n = parent.shape
        if m != n:
            raise DimensionMismatch("matrix is not square")
        self.parent, self.uplo = parent, uplo

    @property
    def shape(self):
        return self.parent.shape

    @property
    def dtype(self):
        return self.parent.dtype

    def bandwidth(self) -> int:
        l, u = self.parent.bandwidths()
        return u if self.uplo == "U" else l

    def bandwidths(self):
        k = self.bandwidth()
        return (k, k)

    def symbandeddata(self) -> BandView:
        """Rows 1:u+1 ('U') or u+1:u+l+1 ('L') of the parent's band array."""
        l, u = self.parent.bandwidths()
        D = self.parent.bandeddata()
        return BandView(D, 0, u + 1) if self.uplo == "U" else BandView(D, u, l + 1)

    def mul(self, y: DeviceArray, x: DeviceArray, alpha=1.0, beta=0.0) -> DeviceArray:
        """mul!(y, Symmetric(A), x, alpha, beta) -> sbmv."""
        n = self.shape[0]
        if x.shape[0] != n or y.shape[0] != n:
            raise DimensionMismatch("*")
        k = self.bandwidth()
        if k < 0:
            from .bandedmatrix import _scale_rows

            _scale_rows(y, 0, n, beta)
            return y
        V = self.symbandeddata()
        cols = [0] if len(x.shape) == 1 else range(x.shape[1])
        isz = x.dtype.itemsize
        for c in cols:
            _lib.call(f"bh_{_prefix(self.dtype)}sbmv", self.uplo.encode(), n, k, alpha, V.ptr, V.ld,
                      x.ptr + c * x.ld * isz, 1, beta, y.ptr + c * y.ld * isz, 1)
        return y

    def __matmul__(self, x: DeviceArray) -> DeviceArray:
        y = DeviceArray(x.shape, x.dtype)
        return self.mul(y, x)

    def to_dense(self) -> np.ndarray:
        M = self.parent.to_dense()
        T = np.triu(M) if self.uplo == "U" else np.tril(M)
        return T + T.T - np.diag(np.diag(M))


@dataclass
class BandedCholesky:
    """factors: (k+1) x n LAPACK symmetric band storage of U ('U') or L ('L')."""

    factors: BandedMatrix
    uplo: str
    info: int

    @property
    def shape(self):
        return self.factors.shape

    def kd(self) -> int:
        return self.factors.u if self.uplo == "U" else self.factors.l


@dataclass
class BandedLDLt:
    """factors: (k+1) x n storage of the triangle: L's entries where A's were, d on the diagonal."""

    factors: BandedMatrix
    uplo: str

    @property
    def shape(self):
        return self.factors.shape

    def kd(self) -> int:
        return self.factors.u if self.uplo == "U" else self.factors.l


def _triangle_copy(S: Symmetric) -> BandedMatrix:
    k = S.bandwidth()
    n = S.shape[0]
    bands = (0, k) if S.uplo == "U" else (k, 0)
    F = BandedMatrix.zeros((n, n), bands, S.dtype)  # corner slots outside the matrix hold 0
    l, u = S.parent.bandwidths()
    tri = BandedMatrix(S.symbandeddata().data.view(0 if S.uplo == "U" else u, 0, (k + 1, n)), n, *bands)
    return copyto_(F, tri)


def ldlt(S: Symmetric) -> BandedLDLt:
    """ldlt(Symmetric(A)): the triangle copied, then factored in place (no pivoting)."""
    F = _triangle_copy(S)
    _lib.call(f"bh_{_prefix(S.dtype)}sbldlt", S.uplo.encode(), S.shape[0], S.bandwidth(), F.data.ptr, F.data.ld)
    return BandedLDLt(F, S.uplo)


def ldlt_ldiv_(F: BandedLDLt, B: DeviceArray) -> DeviceArray:
    """ldiv!(F, B) in place."""
    n = F.shape[0]
    if B.shape[0] != n:
        raise DimensionMismatch(f"Matrix has dimensions {F.shape} but right hand side has first dimension {B.shape[0]}")
    nrhs = 1 if len(B.shape) == 1 else B.shape[1]
    _lib.call(f"bh_{_prefix(B.dtype)}sbldlts", F.uplo.encode(), n, F.kd(), nrhs, F.factors.data.ptr,
              F.factors.data.ld, B.ptr, max(1, B.ld if len(B.shape) == 2 else n))
    return B


def cholesky(S: Symmetric, check: bool = True) -> BandedCholesky:
    """cholesky(Symmetric(A)): the triangle copied into its own (k+1) x n band
    (bands (0, k) or (k, 0)) and factored in place by pbtrf."""
    k = S.bandwidth()
    n = S.shape[0]
    F = _triangle_copy(S)
    info = C.c_int64(0)
    _lib.call(f"bh_{_prefix(S.dtype)}pbtrf", S.uplo.encode(), n, k, F.data.ptr, F.data.ld, C.byref(info))
    if check and info.value > 0:
        raise PosDefException(info.value)
    return BandedCholesky(F, S.uplo, int(info.value))


def cholesky_ldiv_(F: BandedCholesky, B: DeviceArray, method: str = "exact") -> DeviceArray:
    """ldiv!(F, B) in place (pbtrs).

    method="exact": LAPACK's pbtrs operation order.  method="partitioned": an
    upper factor U (its (k+1) x n band is LAPACK's AB layout of a band LU with
    kl = 0, ku = k) with k <= 63 is solved as U^T y = B, U x = y by the
    partitioned band solve (bh_?gbtrs_part, identity pivots; equal to rounding);
    a lower factor or a wider band falls back to the exact solve."""
    n = F.shape[0]
    if B.shape[0] != n:
        raise DimensionMismatch(f"matrix A has dimensions {F.shape}, but right hand side B has {B.shape}")
    if method not in ("exact", "partitioned"):
        raise ValueError(f"unknown solve method {method!r}")
    nrhs = 1 if len(B.shape) == 1 else B.shape[1]
    if method == "partitioned" and F.uplo == "U" and n > 0 and nrhs > 0:
        from .lu import _part_workspace
        from .tribanded import _identity_pivots

        p = _prefix(B.dtype)
        ws = _part_workspace(p, n, 0, F.kd(), nrhs)
        if ws is not None:
            ipiv = _identity_pivots(n)
            for trans in ("T", "N"):
                _lib.call(f"bh_{p}gbtrs_part", trans.encode(), n, 0, F.kd(), nrhs, F.factors.data.ptr,
                          F.factors.data.ld, ipiv.ptr, B.ptr, B.ld, ws.ptr, ws.nbytes)
            return B
    _lib.call(f"bh_{_prefix(B.dtype)}pbtrs", F.uplo.encode(), n, F.kd(), nrhs, F.factors.data.ptr,
              F.factors.data.ld, B.ptr, max(1, B.ld if len(B.shape) == 2 else n))
    return B
