"""Banded QR, mirroring the reference's qr on BandedMatrix.

  qr / qr!            src/banded/bandedqr.jl:1-35 (padded copy, _banded_qr!)
  Q' * B, Q * B       src/banded/bandedqr.jl:38-94,150-172 (banded_qr_lmul!)
  B * Q, B * Q'       src/banded/bandedqr.jl:96-148,210-214 (banded_qr_rmul!)
  ldiv!(F::QR, B)     src/banded/bandedqr.jl:175-206 (Q'B then tbsv of R)
The factors stay in HBM: R in the upper l+u bands of the padded copy, the
reflectors in the lower l band, tau as a device vector.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .bandedmatrix import BandedMatrix, DimensionMismatch, _prefix
from .device import DeviceArray
from .lu import pad_for_lu


@dataclass
class BandedQR:
    """factors: padded BandedMatrix with bands (l, l+u); tau: device vector of min(m, n)."""

    factors: BandedMatrix
    tau: DeviceArray

    @property
    def shape(self):
        return self.factors.shape

    @property
    def Q(self):
        return QRPackedQ(self, adjoint=False)

    def R_dense(self) -> np.ndarray:
        """R as a dense host array (min(m,n) x n), for tests and display."""
        m, n = self.shape
        F = self.factors
        D = F.host_data()
        R = np.zeros((min(m, n), n), D.dtype)
        for j in range(n):
            for i in range(max(0, j - F.u), min(j, min(m, n) - 1) + 1):
                R[i, j] = D[F.u + i - j, j]
        return R


@dataclass
class QRPackedQ:
    parent: BandedQR
    adjoint: bool

    @property
    def T(self):
        return QRPackedQ(self.parent, not self.adjoint)

    def lmul_(self, B: DeviceArray) -> DeviceArray:
        """lmul!(Q, B) / lmul!(Q', B) in place."""
        return qr_lmul_(self.parent, B, "T" if self.adjoint else "N")

    def __matmul__(self, B: DeviceArray) -> DeviceArray:
        return self.lmul_(B.copy())

    def __rmatmul__(self, A: DeviceArray) -> DeviceArray:
        return qr_rmul_(A.copy(), self)


def qr_rmul_(A: DeviceArray, Q: QRPackedQ) -> DeviceArray:
    """rmul!(A, Q) / rmul!(A, Q') in place: A (mA x m) <- A Q or A Q'
    (banded_qr_rmul!, src/banded/bandedqr.jl:96-148,210-214)."""
    F = Q.parent
    m, n = F.shape
    if len(A.shape) != 2 or A.shape[1] != m:
        raise DimensionMismatch(f"matrix A has dimensions {A.shape} but matrix Q has dimensions ({m}, {n})")
    if A.dtype != F.factors.dtype:
        raise TypeError("mixed element types")
    l, u = F.factors.bandwidths()
    _lib.call(f"bh_{_prefix(A.dtype)}gbqrmr", b"T" if Q.adjoint else b"N", A.shape[0], m, n, l, u,
              F.factors.data.ptr, F.factors.data.ld, F.tau.ptr, A.ptr, A.ld)
    return A


def qr_(A: BandedMatrix) -> BandedQR:
    """qr!(A): A already carries the l extra super-diagonals (bands (l, l+u))."""
    m, n = A.shape
    l, u = A.bandwidths()
    tau = DeviceArray((max(min(m, n), 1),), A.dtype)
    _lib.call(f"bh_{_prefix(A.dtype)}gbqrf", m, n, l, u, A.data.ptr, A.data.ld, tau.ptr)
    return BandedQR(A, tau)


def qr(A: BandedMatrix) -> BandedQR:
    """_qr: the padded copy BandedMatrix{T}(A, (l, l+u)), then qr! (bandedqr.jl:1-5)."""
    if A.l < 0 or A.u < 0:
        raise NotImplementedError("qr of a matrix with negative bandwidths")
    return qr_(pad_for_lu(A))


def qr_lmul_(F: BandedQR, B: DeviceArray, trans: str) -> DeviceArray:
    m, n = F.shape
    if B.shape[0] != m:
        raise DimensionMismatch(f"B has {B.shape[0]} rows, Q is {m} x {m}")
    if B.dtype != F.factors.dtype:
        raise TypeError("mixed element types")
    nrhs = 1 if len(B.shape) == 1 else B.shape[1]
    l, u = F.factors.bandwidths()
    _lib.call(f"bh_{_prefix(B.dtype)}gbqrm", trans.encode(), m, n, l, u, F.factors.data.ptr, F.factors.data.ld,
              F.tau.ptr, nrhs, B.ptr, B.ld)
    return B


def qr_ldiv_(F: BandedQR, B: DeviceArray, method: str = "exact") -> DeviceArray:
    """ldiv!(F, B): B (m x nrhs) <- Q'B, then its first n rows <- R \\ (Q'B)[1:n].

    method="exact": the reference's order (reflectors, then tbsv with R).
    method="partitioned": the same Q'B, then R's solve by the partitioned band
    solve (R's band rows are LAPACK's AB layout of a band LU with kl = 0 and
    ku = R's bandwidth; identity pivots), equal to rounding, for R bandwidths
    <= 63; wider R falls back to the exact solve."""
    m, n = F.shape
    if m < n:
        raise ValueError("Not implemented for wide matrices")  # bandedqr.jl:175-177
    if B.shape[0] != m:
        raise DimensionMismatch(f"B has {B.shape[0]} rows, A has {m}")
    if B.dtype != F.factors.dtype:
        raise TypeError("mixed element types")
    if method not in ("exact", "partitioned"):
        raise ValueError(f"unknown solve method {method!r}")
    nrhs = 1 if len(B.shape) == 1 else B.shape[1]
    l, u = F.factors.bandwidths()
    if method == "partitioned" and n > 0 and nrhs > 0:
        from .lu import _part_workspace
        from .tribanded import _identity_pivots

        p = _prefix(B.dtype)
        ws = _part_workspace(p, n, 0, u, nrhs)
        if ws is not None:
            qr_lmul_(F, B, "T")
            _lib.call(f"bh_{p}gbtrs_part", b"N", n, 0, u, nrhs, F.factors.data.ptr, F.factors.data.ld,
                      _identity_pivots(n).ptr, B.ptr, B.ld, ws.ptr, ws.nbytes)
            return B
    _lib.call(f"bh_{_prefix(B.dtype)}gbqrs", m, n, l, u, F.factors.data.ptr, F.factors.data.ld, F.tau.ptr, nrhs,
              B.ptr, B.ld)
    return B


def qr_solve(A: BandedMatrix, b: DeviceArray, method: str = "exact") -> DeviceArray:
    """qr(A) \\ b for m >= n: the least-squares solution (first n rows of the result)."""
    F = qr(A)
    x = qr_ldiv_(F, b.copy(), method)
    return x
