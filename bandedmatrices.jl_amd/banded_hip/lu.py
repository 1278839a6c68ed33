"""Banded LU and solve, mirroring the reference's BandedLU on the hot path.

  BandedLU{T,S}(factors, ipiv, info)  src/banded/BandedLU.jl:10-27
  lu / lu!                            src/banded/BandedLU.jl:90-111
  ldiv!(::BandedLU, B) ('N')          src/banded/linalg.jl:24-30
  ldiv!(transpose(F), B) ('T')        src/banded/linalg.jl:41-47,55-56
  A \\ b (Ldiv materialize!)          src/banded/linalg.jl:5-9,75
Factors stay in HBM; ipiv is kept both on the device (for solves) and as the
host int64 vector the reference's BandedLU carries (BandedLU.jl:12).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .bandedmatrix import BandedMatrix, DimensionMismatch, _prefix
from .device import DeviceArray


class SingularException(Exception):
    """Zero pivot reported by gbtrf (LAPACK info > 0)."""

    def __init__(self, info):
        super().__init__(f"SingularException({info})")
        self.info = info


class BandedLU:
    """factors: the padded BandedMatrix with bands (l, l+u); ipiv 1-based."""

    def __init__(self, factors: BandedMatrix, ipiv_dev: DeviceArray | None, ipiv: np.ndarray, info: int):
        self.factors = factors
        self.ipiv_dev = ipiv_dev
        self.ipiv = ipiv
        self.info = int(info)

    @property
    def shape(self):
        return self.factors.shape

    def size(self, d=None):
        return self.shape if d is None else self.shape[d - 1]

    def issuccess(self) -> bool:
        return self.info == 0

    @property
    def kl(self):
        return self.factors.l

    @property
    def ku(self):
        # the padded factors carry u+l super-diagonals; LAPACK's KU is u
        return self.factors.u - self.factors.l

    def transpose(self) -> "TransposeFact":
        return TransposeFact(self)

    @property
    def T(self):
        return self.transpose()

    @property
    def p(self) -> np.ndarray:
        """Row permutation (0-based) from ipiv, like ipiv2perm."""
        m = self.shape[0]
        perm = np.arange(m)
        for j, q in enumerate(self.ipiv):
            q -= 1
            if q != j:
                perm[j], perm[q] = perm[q], perm[j]
        return perm


class TransposeFact:
    def __init__(self, parent: BandedLU):
        self.parent = parent


def lu_(A: BandedMatrix, check: bool = True) -> BandedLU:
    """lu!(A): A must already carry the l extra super-diagonals (bands (l, l+u))."""
    m, n = A.shape
    l, uu = A.bandwidths()
    ku = uu - l
    p = _prefix(A.dtype)
    mn = min(m, n)
    if m == 0:
        return BandedLU(A, None, np.zeros(0, np.int64), 0)
    ipiv_dev = DeviceArray((max(mn, 1),), np.int64)
    info = C.c_int64(0)
    _lib.call(f"bh_{p}gbtrf", m, n, l, ku, A.data.ptr, A.data.ld, ipiv_dev.ptr, C.byref(info))
    ipiv = ipiv_dev.numpy()[:mn].astype(np.int64)
    F = BandedLU(A, ipiv_dev, ipiv, info.value)
    if check and info.value > 0:
        raise SingularException(info.value)
    return F


def pad_for_lu(A: BandedMatrix, dtype=None) -> BandedMatrix:
    """BandedMatrix{float(T)}(A, (l, l+u)) (src/banded/BandedMatrix.jl:222-232)."""
    m, n = A.shape
    l, u = A.bandwidths()
    dt = np.dtype(dtype or A.dtype)
    ab = DeviceArray((2 * l + u + 1, n), dt)
    if dt == A.dtype:
        _lib.call(f"bh_{_prefix(dt)}gbpad", m, n, l, u, A.data.ptr, A.data.ld, ab.ptr, ab.ld)
    elif A.dtype == np.float64 and dt == np.float32:
        _lib.call("bh_dsgbpad", m, n, l, u, A.data.ptr, A.data.ld, ab.ptr, ab.ld)
    else:
        raise TypeError(f"no padded copy from {A.dtype} to {dt}")
    return BandedMatrix(ab, m, l, l + u)


def lu(A: BandedMatrix, check: bool = True) -> BandedLU:
    """_lu: pad to (l, l+u), then lu! (src/banded/BandedLU.jl:108-111)."""
    if A.l < 0 or A.u < 0:
        raise NotImplementedError("lu of a matrix with negative bandwidths")
    return lu_(pad_for_lu(A), check=check)


_PART_WS: dict = {}
_BLK_WS: dict = {}


def _blocked_workspace(prec: str, n: int, kl: int, ku: int):
    """Device workspace of the MFMA-blocked solve (kept per shape), or None
    when the bands are outside its range."""
    key = (prec, n, kl, ku)
    ws = _BLK_WS.get(key)
    if ws is None:
        nb = C.c_size_t(0)
        _lib.call("bh_gbtrs_blocked_workspace", prec.encode(), n, kl, ku, C.byref(nb))
        if nb.value == 0:
            return None
        ws = DeviceArray((nb.value,), np.uint8)
        _BLK_WS.clear()
        _BLK_WS[key] = ws
    return ws


def _part_workspace(prec: str, n: int, kl: int, ku: int, nrhs: int):
    """Device workspace of the partitioned solve (kept per shape), or None
    when the bands are outside its range."""
    key = (prec, n, kl, ku, nrhs)
    ws = _PART_WS.get(key)
    if ws is None:
        nb = C.c_size_t(0)
        _lib.call("bh_gbtrs_part_workspace", prec.encode(), n, kl, ku, nrhs, C.byref(nb))
        if nb.value == 0:
            return None
        ws = DeviceArray((nb.value,), np.uint8)
        _PART_WS.clear()
        _PART_WS[key] = ws
    return ws


def ldiv_(F, B: DeviceArray, method: str = "exact") -> DeviceArray:
    """ldiv!(F, B) in place; F a BandedLU or transpose(F), or a triangular view
    of a BandedMatrix (UpperTriangular(A) etc., solved with tbsv).

    method="exact": LAPACK's gbtrs operation order on the GPU, bit for bit.
    method="partitioned": the rows split into blocks solved at once across the
    GPU (bh_?gbtrs_part, kl <= 31 and kl + ku <= 63): equal to rounding; falls
    back to the exact solve outside that range.
    method="blocked": 'N' only, panels of 32 rows on the matrix cores
    (bh_?gbtrs_blocked): equal to rounding; for wide bands; falls back to the
    exact solve for 'T' or outside its range.
    method="auto": from n = 4096 rows on, partitioned for narrow bands (kl <= 31)
    and blocked for wider ones ('N'); exact below."""
    from .symbanded import BandedCholesky, BandedLDLt, cholesky_ldiv_, ldlt_ldiv_
    from .tribanded import _Triangular

    # triangular views and Cholesky factors: "partitioned" as asked, every
    # other method their exact solve
    if isinstance(F, _Triangular):
        return F.ldiv_(B, method="partitioned" if method == "partitioned" else "exact")
    if isinstance(F, BandedCholesky):
        return cholesky_ldiv_(F, B, method="partitioned" if method == "partitioned" else "exact")
    if isinstance(F, BandedLDLt):
        return ldlt_ldiv_(F, B)
    trans = "N"
    if isinstance(F, TransposeFact):
        F, trans = F.parent, "T"
    m, n = F.shape
    if m != n:
        raise DimensionMismatch(f"matrix is not square: dimensions are {F.shape}")
    if B.shape[0] != n:
        raise DimensionMismatch(f"B has leading dimension {B.shape[0]}, but needs {n}")
    if B.dtype != F.factors.dtype:
        raise TypeError("mixed element types")
    if m == 0:
        return B
    nrhs = 1 if len(B.shape) == 1 else B.shape[1]
    p = _prefix(B.dtype)
    if method == "auto":
        method = "exact" if n < 4096 else ("partitioned" if F.kl <= 31 else "blocked")
    if method == "blocked":
        ws = _blocked_workspace(p, n, F.kl, F.ku) if trans == "N" else None
        if ws is not None:
            _lib.call(f"bh_{p}gbtrs_blocked", n, F.kl, F.ku, nrhs, F.factors.data.ptr, F.factors.data.ld,
                      F.ipiv_dev.ptr, B.ptr, B.ld, ws.ptr, ws.nbytes)
            return B
        method = "exact"
    if method == "partitioned":
        ws = _part_workspace(p, n, F.kl, F.ku, nrhs)
        if ws is not None:
            _lib.call(f"bh_{p}gbtrs_part", trans.encode(), n, F.kl, F.ku, nrhs, F.factors.data.ptr,
                      F.factors.data.ld, F.ipiv_dev.ptr, B.ptr, B.ld, ws.ptr, ws.nbytes)
            return B
    elif method != "exact":
        raise ValueError(f"unknown solve method {method!r}")
    _lib.call(f"bh_{p}gbtrs", trans.encode(), n, F.kl, F.ku, nrhs, F.factors.data.ptr, F.factors.data.ld,
              F.ipiv_dev.ptr, B.ptr, B.ld)
    return B


def solve(A: BandedMatrix, b: DeviceArray, method: str = "auto") -> DeviceArray:
    """A \\ b: checksquare, factorize, solve a copy of b (src/banded/linalg.jl:5-9).

    The factorisation is LAPACK's (same pivots); the solve takes ldiv_'s
    `method`: by default ("auto") the partitioned solve for long systems with
    narrow bands and the MFMA-blocked one for wide bands (n >= 4096; backward
    error at rounding level, not bit for bit), the bit-exact one otherwise;
    method="exact" reproduces the reference's LAPACK solve bit for bit."""
    m, n = A.shape
    if m != n:
        raise DimensionMismatch(f"matrix is not square: dimensions are {A.shape}")
    if b.shape[0] != m:
        raise DimensionMismatch(f"b has length {b.shape[0]}, A has {m} rows")
    x = b.copy()
    return ldiv_(lu(A), x, method)
