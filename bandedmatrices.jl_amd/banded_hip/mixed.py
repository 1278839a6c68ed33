"""Mixed precision A \\ B: FP32 band LU + FP64 iterative refinement.

No reference counterpart (BandedMatrices.jl factors in the input precision,
src/banded/BandedLU.jl:108-111); this is an alternative drop-in for A \\ b
(src/banded/linalg.jl:5-9) on FP64 data.  Residuals and the solution stay
FP64; the result meets LAPACK dsgesv's backward-error test or comes from the
FP64 fallback (``iters < 0``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .bandedmatrix import BandedMatrix, DimensionMismatch
from .device import DeviceArray


@dataclass
class MixedResult:
    x: DeviceArray
    iters: int          # refinement sweeps; < 0 means the FP64 fallback ran
    info: int           # info of the FP64 fallback factorisation (0 otherwise)
    ipiv32: np.ndarray  # pivots of the FP32 factorisation (1-based)


class MixedSolver:
    """Holds the device workspace so repeated solves allocate nothing new."""

    def __init__(self, n: int, kl: int, ku: int, nrhs: int):
        self.n, self.kl, self.ku, self.nrhs = n, kl, ku, nrhs
        nb = C.c_size_t()
        _lib.call("bh_dsgbsv_workspace", n, kl, ku, nrhs, C.byref(nb))
        self.lwork = nb.value
        self.work = DeviceArray((max(1, self.lwork),), np.uint8)
        self.ipiv = DeviceArray((max(1, n),), np.int64)

    def solve(self, A: BandedMatrix, B: DeviceArray, X: DeviceArray | None = None, want_pivots=False) -> MixedResult:
        n, kl, ku = self.n, self.kl, self.ku
        if A.shape != (n, n) or A.bandwidths() != (kl, ku) or A.dtype != np.float64:
            raise DimensionMismatch(f"solver set up for {n}x{n} ({kl},{ku}) Float64, got {A}")
        nrhs = 1 if len(B.shape) == 1 else B.shape[1]
        if B.shape[0] != n or nrhs != self.nrhs:
            raise DimensionMismatch(f"B is {B.shape}")
        if X is None:
            X = DeviceArray(B.shape, np.float64)
        iters, info = C.c_int64(), C.c_int64()
        _lib.call("bh_dsgbsv", n, kl, ku, nrhs, A.data.ptr, A.data.ld, B.ptr, max(1, n), X.ptr, max(1, n),
                  self.work.ptr, self.lwork, self.ipiv.ptr, C.byref(iters), C.byref(info))
        piv = self.ipiv.numpy()[:n] if want_pivots else np.zeros(0, np.int64)
        return MixedResult(X, int(iters.value), int(info.value), piv)


def solve_mixed(A: BandedMatrix, B: DeviceArray) -> MixedResult:
    n = A.shape[0]
    if A.shape[0] != A.shape[1]:
        raise DimensionMismatch(f"matrix is not square: dimensions are {A.shape}")
    nrhs = 1 if len(B.shape) == 1 else B.shape[1]
    return MixedSolver(n, A.l, A.u, nrhs).solve(A, B, want_pivots=True)
