"""banded_hip: the MI355X (gfx950) banded hot path of BandedMatrices.jl.

The arithmetic runs in libbandedhip.so (HIP kernels behind the C ABI in
include/bandedhip.h); this package is the host-side mirror of the reference's
operator interface for the path: BandedMatrix storage, mul!/A*x, A*B, lu,
ldiv!/\\, qr and the triangular band solve/product.
"""
from ._lib import BandedHIPError, load as load_library, LIB_PATH
from .device import DeviceArray, Timer, sync
from .bandedmatrix import (BandedMatrix, BandError, DimensionMismatch, _BandedMatrix, axpy_, brand, convert, copyto_,
                           gbmm, mul, prodbandwidths)
from .lu import BandedLU, SingularException, TransposeFact, ldiv_, lu, lu_, pad_for_lu, solve
from .mixed import MixedResult, MixedSolver, solve_mixed
from .adjoint import band_outer_, loss_and_grad, mul_pullback, solve_pullback
from .batch import BatchedGBSV, BatchGraph, c4_batch
from .qr import BandedQR, QRPackedQ, qr, qr_, qr_ldiv_, qr_lmul_, qr_rmul_, qr_solve
from .symbanded import (BandedCholesky, BandedLDLt, PosDefException, Symmetric, cholesky, cholesky_ldiv_, ldlt,
                        ldlt_ldiv_)
from .tribanded import (BandView, LowerTriangular, UnitLowerTriangular, UnitUpperTriangular, UpperTriangular, tbmv_,
                        tbsv_)

__all__ = [
    "BandedHIPError", "load_library", "LIB_PATH", "DeviceArray", "Timer", "sync", "BandedMatrix", "BandError",
    "DimensionMismatch", "_BandedMatrix", "brand", "gbmm", "mul", "prodbandwidths", "BandedLU",
    "SingularException", "TransposeFact", "ldiv_", "lu", "lu_", "pad_for_lu", "solve", "MixedResult",
    "MixedSolver", "solve_mixed", "BandedQR", "QRPackedQ", "qr", "qr_", "qr_ldiv_", "qr_lmul_", "qr_rmul_", "qr_solve",
    "BatchedGBSV", "BatchGraph", "c4_batch", "band_outer_", "loss_and_grad", "mul_pullback", "solve_pullback",
    "axpy_", "convert", "copyto_", "Symmetric", "BandedCholesky", "PosDefException", "cholesky", "cholesky_ldiv_", "BandedLDLt", "ldlt", "ldlt_ldiv_",
    "BandView", "UpperTriangular", "LowerTriangular", "UnitUpperTriangular", "UnitLowerTriangular", "tbsv_", "tbmv_",
]
