"""Reverse-mode rules for the banded hot path (no reference counterpart: the
reference defines no rrule; its transposed solve src/banded/linalg.jl:41-47 is
the adjoint solve reused here).

  solve_pullback(F, x, xbar)  for x = A \\ b:  bbar = A^{-T} xbar,  Abar = -P(bbar x')
  mul_pullback(A, x0, ybar)   for y = A x0:    Abar = P(ybar x0'),   x0bar = A' ybar
  loss_and_grad(A, b, w, x0)  L = w' A^{-1} b + 1/2 ||A x0||^2 and (dL/dA band, dL/db)
P restricts an outer product to A's band (band storage, like A.data).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .bandedmatrix import BandedMatrix, _prefix, mul
from .device import DeviceArray
from .lu import BandedLU, ldiv_, lu


def band_outer_(G: BandedMatrix, alpha1, u1, v1, alpha2=0.0, u2=None, v2=None, beta=1.0) -> BandedMatrix:
    """G <- beta G + P(alpha1 u1 v1' + alpha2 u2 v2') on G's band, in place."""
    m, n = G.shape
    p = _prefix(G.dtype)
    _lib.call(f"bh_{p}gb_band_outer", m, n, G.l, G.u, alpha1, u1.ptr if u1 is not None else None,
              v1.ptr if v1 is not None else None, alpha2, u2.ptr if u2 is not None else None,
              v2.ptr if v2 is not None else None, beta, G.data.ptr, G.data.ld)
    return G


def solve_pullback(F: BandedLU, x: DeviceArray, xbar: DeviceArray, bands):
    """Pullback of x = A \\ b (A square with bandwidths `bands`): (Abar, bbar)."""
    bbar = ldiv_(F.transpose(), xbar.copy())
    n = F.shape[0]
    l, u = bands
    Abar = BandedMatrix.zeros((n, n), (l, u), F.factors.dtype)
    band_outer_(Abar, -1.0, bbar, x, beta=0.0)
    return Abar, bbar


def mul_pullback(A: BandedMatrix, x0: DeviceArray, ybar: DeviceArray):
    """Pullback of y = A x0: (Abar, x0bar)."""
    Abar = BandedMatrix.zeros(A.shape, A.bandwidths(), A.dtype)
    band_outer_(Abar, 1.0, ybar, x0, beta=0.0)
    x0bar = DeviceArray((A.shape[1],), A.dtype)
    mul(x0bar, A, ybar, trans="T")
    return Abar, x0bar


def loss_and_grad(A: BandedMatrix, b: DeviceArray, w: DeviceArray, x0: DeviceArray, method: str = "auto"):
    """L = w' A^{-1} b + 1/2 ||A x0||^2 with its gradient: returns (L, dA, db, x).
    method: the two solves' method (ldiv_'s "auto" -- partitioned from n = 4096
    on --, "exact" or "partitioned").

    Forward: band LU, 'N' solve, gbmv.  Backward: 'T' solve with the same
    factors and pivots, then one fused band kernel
    dA_ij = fma(y0_i, x0_j, -(lambda_i x_j))."""
    n = A.shape[0]
    F = lu(A)
    x = ldiv_(F, b.copy(), method)
    y0 = DeviceArray((n,), A.dtype)
    mul(y0, A, x0)
    lam = ldiv_(F.transpose(), w.copy(), method)
    dA = BandedMatrix.zeros((n, n), A.bandwidths(), A.dtype)
    _lib.call(f"bh_{_prefix(A.dtype)}gb_band_grad", n, n, A.l, A.u, lam.ptr, x.ptr, y0.ptr, x0.ptr, dA.data.ptr,
              dA.data.ld)
    xh, wh, yh = x.numpy(), w.numpy(), y0.numpy()
    L = float(np.dot(wh, xh) + 0.5 * np.dot(yh, yh))
    return L, dA, lam, x
