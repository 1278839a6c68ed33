"""ctypes front end of the CPU oracle (oracle/bandoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``, always as the checker.  The
product package never imports this module.

Every function copies its inputs (numpy, Fortran order) and returns fresh
arrays, so the oracle never aliases the arrays under test.  Band storage is
the BandedMatrices.jl layout (src/banded/BandedMatrix.jl:16-28): a
(kl+ku+1) x n column-major array with A[i,j] at data[ku+i-j, j].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int64)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the recipe in oracle/Makefile."""
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(
        os.path.join(_HERE, "bandoracle.c")
    ):
        cmd = ["make", "-s"] + (["-B"] if force else []) + ["liboracle.so"]
        subprocess.run(cmd, cwd=_HERE, check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        for pre, fp, ft in (("d", _dp, C.c_double), ("s", _fp, C.c_float)):
            f = getattr(L, f"bo_{pre}gbmv")
            f.argtypes = [C.c_char, i64, i64, i64, i64, ft, fp, i64, fp, i64, ft, fp, i64]
            f.restype = None
            f = getattr(L, f"bo_{pre}gbtf2")
            f.argtypes = [i64, i64, i64, i64, fp, i64, _ip]
            f.restype = i64
            f = getattr(L, f"bo_{pre}gbtrs")
            f.argtypes = [C.c_char, i64, i64, i64, i64, fp, i64, _ip, fp, i64]
            f.restype = None
            f = getattr(L, f"bo_{pre}tbsv_u")
            f.argtypes = [C.c_char, i64, i64, fp, i64, fp, i64]
            f.restype = None
            for nm in ("tbsv", "tbmv"):
                f = getattr(L, f"bo_{pre}{nm}")
                f.argtypes = [C.c_char, C.c_char, C.c_char, i64, i64, fp, i64, fp, i64]
                f.restype = None
            f = getattr(L, f"bo_{pre}pbtf2")
            f.argtypes = [C.c_char, i64, i64, fp, i64]
            f.restype = i64
            f = getattr(L, f"bo_{pre}sbldlt")
            f.argtypes = [C.c_char, i64, i64, fp, i64]
            f.restype = None
            f = getattr(L, f"bo_{pre}sbmv")
            f.argtypes = [C.c_char, i64, i64, ft, fp, i64, fp, ft, fp]
            f.restype = None
        L.bo_dsbldlts.argtypes = [C.c_char, i64, i64, _dp, i64, _dp]
        L.bo_dsbldlts.restype = None
        L.bo_dgbtf2_planted.argtypes = [C.c_int, i64, i64, i64, i64, _dp, i64, _ip]
        L.bo_dgbtf2_planted.restype = i64
        L.bo_dgbmm.argtypes = [i64, i64, i64, C.c_double, _dp, i64, i64, i64, _dp, i64, i64, i64,
                               C.c_double, _dp, i64, i64, i64]
        L.bo_dgbmm.restype = C.c_int
        L.bo_dgbqrf.argtypes = [i64, i64, i64, i64, _dp, i64, _dp]
        L.bo_dgbqrf.restype = None
        for nm in ("bo_dgbqr_qtmul", "bo_dgbqr_qmul"):
            getattr(L, nm).argtypes = [i64, i64, i64, i64, _dp, i64, _dp, i64, i64, _dp, i64]
            getattr(L, nm).restype = None
        L.bo_dgbqr_rmul.argtypes = [C.c_char, i64, i64, i64, i64, i64, _dp, i64, _dp, _dp, i64]
        L.bo_dgbqr_rmul.restype = None
        L.bo_dgbqrs.argtypes = [i64, i64, i64, i64, _dp, i64, _dp, i64, _dp, i64]
        L.bo_dgbqrs.restype = None
        L.bo_dgb_band_grad.argtypes = [i64, i64, i64, _dp, _dp, _dp, _dp, _dp, i64]
        L.bo_dgb_band_grad.restype = None
        _lib = L
    return _lib


def _pre(dt):
    dt = np.dtype(dt)
    if dt == np.float64:
        return "d", _dp
    if dt == np.float32:
        return "s", _fp
    raise TypeError(f"oracle supports float64/float32, got {dt}")


def _f(a, dt=None):
    return np.array(a, dtype=dt if dt is not None else a.dtype, order="F", copy=True)


def _ptr(a, p):
    return a.ctypes.data_as(p)


def gbmv(trans, m, n, kl, ku, alpha, data, x, beta=0.0, y=None):
    """y <- alpha*op(A)*x + beta*y (OpenBLAS ?gbmv restated)."""
    pre, P = _pre(data.dtype)
    a = _f(data)
    xx = np.ascontiguousarray(x, dtype=a.dtype).copy()
    ylen = m if trans in ("N", "n") else n
    yy = np.zeros(ylen, a.dtype) if y is None else np.ascontiguousarray(y, dtype=a.dtype).copy()
    getattr(lib(), f"bo_{pre}gbmv")(trans.encode(), m, n, kl, ku, alpha, _ptr(a, P), a.shape[0],
                                    _ptr(xx, P), 1, beta, _ptr(yy, P), 1)
    return yy


def gbmm(n, nu, m, alpha, A, Al, Au, B, Bl, Bu, beta, Cdata, Cl, Cu):
    """gbmm! restated (float64).  Returns (C, status); status -1 is BandError."""
    a, b, c = _f(A, np.float64), _f(B, np.float64), _f(Cdata, np.float64)
    st = lib().bo_dgbmm(n, nu, m, alpha, _ptr(a, _dp), max(a.shape[0], 1), Al, Au,
                        _ptr(b, _dp), max(b.shape[0], 1), Bl, Bu, beta,
                        _ptr(c, _dp), max(c.shape[0], 1), Cl, Cu)
    return c, st


def gbtf2(m, n, kl, ku, ab):
    """Unblocked band LU (LAPACK ?gbtf2 restated).  ab is (2kl+ku+1) x n.
    Returns (factors, ipiv 1-based int64, info)."""
    pre, P = _pre(ab.dtype)
    f = _f(ab)
    ipiv = np.zeros(max(min(m, n), 0), np.int64)
    info = getattr(lib(), f"bo_{pre}gbtf2")(m, n, kl, ku, _ptr(f, P), f.shape[0], _ptr(ipiv, _ip))
    return f, ipiv, int(info)


def gbtf2_planted(variant, m, n, kl, ku, ab):
    """gbtf2 with a planted mistake (variant 1: divide, 2: unfused update);
    used only to show that the LAPACK bit pin is discriminating."""
    f = _f(ab, np.float64)
    ipiv = np.zeros(max(min(m, n), 0), np.int64)
    info = lib().bo_dgbtf2_planted(variant, m, n, kl, ku, _ptr(f, _dp), f.shape[0], _ptr(ipiv, _ip))
    return f, ipiv, int(info)


def gbtrs(trans, n, kl, ku, ab, ipiv, b):
    """Solve with gbtf2 factors (LAPACK ?gbtrs restated).  b is n or n x nrhs."""
    pre, P = _pre(ab.dtype)
    f = _f(ab)
    vec = b.ndim == 1
    bb = _f(b.reshape(n, -1), ab.dtype)
    piv = np.ascontiguousarray(ipiv, dtype=np.int64)
    getattr(lib(), f"bo_{pre}gbtrs")(trans.encode(), n, kl, ku, bb.shape[1], _ptr(f, P), f.shape[0],
                                     _ptr(piv, _ip), _ptr(bb, P), max(n, 1))
    return bb[:, 0].copy() if vec else bb


def tbsv_u(trans, n, k, a, x):
    pre, P = _pre(a.dtype)
    aa = _f(a)
    xx = np.ascontiguousarray(x, dtype=a.dtype).copy()
    getattr(lib(), f"bo_{pre}tbsv_u")(trans.encode(), n, k, _ptr(aa, P), aa.shape[0], _ptr(xx, P), 1)
    return xx


def tbsv(uplo, trans, diag, n, k, a, x):
    """x <- op(A)^{-1} x for the triangular band A stored in a ((k+1) x n, BLAS layout)."""
    pre, P = _pre(a.dtype)
    aa = _f(a)
    xx = np.ascontiguousarray(x, dtype=a.dtype).copy()
    getattr(lib(), f"bo_{pre}tbsv")(uplo.encode(), trans.encode(), diag.encode(), n, k, _ptr(aa, P), aa.shape[0],
                                    _ptr(xx, P), 1)
    return xx


def tbmv(uplo, trans, diag, n, k, a, x):
    """x <- op(A) x for the triangular band A stored in a ((k+1) x n, BLAS layout)."""
    pre, P = _pre(a.dtype)
    aa = _f(a)
    xx = np.ascontiguousarray(x, dtype=a.dtype).copy()
    getattr(lib(), f"bo_{pre}tbmv")(uplo.encode(), trans.encode(), diag.encode(), n, k, _ptr(aa, P), aa.shape[0],
                                    _ptr(xx, P), 1)
    return xx


def pbtf2(uplo, n, kd, ab):
    """Band Cholesky of the symmetric band ab ((kd+1) x n, LAPACK storage): (factor, info)."""
    pre, P = _pre(ab.dtype)
    f = _f(ab).copy(order="F")
    info = getattr(lib(), f"bo_{pre}pbtf2")(uplo.encode(), n, kd, _ptr(f, P), f.shape[0])
    return f, int(info)


def pbtrs(uplo, n, kd, f, b):
    """Solve with the band Cholesky factor: two triangular band solves, as xPBTRS does."""
    if uplo == "U":
        y = tbsv("U", "T", "N", n, kd, f, b)
        return tbsv("U", "N", "N", n, kd, f, y)
    y = tbsv("L", "N", "N", n, kd, f, b)
    return tbsv("L", "T", "N", n, kd, f, y)


def sbldlt(uplo, n, kd, ab):
    """Banded LDL' of the stored triangle (ldlt!): L's entries where A's were, d on the diagonal."""
    pre, P = _pre(ab.dtype)
    f = _f(ab).copy(order="F")
    getattr(lib(), f"bo_{pre}sbldlt")(uplo.encode(), n, kd, _ptr(f, P), f.shape[0])
    return f


def sbldlts(uplo, n, kd, f, b):
    """ldiv!(::LDLt, b) with the factor of sbldlt (FP64)."""
    ff = _f(f)
    bb = np.ascontiguousarray(b, dtype=np.float64).copy()
    lib().bo_dsbldlts(uplo.encode(), n, kd, _ptr(ff, _dp), ff.shape[0], _ptr(bb, _dp))
    return bb


def sbmv(uplo, n, k, alpha, ab, x, beta=0.0, y=None):
    pre, P = _pre(ab.dtype)
    aa = _f(ab)
    xx = np.ascontiguousarray(x, dtype=ab.dtype)
    yy = np.zeros(n, ab.dtype) if y is None else np.ascontiguousarray(y, dtype=ab.dtype).copy()
    getattr(lib(), f"bo_{pre}sbmv")(uplo.encode(), n, k, alpha, _ptr(aa, P), aa.shape[0], _ptr(xx, P), beta,
                                    _ptr(yy, P))
    return yy


def gbqrf(m, n, l, u, D):
    """_banded_qr! restated on the padded copy (bands (l, u) with u already
    including the l fill-in rows).  Returns (factors, tau)."""
    d = _f(D, np.float64)
    tau = np.zeros(min(m, n), np.float64)
    lib().bo_dgbqrf(m, n, l, u, _ptr(d, _dp), d.shape[0], _ptr(tau, _dp))
    return d, tau


def gbqr_qtmul(m, n, l, u, D, tau, B):
    d = _f(D, np.float64)
    t = np.ascontiguousarray(tau, np.float64)
    vec = B.ndim == 1
    bb = _f(B.reshape(m, -1), np.float64)
    lib().bo_dgbqr_qtmul(m, n, l, u, _ptr(d, _dp), d.shape[0], _ptr(t, _dp), len(t), bb.shape[1],
                         _ptr(bb, _dp), max(m, 1))
    return bb[:, 0].copy() if vec else bb


def gbqr_qmul(m, n, l, u, D, tau, B):
    d = _f(D, np.float64)
    t = np.ascontiguousarray(tau, np.float64)
    vec = B.ndim == 1
    bb = _f(B.reshape(m, -1), np.float64)
    lib().bo_dgbqr_qmul(m, n, l, u, _ptr(d, _dp), d.shape[0], _ptr(t, _dp), len(t), bb.shape[1],
                        _ptr(bb, _dp), max(m, 1))
    return bb[:, 0].copy() if vec else bb


def gbqr_rmul(trans, mQ, nQ, l, u, D, tau, A):
    """A (mA x mQ) <- A Q ('N') or A Q' ('T'), banded_qr_rmul! restated."""
    d = _f(D, np.float64)
    t = np.ascontiguousarray(tau, np.float64)
    aa = _f(A, np.float64).copy(order="F")
    lib().bo_dgbqr_rmul(trans.encode(), aa.shape[0], mQ, nQ, l, u, _ptr(d, _dp), d.shape[0], _ptr(t, _dp),
                        _ptr(aa, _dp), max(aa.shape[0], 1))
    return aa


def gbqrs(m, n, l, u, D, tau, B):
    """QR solve for m >= n; returns the first n rows of the overwritten B."""
    d = _f(D, np.float64)
    t = np.ascontiguousarray(tau, np.float64)
    vec = B.ndim == 1
    bb = _f(B.reshape(m, -1), np.float64)
    lib().bo_dgbqrs(m, n, l, u, _ptr(d, _dp), d.shape[0], _ptr(t, _dp), bb.shape[1], _ptr(bb, _dp),
                    max(m, 1))
    return bb[:n, 0].copy() if vec else bb[:n]


def band_grad(n, kl, ku, lam, x, y0, x0):
    g = np.zeros((kl + ku + 1, n), np.float64, order="F")
    args = [np.ascontiguousarray(v, np.float64) for v in (lam, x, y0, x0)]
    lib().bo_dgb_band_grad(n, kl, ku, *[_ptr(v, _dp) for v in args], _ptr(g, _dp), kl + ku + 1)
    return g


# ------------------------------------------------------------------ helpers
def band_to_dense(data, m, n, kl, ku):
    """Dense m x n matrix of a band array (entries outside the band are 0)."""
    M = np.zeros((m, n), dtype=data.dtype)
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            M[i, j] = data[ku + i - j, j]
    return M


def dense_to_band(M, kl, ku, fill=0.0):
    m, n = M.shape
    d = np.full((kl + ku + 1, n), fill, dtype=M.dtype, order="F")
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            d[ku + i - j, j] = M[i, j]
    return d


def pad_for_lu(data, m, n, kl, ku, dtype=None):
    """The padded copy BandedMatrix{T}(A, (l, l+u)) (BandedMatrix.jl:222-232):
    (2kl+ku+1) x n with kl zero fill-in rows on top (LAPACK AB, LDAB=2kl+ku+1)."""
    dt = dtype or data.dtype
    ab = np.zeros((2 * kl + ku + 1, n), dtype=dt, order="F")
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            ab[kl + ku + i - j, j] = data[ku + i - j, j]
    return ab
