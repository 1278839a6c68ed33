"""Triangular band solve and product (tbsv!/tbmv!, src/blas.jl:71-147, as
src/tribanded.jl:47-54,75-96 reaches them).

CPU: the oracle against scipy's OpenBLAS ?tbsv / ?tbmv, the routines the
reference's ccall binds:
  solve 'N', every uplo/diag: bit for bit;
  solve 'T': 1e-13 (f64) / 2e-5 (f32) relative (OpenBLAS blocks its dots);
  product: exactly on integer data (every summation order is exact there) and
  1e-14 relative on random data (OpenBLAS's order is not the oracle's);
  plus a dense triangular solve / product and planted operand mistakes that
  the pins must catch.
GPU (C ABI): solve 'N' and the product bit for bit against the oracle, solve
'T' to the same tolerance, every uplo/trans/diag, bandwidths on both sides of
the register-window limits (64, 128, 256) and past them (the workgroup path),
strided x, triangles that are views into a BandedMatrix, and a 2^20 round
trip.  Unit-diagonal data carries NaN on the diagonal and NaN in the corner
slots outside the triangle: neither may be read.
"""
import ctypes as C
import itertools

import numpy as np
import pytest
import scipy.linalg.blas as sblas

from _util import bits_equal

UPLO = ["U", "L"]
TRANS = ["N", "T"]
DIAG = ["N", "U"]


def tri_band(rng, n, k, uplo, diag, dtype=np.float64, integer=False):
    """(k+1) x n BLAS triangular band storage; corner slots outside the matrix
    and (for a unit diagonal) the diagonal row hold NaN."""
    if integer:
        a = rng.integers(-3, 4, size=(k + 1, n)).astype(dtype)
    else:
        a = (rng.uniform(-1.0, 1.0, size=(k + 1, n)) * (0.5 / max(k, 1))).astype(dtype)
    a = np.asfortranarray(a)
    dr = k if uplo == "U" else 0
    if diag == "U":
        a[dr, :] = np.nan
    elif not integer:
        a[dr, :] = (1.0 + rng.random(n)).astype(dtype) * np.where(rng.random(n) < 0.5, -1, 1)
    for j in range(n):
        for t in range(1, k + 1):
            if uplo == "U" and j - t < 0:
                a[k - t, j] = np.nan
            if uplo == "L" and j + t >= n:
                a[t, j] = np.nan
    return a


def dense_tri(a, n, k, uplo, diag):
    M = np.zeros((n, n), a.dtype)
    for j in range(n):
        for t in range(0, k + 1):
            i = j - t if uplo == "U" else j + t
            if 0 <= i < n:
                M[i, j] = a[k - t, j] if uplo == "U" else a[t, j]
    if diag == "U":
        np.fill_diagonal(M, 1.0)
    return M


def _scipy(fn, a, x, k, uplo, trans, diag):
    f = getattr(sblas, ("d" if a.dtype == np.float64 else "s") + fn)
    return f(k, np.nan_to_num(a, nan=0.0) if diag == "N" else _unit_clean(a, k, uplo), x.copy(),
             lower=int(uplo == "L"), trans=int(trans == "T"), diag=int(diag == "U"))


def _unit_clean(a, k, uplo):
    # OpenBLAS never reads the unit diagonal either; keep its NaN there, zero the corners
    b = np.nan_to_num(a, nan=0.0)
    b[k if uplo == "U" else 0, :] = np.nan
    return np.asfortranarray(b)


SIZES = [(50, 3), (70, 7), (40, 17), (200, 31), (30, 0), (10, 12)]


# --------------------------------------------------------------- oracle (CPU)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo,diag", list(itertools.product(UPLO, DIAG)))
def test_oracle_tbsv_N_bitexact_vs_openblas(oracle, dtype, uplo, diag):
    rng = np.random.default_rng(3)
    for n, k in SIZES:
        a = tri_band(rng, n, k, uplo, diag, dtype)
        x = rng.standard_normal(n).astype(dtype)
        ref = _scipy("tbsv", a, x, k, uplo, "N", diag)
        assert bits_equal(oracle.tbsv(uplo, "N", diag, n, k, a, x), ref), (n, k)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo,diag", list(itertools.product(UPLO, DIAG)))
def test_oracle_tbsv_T_vs_openblas(oracle, dtype, uplo, diag):
    rng = np.random.default_rng(4)
    tol = 1e-13 if dtype == np.float64 else 2e-5
    for n, k in SIZES:
        a = tri_band(rng, n, k, uplo, diag, dtype)
        x = rng.standard_normal(n).astype(dtype)
        ref = _scipy("tbsv", a, x, k, uplo, "T", diag)
        got = oracle.tbsv(uplo, "T", diag, n, k, a, x)
        assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (n, k)


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product(UPLO, TRANS, DIAG)))
def test_oracle_tbmv_exact_on_integers(oracle, uplo, trans, diag):
    rng = np.random.default_rng(5)
    for n, k in SIZES:
        a = tri_band(rng, n, k, uplo, diag, integer=True)
        x = rng.integers(-5, 6, size=n).astype(np.float64)
        ref = _scipy("tbmv", a, x, k, uplo, trans, diag)
        assert np.array_equal(oracle.tbmv(uplo, trans, diag, n, k, a, x), ref), (n, k)
        M = dense_tri(np.nan_to_num(a, nan=0.0), n, k, uplo, diag)
        assert np.array_equal(ref, (M.T if trans == "T" else M) @ x)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product(UPLO, TRANS, DIAG)))
def test_oracle_tbmv_vs_openblas(oracle, dtype, uplo, trans, diag):
    rng = np.random.default_rng(6)
    tol = 1e-14 if dtype == np.float64 else 5e-6
    for n, k in SIZES:
        a = tri_band(rng, n, k, uplo, diag, dtype)
        x = rng.standard_normal(n).astype(dtype)
        ref = _scipy("tbmv", a, x, k, uplo, trans, diag)
        got = oracle.tbmv(uplo, trans, diag, n, k, a, x)
        assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max()), (n, k)


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product(UPLO, TRANS, DIAG)))
def test_oracle_tb_dense_and_round_trip(oracle, uplo, trans, diag):
    rng = np.random.default_rng(7)
    n, k = 120, 9
    a = tri_band(rng, n, k, uplo, diag)
    M = dense_tri(np.nan_to_num(a, nan=0.0), n, k, uplo, diag)
    op = M.T if trans == "T" else M
    x = rng.standard_normal(n)
    assert np.allclose(oracle.tbsv(uplo, trans, diag, n, k, a, x), np.linalg.solve(op, x), rtol=1e-12, atol=1e-13)
    assert np.allclose(oracle.tbmv(uplo, trans, diag, n, k, a, x), op @ x, rtol=1e-13, atol=1e-14)
    y = oracle.tbmv(uplo, trans, diag, n, k, a, x)
    assert np.allclose(oracle.tbsv(uplo, trans, diag, n, k, a, y), x, rtol=1e-12, atol=1e-13)


def test_pins_catch_planted_operand_mistakes(oracle):
    """A swapped triangle, a dropped transpose or a wrongly read diagonal each
    fail the OpenBLAS comparison the oracle passes."""
    rng = np.random.default_rng(8)
    n, k = 60, 5
    a = tri_band(rng, n, k, "U", "N")
    x = rng.standard_normal(n)
    ref = _scipy("tbsv", a, x, k, "U", "N", "N")
    assert bits_equal(oracle.tbsv("U", "N", "N", n, k, a, x), ref)
    assert not np.allclose(oracle.tbsv("U", "T", "N", n, k, a, x), ref)
    assert not np.allclose(oracle.tbsv("U", "N", "U", n, k, a, x), ref)
    refL = _scipy("tbsv", a, x, k, "L", "N", "N")
    assert not np.allclose(oracle.tbsv("U", "N", "N", n, k, a, x), refL)


def test_argument_errors_blas_style():
    """BLAS xerbla numbering, checked on the host before any device work."""
    from banded_hip import _lib

    L = _lib.load()
    for fn in (L.bh_dtbsv, L.bh_dtbmv, L.bh_stbsv, L.bh_stbmv):
        assert fn(b"X", b"N", b"N", 4, 1, None, 2, None, 1) == -1
        assert fn(b"U", b"Q", b"N", 4, 1, None, 2, None, 1) == -2
        assert fn(b"U", b"N", b"Z", 4, 1, None, 2, None, 1) == -3
        assert fn(b"U", b"N", b"N", -1, 1, None, 2, None, 1) == -4
        assert fn(b"U", b"N", b"N", 4, -1, None, 2, None, 1) == -5
        assert fn(b"U", b"N", b"N", 4, 3, None, 3, None, 1) == -7
        assert fn(b"U", b"N", b"N", 4, 1, None, 2, None, 0) == -9
        assert fn(b"L", b"T", b"U", 0, 1, None, 2, None, 1) == 0  # empty: nothing to do
    for fn in (L.bh_dtbsv_multi, L.bh_stbsv_multi):
        assert fn(b"U", b"N", b"N", 4, 1, None, 2, -1, None, 4) == -8
        assert fn(b"U", b"N", b"N", 4, 1, None, 2, 2, None, 3) == -10
        assert fn(b"U", b"N", b"N", 4, 1, None, 2, 0, None, 4) == 0


# ------------------------------------------------------------------ GPU (C ABI)
GPU_SIZES = [(1, 0), (7, 3), (64, 1), (65, 64), (300, 65), (1000, 128), (777, 200), (2000, 256), (600, 300),
             (130, 1000)]


def _gpu_tb(bh, fn, uplo, trans, diag, n, k, a, x, incx=1):
    A = bh.DeviceArray.from_numpy(a)
    if incx == 1:
        X = bh.DeviceArray.from_numpy(x)
    else:
        buf = np.full(((n - 1) * incx + 1,), np.nan, x.dtype)
        buf[::incx] = x
        X = bh.DeviceArray.from_numpy(buf)
    pre = "d" if a.dtype == np.float64 else "s"
    bh._lib.call(f"bh_{pre}{fn}", uplo.encode(), trans.encode(), diag.encode(), n, k, A.ptr, A.ld, X.ptr, incx)
    out = X.numpy()
    if incx != 1:
        assert np.isnan(np.delete(out, np.arange(0, len(out), incx))).all(), "gaps of a strided x were written"
        out = out[::incx]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product(UPLO, TRANS, DIAG)))
def test_gpu_tbsv_matches_oracle(bh, oracle, dtype, uplo, trans, diag):
    rng = np.random.default_rng(11)
    tol = 1e-13 if dtype == np.float64 else 2e-5
    for n, k in GPU_SIZES:
        a = tri_band(rng, n, k, uplo, diag, dtype)
        x = rng.standard_normal(n).astype(dtype)
        print(f"tbsv {uplo}{trans}{diag} {np.dtype(dtype).name} n={n} k={k}", flush=True)
        got = _gpu_tb(bh, "tbsv", uplo, trans, diag, n, k, a, x)
        ref = oracle.tbsv(uplo, trans, diag, n, k, a, x)
        if trans == "N":
            assert bits_equal(got, ref), (n, k)
        else:
            assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (n, k)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product(UPLO, TRANS, DIAG)))
def test_gpu_tbmv_bitexact(bh, oracle, dtype, uplo, trans, diag):
    rng = np.random.default_rng(12)
    for n, k in GPU_SIZES:
        a = tri_band(rng, n, k, uplo, diag, dtype)
        x = rng.standard_normal(n).astype(dtype)
        print(f"tbmv {uplo}{trans}{diag} {np.dtype(dtype).name} n={n} k={k}", flush=True)
        got = _gpu_tb(bh, "tbmv", uplo, trans, diag, n, k, a, x)
        assert bits_equal(got, oracle.tbmv(uplo, trans, diag, n, k, a, x)), (n, k)


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["tbsv", "tbmv"])
def test_gpu_tb_strided_x(bh, oracle, fn):
    rng = np.random.default_rng(13)
    for (uplo, trans, diag), (n, k) in zip(itertools.product(UPLO, TRANS, DIAG),
                                           [(90, 4), (500, 70), (129, 128), (64, 0), (333, 200), (70, 3),
                                            (1000, 60), (40, 300)]):
        a = tri_band(rng, n, k, uplo, diag)
        x = rng.standard_normal(n)
        got = _gpu_tb(bh, fn, uplo, trans, diag, n, k, a, x, incx=3)
        ref = getattr(oracle, fn)(uplo, trans, diag, n, k, a, x)
        if fn == "tbsv" and trans == "T":
            assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
        else:
            assert bits_equal(got, ref), (uplo, trans, diag, n, k)


@pytest.mark.gpu
@pytest.mark.parametrize("cls", ["UpperTriangular", "LowerTriangular", "UnitUpperTriangular",
                                 "UnitLowerTriangular"])
def test_gpu_triangular_views_of_banded_matrix(bh, oracle, cls):
    """ldiv!/lmul! on a triangle of a BandedMatrix pass a pointer into the
    parent's band (rows 1:u+1 or u+1:l+u+1, src/tribanded.jl:19-26)."""
    rng = np.random.default_rng(14)
    n, l, u = 700, 5, 9
    D = np.asfortranarray(rng.uniform(-0.05, 0.05, size=(l + u + 1, n)))
    D[u, :] = 1.0 + rng.random(n)
    A = bh.BandedMatrix.from_host(D, n, l, u)
    T = getattr(bh, cls)(A)
    upper = cls.endswith("UpperTriangular")
    unit = cls.startswith("Unit")
    k = u if upper else l
    tri = np.asfortranarray(D[: u + 1] if upper else D[u:])
    x = rng.standard_normal(n)
    uplo, diag = ("U" if upper else "L"), ("U" if unit else "N")
    got = bh.ldiv_(T, bh.DeviceArray.from_numpy(x)).numpy()
    assert bits_equal(got, oracle.tbsv(uplo, "N", diag, n, k, tri, x))
    dense = T.to_dense()
    assert np.allclose(got, np.linalg.solve(dense, x), rtol=1e-12, atol=1e-12)
    y = (T @ bh.DeviceArray.from_numpy(x)).numpy()
    assert bits_equal(y, oracle.tbmv(uplo, "N", diag, n, k, tri, x))
    assert np.allclose(y, dense @ x, rtol=1e-13, atol=1e-13)
    X = rng.standard_normal((n, 3))
    got = T.ldiv_(bh.DeviceArray.from_numpy(X), trans="T").numpy()
    assert np.allclose(got, np.linalg.solve(dense.T, X), rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,l,u,nrhs", [(20000, 3, 8, 1), (5000, 0, 40, 3), (3000, 2, 63, 1), (100, 1, 5, 2),
                                        (1, 0, 0, 1)])
def test_gpu_triangular_partitioned_solve(bh, oracle, dtype, n, l, u, nrhs):
    """ldiv_(UpperTriangular(A), x, method="partitioned") -- the partitioned
    band solve with kl = 0 and no interchanges -- equals tbsv to rounding for
    'N' and 'T' (max-norm relative 1e-12 FP64 / 1e-4 FP32 against the oracle's
    tbsv on a well-conditioned triangle); the sub-diagonal rows of the parent
    (NaN here) are never read; lower and unit-diagonal triangles and k > 63
    fall back to the exact tbsv (bit for bit)."""
    rng = np.random.default_rng(n + u)
    D = np.asfortranarray(rng.uniform(-0.5 / max(u, 1), 0.5 / max(u, 1), size=(l + u + 1, n)))
    D[u, :] = 1.0 + rng.random(n)
    D[u + 1:, :] = np.nan  # the lower triangle must not be touched
    Dd = D.astype(dtype)
    A = bh.BandedMatrix.from_host(Dd, n, l, u)
    T = bh.UpperTriangular(A)
    tri = np.asfortranarray(Dd[: u + 1].astype(np.float64))
    B = rng.standard_normal((n, nrhs)).astype(dtype)
    tol = 1e-12 if dtype == np.float64 else 1e-4
    for trans in ("N", "T"):
        xb = bh.DeviceArray.from_numpy(np.asfortranarray(B if nrhs > 1 else B[:, 0]))
        got = T.ldiv_(xb, trans=trans, method="partitioned").numpy().reshape(n, nrhs).astype(np.float64)
        for c in range(nrhs):
            ref = oracle.tbsv("U", trans, "N", n, u, tri, B[:, c].astype(np.float64))
            assert np.abs(got[:, c] - ref).max() <= tol * np.abs(ref).max(), (trans, c)
    # fallbacks: bit for bit with the exact solve
    L = bh.LowerTriangular(bh.BandedMatrix.from_host(np.nan_to_num(Dd), n, l, u))
    x = rng.standard_normal(n).astype(dtype)
    a = L.ldiv_(bh.DeviceArray.from_numpy(x), method="partitioned").numpy()
    b = L.ldiv_(bh.DeviceArray.from_numpy(x)).numpy()
    assert bits_equal(a, b)
    with pytest.raises(ValueError):
        T.ldiv_(bh.DeviceArray.from_numpy(x), method="nope")


@pytest.mark.gpu
def test_gpu_triangular_partitioned_wide_band_falls_back(bh):
    """k = 64 is past the partitioned solve's range: the exact tbsv runs."""
    rng = np.random.default_rng(3)
    n, u = 3000, 64
    D = np.asfortranarray(rng.uniform(-0.01, 0.01, size=(u + 1, n)))
    D[u, :] = 1.0 + rng.random(n)
    T = bh.UpperTriangular(bh.BandedMatrix.from_host(D, n, 0, u))
    x = rng.standard_normal(n)
    a = T.ldiv_(bh.DeviceArray.from_numpy(x), method="partitioned").numpy()
    b = T.ldiv_(bh.DeviceArray.from_numpy(x)).numpy()
    assert bits_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product(UPLO, TRANS, DIAG)))
def test_gpu_tbsv_multi_equals_single(bh, oracle, uplo, trans, diag):
    """All columns in one launch: each equals the single-vector call bit for bit
    (and the oracle, bitwise for 'N'), with ldx > n."""
    rng = np.random.default_rng(16)
    for n, k, nrhs in [(257, 9, 5), (1000, 130, 3), (90, 400, 2)]:
        a = tri_band(rng, n, k, uplo, diag)
        X = rng.standard_normal((n + 3, nrhs))
        A = bh.DeviceArray.from_numpy(a)
        DX = bh.DeviceArray.from_numpy(X)
        bh._lib.call("bh_dtbsv_multi", uplo.encode(), trans.encode(), diag.encode(), n, k, A.ptr, A.ld, nrhs,
                     DX.ptr, n + 3)
        got = DX.numpy()
        assert bits_equal(got[n:], X[n:]), "rows below n were written"
        for c in range(nrhs):
            single = _gpu_tb(bh, "tbsv", uplo, trans, diag, n, k, a, X[:n, c])
            assert bits_equal(got[:n, c], single), (n, k, c)
            if trans == "N":
                assert bits_equal(single, oracle.tbsv(uplo, trans, diag, n, k, a, X[:n, c]))


@pytest.mark.gpu
def test_gpu_tb_shape_errors(bh):
    A = bh.BandedMatrix.from_host(np.ones((3, 10)), 10, 1, 1)
    with pytest.raises(bh.DimensionMismatch):
        bh.ldiv_(bh.UpperTriangular(A), bh.DeviceArray.from_numpy(np.ones(9)))
    with pytest.raises(ValueError):
        bh.tbsv_("U", "N", "N", 10, 3, A.bandeddata(), bh.DeviceArray.from_numpy(np.ones(10)))


@pytest.mark.gpu
@pytest.mark.parametrize("uplo,trans", list(itertools.product(UPLO, TRANS)))
def test_gpu_tb_round_trip_2p20(bh, uplo, trans):
    """n = 2^20, k = 64: tbsv(tbmv(x)) returns x (a size-independent property)."""
    rng = np.random.default_rng(15)
    n, k = 1 << 20, 64
    a = np.asfortranarray(rng.uniform(-1.0, 1.0, size=(k + 1, n)) * (0.5 / k))
    a[k if uplo == "U" else 0, :] = 1.0 + rng.random(n)
    x = rng.standard_normal(n)
    A = bh.DeviceArray.from_numpy(a)
    X = bh.DeviceArray.from_numpy(x)
    bh.tbmv_(uplo, trans, "N", n, k, A, X)
    bh.tbsv_(uplo, trans, "N", n, k, A, X)
    assert np.abs(X.numpy() - x).max() <= 1e-12 * np.abs(x).max()
