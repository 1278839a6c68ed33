"""Symmetric band Cholesky (pbtrf!/pbtrs!, src/lapack.jl:268-333, as
cholesky(Symmetric(A)) and ldiv! reach them, src/symbanded/BandedCholesky.jl:1-79)
and the symmetric band product (BLAS.sbmv!, src/symbanded/symbanded.jl:72-91).

CPU: the oracle against scipy's OpenBLAS/LAPACK: dpbtrf factor and info bit for
bit (upper and lower, kd on both sides of LAPACK's 32 block switch, a
non-positive pivot, a NaN that LAPACK lets through); dpbtrs to 1e-14; dsbmv
exactly on integer data and to 1e-14 on random data; U'U (L L') reproduces A.
GPU (C ABI through the mirror): pbtrf bit for bit against the oracle (FP64 and
FP32, kd up to the LDS limit), pbtrs against the oracle to rounding and against
a dense solve, sbmv bit for bit against the oracle.
"""
import itertools

import numpy as np
import pytest
import scipy.linalg.blas as sblas
import scipy.linalg.lapack as slap

from _util import bits_equal


def spd_band(rng, n, kd, uplo, dtype=np.float64):
    """A diagonally dominant symmetric band in LAPACK storage; corner slots 0."""
    a = rng.uniform(-1.0, 1.0, size=(kd + 1, n)).astype(dtype)
    dr = kd if uplo == "U" else 0
    a[dr, :] = (kd + 1 + rng.random(n)).astype(dtype)
    for j in range(n):
        for t in range(1, kd + 1):
            if uplo == "U" and j - t < 0:
                a[kd - t, j] = 0
            if uplo == "L" and j + t >= n:
                a[t, j] = 0
    return np.asfortranarray(a)


def sym_dense(a, n, kd, uplo):
    M = np.zeros((n, n), a.dtype)
    for j in range(n):
        for t in range(kd + 1):
            i = j - t if uplo == "U" else j + t
            if 0 <= i < n:
                v = a[kd - t, j] if uplo == "U" else a[t, j]
                M[i, j] = M[j, i] = v
    return M


CPU_SIZES = [(50, 3), (70, 7), (100, 17), (200, 31), (300, 32), (300, 40), (10, 0)]


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_pbtf2_bitexact_vs_lapack(oracle, uplo, dtype):
    rng = np.random.default_rng(1)
    f_lap = slap.dpbtrf if dtype == np.float64 else slap.spbtrf
    for n, kd in CPU_SIZES:
        a = spd_band(rng, n, kd, uplo, dtype)
        c, info = f_lap(a.copy(order="F"), lower=int(uplo == "L"))
        f, inf2 = oracle.pbtf2(uplo, n, kd, a)
        assert info == inf2 == 0
        assert bits_equal(f, np.asfortranarray(c)), (n, kd)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_oracle_pbtf2_info_cases(oracle, uplo):
    rng = np.random.default_rng(2)
    n, kd = 20, 3
    a = spd_band(rng, n, kd, uplo)
    dr = kd if uplo == "U" else 0
    for bad, want in ((-1.0, 8), (0.0, 8), (np.nan, 0)):  # LAPACK lets a NaN pivot through
        b = a.copy(order="F")
        b[dr, 7] = bad
        c, info = slap.dpbtrf(b.copy(order="F"), lower=int(uplo == "L"))
        f, inf2 = oracle.pbtf2(uplo, n, kd, b)
        assert info == inf2 == want
        assert bits_equal(f, np.asfortranarray(c))


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_oracle_pbtrs_and_factor_identity(oracle, uplo):
    rng = np.random.default_rng(3)
    for n, kd in [(60, 4), (120, 33)]:
        a = spd_band(rng, n, kd, uplo)
        f, info = oracle.pbtf2(uplo, n, kd, a)
        b = rng.standard_normal(n)
        xs, _ = slap.dpbtrs(f, b, lower=int(uplo == "L"))
        xo = oracle.pbtrs(uplo, n, kd, f, b)
        assert np.abs(xo - xs).max() <= 1e-14 * np.abs(xs).max()
        M = sym_dense(a, n, kd, uplo)
        assert np.allclose(M @ xo, b, rtol=1e-12, atol=1e-12)
        T = np.zeros((n, n))
        for j in range(n):
            for t in range(kd + 1):
                i = j - t if uplo == "U" else j + t
                if 0 <= i < n:
                    T[i, j] = f[kd - t, j] if uplo == "U" else f[t, j]
        R = T.T @ T if uplo == "U" else T @ T.T
        assert np.allclose(R, M, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_oracle_sbmv_vs_blas(oracle, uplo):
    rng = np.random.default_rng(4)
    for n, k in [(40, 3), (90, 12), (33, 0)]:
        a = np.asfortranarray(rng.integers(-4, 5, size=(k + 1, n)).astype(np.float64))
        x = rng.integers(-4, 5, size=n).astype(np.float64)
        y = rng.integers(-4, 5, size=n).astype(np.float64)
        ref = sblas.dsbmv(k, 2.0, a, x, beta=3.0, y=y.copy(), lower=int(uplo == "L"))
        assert np.array_equal(oracle.sbmv(uplo, n, k, 2.0, a, x, 3.0, y), ref)
        assert np.array_equal(ref, 2.0 * (sym_dense(a, n, k, uplo) @ x) + 3.0 * y)
        a = spd_band(rng, n, k, uplo)
        x = rng.standard_normal(n)
        ref = sblas.dsbmv(k, 1.5, a, x, lower=int(uplo == "L"))
        assert np.abs(oracle.sbmv(uplo, n, k, 1.5, a, x) - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_oracle_ldlt_closed_form_and_identity(oracle, uplo):
    """[-1, 2, -1]: d_j = (j+2)/(j+1), l = -(j+1)/(j+2) (0-based); L D L' = A and
    the solve on a random diagonally dominant band."""
    n = 12
    a = np.zeros((2, n), order="F")
    if uplo == "U":
        a[1, :], a[0, 1:] = 2, -1
    else:
        a[0, :], a[1, :-1] = 2, -1
    f = oracle.sbldlt(uplo, n, 1, a)
    j = np.arange(n)
    d, lv = (f[1], f[0, 1:]) if uplo == "U" else (f[0], f[1, :-1])
    assert np.allclose(d, (j + 2) / (j + 1), rtol=1e-15)
    assert np.allclose(lv, -(j[:-1] + 1) / (j[:-1] + 2), rtol=1e-15)
    rng = np.random.default_rng(5)
    n, kd = 80, 6
    a = spd_band(rng, n, kd, uplo)
    f = oracle.sbldlt(uplo, n, kd, a)
    L = np.eye(n)
    D = np.zeros(n)
    for c in range(n):
        D[c] = f[kd, c] if uplo == "U" else f[0, c]
        for t in range(1, kd + 1):
            if c + t < n:
                L[c + t, c] = f[kd - t, c + t] if uplo == "U" else f[t, c]
    M = sym_dense(a, n, kd, uplo)
    assert np.allclose(L @ np.diag(D) @ L.T, M, rtol=1e-13, atol=1e-13)
    b = rng.standard_normal(n)
    assert np.allclose(oracle.sbldlts(uplo, n, kd, f, b), np.linalg.solve(M, b), rtol=1e-12, atol=1e-13)


def test_argument_errors():
    from banded_hip import _lib
    import ctypes

    L = _lib.load()
    info = ctypes.c_int64()
    assert L.bh_dpbtrf(b"X", 4, 1, None, 2, ctypes.byref(info)) == -1
    assert L.bh_dpbtrf(b"U", -1, 1, None, 2, ctypes.byref(info)) == -2
    assert L.bh_dpbtrf(b"U", 4, 2, None, 2, ctypes.byref(info)) == -5
    assert L.bh_dpbtrs(b"U", 4, 1, 1, None, 2, None, 3) == -8
    assert L.bh_dsbmv(b"U", 4, 1, 1.0, None, 2, None, 0, 0.0, None, 1) == -8
    assert L.bh_dsbldlt(b"U", 4, 3, None, 3) == -5
    assert L.bh_dsbldlts(b"L", 4, 1, 1, None, 2, None, 2) == -8


# ------------------------------------------------------------------ GPU
GPU_SIZES = [(1, 0), (10, 0), (50, 3), (300, 17), (1000, 31), (500, 64), (400, 100), (300, 127), (2000, 8), (100, 40)]


@pytest.mark.gpu
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gpu_pbtrf_bitexact(bh, oracle, uplo, dtype):
    rng = np.random.default_rng(11)
    import ctypes

    for n, kd in GPU_SIZES:
        if dtype == np.float32 and kd > 180:
            continue
        a = spd_band(rng, n, kd, uplo, dtype)
        print(f"pbtrf {uplo} {np.dtype(dtype).name} n={n} kd={kd}", flush=True)
        A = bh.DeviceArray.from_numpy(a)
        info = ctypes.c_int64(-1)
        pre = "d" if dtype == np.float64 else "s"
        bh._lib.call(f"bh_{pre}pbtrf", uplo.encode(), n, kd, A.ptr, A.ld, ctypes.byref(info))
        f, inf2 = oracle.pbtf2(uplo, n, kd, a)
        assert info.value == inf2 == 0
        assert bits_equal(A.numpy(), f), (n, kd)


@pytest.mark.gpu
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_gpu_pbtrf_info(bh, oracle, uplo):
    import ctypes

    rng = np.random.default_rng(12)
    n, kd = 300, 9
    a = spd_band(rng, n, kd, uplo)
    dr = kd if uplo == "U" else 0
    for col, bad in ((0, -2.0), (137, 0.0), (299, -1.0), (50, np.nan)):
        b = a.copy(order="F")
        b[dr, col] = bad
        A = bh.DeviceArray.from_numpy(b)
        info = ctypes.c_int64(-1)
        bh._lib.call("bh_dpbtrf", uplo.encode(), n, kd, A.ptr, A.ld, ctypes.byref(info))
        f, inf2 = oracle.pbtf2(uplo, n, kd, b)
        assert info.value == inf2
        got = A.numpy()
        # a NaN pivot spreads NaN; IEEE leaves NaN payload and sign open, so NaN
        # positions must agree and every other entry bit for bit
        assert np.array_equal(np.isnan(got), np.isnan(f)), (col, bad)
        fin = ~np.isnan(f)
        assert bits_equal(got[fin], f[fin]), (col, bad)


@pytest.mark.gpu
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_gpu_cholesky_solve(bh, oracle, uplo):
    """cholesky(Symmetric(A)) \\ B through the mirror: factor bits, solve to rounding."""
    rng = np.random.default_rng(13)
    n, l, u = 800, 6, 6
    D = np.asfortranarray(rng.uniform(-1, 1, size=(l + u + 1, n)))
    D[u, :] = 13 + rng.random(n)
    A = bh.BandedMatrix.from_host(D, n, l, u)
    S = bh.Symmetric(A, uplo)
    F = bh.cholesky(S)
    tri = np.asfortranarray(D[: u + 1] if uplo == "U" else D[u:])
    ref = tri.copy(order="F")
    for j in range(n):  # the corner slots the mirror zeroes
        for t in range(1, 7):
            if uplo == "U" and j - t < 0:
                ref[6 - t, j] = 0
            if uplo == "L" and j + t >= n:
                ref[t, j] = 0
    f, info = oracle.pbtf2(uplo, n, 6, ref)
    assert F.info == info == 0 and bits_equal(F.factors.host_data(), f)
    B = rng.standard_normal((n, 3))
    X = bh.ldiv_(F, bh.DeviceArray.from_numpy(B)).numpy()
    for c in range(3):
        xo = oracle.pbtrs(uplo, n, 6, f, B[:, c])
        assert np.abs(X[:, c] - xo).max() <= 1e-13 * np.abs(xo).max()
    assert np.allclose(S.to_dense() @ X, B, rtol=1e-11, atol=1e-11)
    with pytest.raises(bh.PosDefException):
        D2 = D.copy(order="F")
        D2[u, 5] = -100.0
        bh.cholesky(bh.Symmetric(bh.BandedMatrix.from_host(D2, n, l, u), uplo))


@pytest.mark.gpu
@pytest.mark.parametrize("n,kd,nrhs", [(20000, 8, 1), (6000, 40, 3), (500, 63, 2)])
def test_gpu_cholesky_partitioned_solve(bh, oracle, n, kd, nrhs):
    """ldiv_(cholesky(Symmetric(A)), B, method="partitioned"): an upper factor
    U solved as U^T y = B, U x = y by the partitioned band solve (kl = 0),
    within 1e-12 (max norm, relative) of the oracle's pbtrs on the same
    factor; a lower factor falls back to the exact pbtrs, bit for bit."""
    rng = np.random.default_rng(n + kd)
    D = np.asfortranarray(rng.uniform(-1, 1, size=(2 * kd + 1, n)))
    D[kd, :] = 2.0 * kd + 1.0 + rng.random(n)
    A = bh.BandedMatrix.from_host(D, n, kd, kd)
    F = bh.cholesky(bh.Symmetric(A, "U"))
    f = F.factors.host_data()
    B = rng.standard_normal((n, nrhs))
    X = bh.ldiv_(F, bh.DeviceArray.from_numpy(np.asfortranarray(B)), method="partitioned").numpy().reshape(n, nrhs)
    for c in range(nrhs):
        xo = oracle.pbtrs("U", n, kd, f, B[:, c])
        assert np.abs(X[:, c] - xo).max() <= 1e-12 * np.abs(xo).max(), c
    FL = bh.cholesky(bh.Symmetric(A, "L"))
    a = bh.ldiv_(FL, bh.DeviceArray.from_numpy(B[:, 0].copy()), method="partitioned").numpy()
    b = bh.ldiv_(FL, bh.DeviceArray.from_numpy(B[:, 0].copy())).numpy()
    assert bits_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gpu_sbmv_bitexact(bh, oracle, uplo, dtype):
    rng = np.random.default_rng(14)
    for n, k in [(1, 0), (100, 3), (5000, 17), (3000, 130)]:
        a = np.asfortranarray(rng.standard_normal((k + 1, n)).astype(dtype))
        x = rng.standard_normal(n).astype(dtype)
        y0 = rng.standard_normal(n).astype(dtype)
        pre = "d" if dtype == np.float64 else "s"
        A, X, Y = bh.DeviceArray.from_numpy(a), bh.DeviceArray.from_numpy(x), bh.DeviceArray.from_numpy(y0)
        bh._lib.call(f"bh_{pre}sbmv", uplo.encode(), n, k, dtype(0.75), A.ptr, A.ld, X.ptr, 1, dtype(-2.0), Y.ptr, 1)
        assert bits_equal(Y.numpy(), oracle.sbmv(uplo, n, k, dtype(0.75), a, x, dtype(-2.0), y0)), (n, k)


@pytest.mark.gpu
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_gpu_ldlt(bh, oracle, uplo):
    """ldlt(Symmetric(A)) \\ B: the factor against the oracle's restatement of the
    reference's loops to rounding (the update order differs), the solve against
    the oracle and a dense solve."""
    rng = np.random.default_rng(16)
    for n, l, u in [(500, 4, 4), (300, 40, 40), (64, 0, 0)]:
        D = np.asfortranarray(rng.uniform(-1, 1, size=(l + u + 1, n)))
        D[u, :] = 2 * (max(l, u) + 1) + rng.random(n)
        A = bh.BandedMatrix.from_host(D, n, l, u)
        S = bh.Symmetric(A, uplo)
        F = bh.ldlt(S)
        kd = S.bandwidth()
        tri = np.asfortranarray(D[: u + 1] if uplo == "U" else D[u:]).copy()
        for j in range(n):  # the corner slots the mirror zeroes
            for t in range(1, kd + 1):
                if uplo == "U" and j - t < 0:
                    tri[kd - t, j] = 0
                if uplo == "L" and j + t >= n:
                    tri[t, j] = 0
        f = oracle.sbldlt(uplo, n, kd, tri)
        got = F.factors.host_data()
        assert np.abs(got - f).max() <= 1e-13 * np.abs(f).max(), (n, l, u)
        B = rng.standard_normal((n, 2))
        X = bh.ldiv_(F, bh.DeviceArray.from_numpy(B)).numpy()
        for c in range(2):
            xo = oracle.sbldlts(uplo, n, kd, f, B[:, c])
            assert np.abs(X[:, c] - xo).max() <= 1e-12 * np.abs(xo).max()
        assert np.allclose(S.to_dense() @ X, B, rtol=1e-11, atol=1e-11)


@pytest.mark.gpu
def test_gpu_symmetric_view_mul(bh):
    rng = np.random.default_rng(15)
    n, l, u = 400, 4, 5
    D = np.asfortranarray(rng.standard_normal((l + u + 1, n)))
    A = bh.BandedMatrix.from_host(D, n, l, u)
    for uplo in ("U", "L"):
        S = bh.Symmetric(A, uplo)
        x = rng.standard_normal(n)
        y = (S @ bh.DeviceArray.from_numpy(x)).numpy()
        assert np.allclose(y, S.to_dense() @ x, rtol=1e-13, atol=1e-13)
