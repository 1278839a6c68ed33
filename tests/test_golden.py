"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py from scipy's OpenBLAS and from the reference
tests' exact integer cases) against the oracle (CPU) and the library (GPU).

Bars: gbmv 'N', gbtrf factors + pivots, gbtrs 'N' bit for bit (the reference
calls the same OpenBLAS routines the fixtures come from); gbtrs 'T' 1e-13;
integer closed forms exactly; QR's R and tau against dgeqrf to 1e-12 (the
Householder convention is the same, the operation order is not); tbsv 'N' and
pbtrf bit for bit, tbsv 'T', tbmv, pbtrs and sbmv to 1e-13 (OpenBLAS blocks
those sums)."""
import os

import numpy as np
import pytest

from _util import bits_equal

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name))


GBMV_TAGS = ["sq_d", "wide_d", "tall_d", "sq_s"]
LU_TAGS = ["piv_d", "piv_s", "wide_d", "ties_d", "singular_d", "b15"]


def _r_from_padded(F, n, l, u):
    lu_ = l + u
    R = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - lu_), j + 1):
            R[i, j] = F[lu_ + i - j, j]
    return R


# ---------------------------------------------------------------- oracle (CPU)
@pytest.mark.parametrize("tag", GBMV_TAGS)
def test_oracle_gbmv_golden(oracle, tag):
    z = load("gbmv_openblas.npz")
    m, n, kl, ku = (int(v) for v in z[f"{tag}_dims"])
    alpha, beta = z[f"{tag}_ab"]
    D = z[f"{tag}_D"]
    dt = D.dtype.type
    out = oracle.gbmv("N", m, n, kl, ku, dt(alpha), D, z[f"{tag}_x"], dt(beta), z[f"{tag}_y"].copy())
    assert bits_equal(out, z[f"{tag}_ref"])


@pytest.mark.parametrize("tag", LU_TAGS)
def test_oracle_lu_golden(oracle, tag):
    z = load("lu_openblas.npz")
    n, kl, ku = (int(v) for v in z[f"{tag}_dims"])
    D = z[f"{tag}_D"]
    ab = oracle.pad_for_lu(np.nan_to_num(D, nan=0.0), n, n, kl, ku)
    f, ipiv, info = oracle.gbtf2(n, n, kl, ku, ab)
    assert info == int(z[f"{tag}_info"][0])
    assert np.array_equal(ipiv, z[f"{tag}_ipiv"])
    assert bits_equal(f, z[f"{tag}_lu"])
    if f"{tag}_xN" in z:
        xN = oracle.gbtrs("N", n, kl, ku, f, ipiv, z[f"{tag}_B"].copy())
        assert bits_equal(xN, z[f"{tag}_xN"])
    if f"{tag}_xT" in z:
        xT = oracle.gbtrs("T", n, kl, ku, f, ipiv, z[f"{tag}_B"].copy())
        ref = z[f"{tag}_xT"]
        tol = 1e-13 if ref.dtype == np.float64 else 1e-5
        assert np.abs(xT - ref).max() <= tol * np.abs(ref).max()


def test_oracle_lu_b15_matches_dense_solve(oracle):
    z = load("lu_openblas.npz")
    assert np.allclose(z["b15_xN"][:, 0], z["b15_dense_x"], rtol=1e-14, atol=0)


def test_oracle_integer_closed_forms(oracle):
    z = load("integer_closed_form.npz")
    D = z["sym_D"]
    assert np.array_equal(oracle.gbmv("N", 10, 10, 2, 2, 1.0, D, z["sym_v"], 0.0), z["sym_Mv"])
    C, st = oracle.gbmm(10, 10, 10, 1.0, np.nan_to_num(D, nan=0.0), 2, 2, np.nan_to_num(D, nan=0.0), 2, 2, 0.0,
                        np.zeros((9, 10)), 4, 4)
    assert st == 0
    ref = np.nan_to_num(z["sym_MM"], nan=0.0)
    assert np.array_equal(C, ref)
    S = z["strided_D"]
    assert np.array_equal(oracle.gbmv("N", 101, 101, 1, 2, 1.0, S, z["strided_x"], 0.0), z["strided_Bx"])


@pytest.mark.parametrize("tag", ["a", "b"])
def test_oracle_qr_golden(oracle, tag):
    z = load("qr_lapack.npz")
    n, l, u = (int(v) for v in z[f"{tag}_dims"])
    D = z[f"{tag}_D"]
    Dp = oracle.pad_for_lu(np.nan_to_num(D, nan=0.0), n, n, l, u)
    F, tau = oracle.gbqrf(n, n, l, l + u, Dp)
    R = _r_from_padded(F, n, l, u)
    scale = np.abs(z[f"{tag}_R"]).max()
    assert np.abs(R - z[f"{tag}_R"]).max() <= 1e-12 * scale
    assert np.abs(tau - z[f"{tag}_tau"]).max() <= 1e-12


# ---------------------------------------------------------------- library (GPU)
@pytest.mark.gpu
@pytest.mark.parametrize("tag", GBMV_TAGS)
def test_gpu_gbmv_golden(bh, tag):
    z = load("gbmv_openblas.npz")
    m, n, kl, ku = (int(v) for v in z[f"{tag}_dims"])
    alpha, beta = z[f"{tag}_ab"]
    D = z[f"{tag}_D"]
    y = bh.DeviceArray.from_numpy(z[f"{tag}_y"].copy())
    bh.mul(y, bh.BandedMatrix.from_host(D, m, kl, ku), bh.DeviceArray.from_numpy(z[f"{tag}_x"]), alpha, beta)
    assert bits_equal(y.numpy(), z[f"{tag}_ref"])


@pytest.mark.gpu
@pytest.mark.parametrize("tag", LU_TAGS)
def test_gpu_lu_golden(bh, tag):
    z = load("lu_openblas.npz")
    n, kl, ku = (int(v) for v in z[f"{tag}_dims"])
    D = z[f"{tag}_D"]
    F = bh.lu(bh.BandedMatrix.from_host(D, n, kl, ku), check=False)
    assert F.info == int(z[f"{tag}_info"][0])
    assert np.array_equal(F.ipiv, z[f"{tag}_ipiv"])
    assert bits_equal(F.factors.data.numpy(), z[f"{tag}_lu"])
    if f"{tag}_xN" in z:
        B = bh.DeviceArray.from_numpy(z[f"{tag}_B"].copy())
        bh.ldiv_(F, B)
        assert bits_equal(B.numpy(), z[f"{tag}_xN"])
    if f"{tag}_xT" in z:
        # the transposed solve on LAPACK's factors, row interchanges included
        ref = z[f"{tag}_xT"]
        B = bh.DeviceArray.from_numpy(z[f"{tag}_B"].copy())
        bh.ldiv_(F.transpose(), B)
        tol = 1e-13 if ref.dtype == np.float64 else 1e-5
        assert np.abs(B.numpy() - ref).max() <= tol * (np.abs(ref).max() + 1)


@pytest.mark.gpu
def test_gpu_integer_closed_forms(bh):
    z = load("integer_closed_form.npz")
    y = bh.DeviceArray((10,), np.float64)
    bh.mul(y, bh.BandedMatrix.from_host(z["sym_D"], 10, 2, 2), bh.DeviceArray.from_numpy(z["sym_v"]))
    assert np.array_equal(y.numpy(), z["sym_Mv"])
    y = bh.DeviceArray((101,), np.float64)
    bh.mul(y, bh.BandedMatrix.from_host(z["strided_D"], 101, 1, 2), bh.DeviceArray.from_numpy(z["strided_x"]))
    assert np.array_equal(y.numpy(), z["strided_Bx"])
    A = bh.BandedMatrix.from_host(np.nan_to_num(z["sym_D"], nan=0.0), 10, 2, 2)
    C = bh.gbmm(A, A)
    ref = z["sym_MM"]
    inside = ~np.isnan(ref)  # corner slots of band storage lie outside the matrix and are never written
    assert np.array_equal(C.data.numpy()[inside], ref[inside])


TRI_TAGS = ["UN", "UU", "LN", "LU"]
SYM_TAGS = ["U4", "U20", "L4", "L20"]


def _close(got, ref, tol=1e-13):
    return np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("tag", TRI_TAGS)
def test_oracle_triangular_golden(oracle, tag):
    z = load("tri_openblas.npz")
    n, k = (int(v) for v in z[f"{tag}_dims"])
    uplo, diag = tag[0], tag[1]
    a, x = z[f"{tag}_a"], z[f"{tag}_x"]
    assert bits_equal(oracle.tbsv(uplo, "N", diag, n, k, a, x), z[f"{tag}_svN"])
    assert _close(oracle.tbsv(uplo, "T", diag, n, k, a, x), z[f"{tag}_svT"])
    for t in "NT":
        assert _close(oracle.tbmv(uplo, t, diag, n, k, a, x), z[f"{tag}_mv{t}"])


@pytest.mark.parametrize("tag", SYM_TAGS)
def test_oracle_symmetric_golden(oracle, tag):
    z = load("sym_lapack.npz")
    n, kd = (int(v) for v in z[f"{tag}_dims"])
    uplo = tag[0]
    f, info = oracle.pbtf2(uplo, n, kd, z[f"{tag}_a"])
    assert info == int(z[f"{tag}_info"][0]) == 0
    assert bits_equal(f, z[f"{tag}_chol"])
    assert _close(oracle.pbtrs(uplo, n, kd, f, z[f"{tag}_b"]), z[f"{tag}_x"])
    assert _close(oracle.sbmv(uplo, n, kd, 1.25, z[f"{tag}_a"], z[f"{tag}_sx"]), z[f"{tag}_sy"])


@pytest.mark.gpu
@pytest.mark.parametrize("tag", TRI_TAGS)
def test_gpu_triangular_golden(bh, tag):
    z = load("tri_openblas.npz")
    n, k = (int(v) for v in z[f"{tag}_dims"])
    uplo, diag = tag[0], tag[1]
    A = bh.DeviceArray.from_numpy(z[f"{tag}_a"])
    for t in "NT":
        X = bh.DeviceArray.from_numpy(z[f"{tag}_x"])
        bh.tbsv_(uplo, t, diag, n, k, A, X)
        if t == "N":
            assert bits_equal(X.numpy(), z[f"{tag}_svN"])
        else:
            assert _close(X.numpy(), z[f"{tag}_svT"])
        X = bh.DeviceArray.from_numpy(z[f"{tag}_x"])
        bh.tbmv_(uplo, t, diag, n, k, A, X)
        assert _close(X.numpy(), z[f"{tag}_mv{t}"])


@pytest.mark.gpu
@pytest.mark.parametrize("tag", SYM_TAGS)
def test_gpu_symmetric_golden(bh, tag):
    import ctypes

    z = load("sym_lapack.npz")
    n, kd = (int(v) for v in z[f"{tag}_dims"])
    uplo = tag[0].encode()
    A = bh.DeviceArray.from_numpy(z[f"{tag}_a"])
    info = ctypes.c_int64(-1)
    bh._lib.call("bh_dpbtrf", uplo, n, kd, A.ptr, A.ld, ctypes.byref(info))
    assert info.value == 0 and bits_equal(A.numpy(), z[f"{tag}_chol"])
    B = bh.DeviceArray.from_numpy(z[f"{tag}_b"])
    bh._lib.call("bh_dpbtrs", uplo, n, kd, 1, A.ptr, A.ld, B.ptr, n)
    assert _close(B.numpy(), z[f"{tag}_x"])
    A0 = bh.DeviceArray.from_numpy(z[f"{tag}_a"])
    X = bh.DeviceArray.from_numpy(z[f"{tag}_sx"])
    Y = bh.DeviceArray((n,), np.float64)
    bh._lib.call("bh_dsbmv", uplo, n, kd, 1.25, A0.ptr, A0.ld, X.ptr, 1, 0.0, Y.ptr, 1)
    assert _close(Y.numpy(), z[f"{tag}_sy"])
