"""CPU: pin the oracle (oracle/bandoracle.c) before trusting it.

Pins, per function:
  gbmv 'N'     bit-exact vs OpenBLAS dgbmv/sgbmv (single thread) -- the routine
               BLAS.gbmv! reaches (matmul.jl:21-23); dense product on tiny inputs.
  gbmv 'T'     1e-14 vs OpenBLAS (its dot kernel order is unspecified); dense.
  gbtf2        bit-exact factors + pivots vs OpenBLAS dgbtrf for ku <= 64 (the
               unblocked path LAPACK.gbtrf! runs, BandedLU.jl:98); pivots equal
               and factors within 1e-11 for ku >= 65 (blocked dgbtrf); P*L*U = A;
               pivots equal dense dgetrf's on small cases (test_bandedlu.jl:34-38).
  gbtrs 'N'    bit-exact vs OpenBLAS dgbtrs; 'T' within 1e-12.
  tbsv upper   bit-exact vs OpenBLAS dtbsv 'N'.
  gbmm         dense product (integer-valued inputs: exact) + the reference's
               own closed-form and BandError cases (test_linalg.jl:272-295).
  QR           factors and tau vs LAPACK dgeqrf on the densified matrix
               (test_bandedqr.jl:35-36), Q'Q = I, QR = A, the closed-form
               lower-banded Q of test_bandedqr.jl:146-152, and the reflector
               convention on an already-triangular column.
  band grad    central finite differences of the loss.
A planted mistake in each restatement (tests at the bottom) must fail its pin.
"""
import numpy as np
import pytest
import scipy.linalg.blas as sblas
import scipy.linalg.lapack as slap
from threadpoolctl import threadpool_limits

from _util import band, bits_equal


@pytest.fixture(autouse=True)
def _one_blas_thread():
    # OpenBLAS splits gbmv over threads and sums partial vectors; the
    # reference recipe is the single-thread kernel.
    with threadpool_limits(1):
        yield


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,kl,ku,alpha,beta", [
    (2000, 2000, 8, 8, 1.0, 0.0), (2000, 2000, 8, 8, 2.0, 3.0), (10000, 10000, 4, 3, 1.0, 0.0),
    (500, 500, 0, 5, 1.5, 0.5), (500, 500, 7, 0, 1.0, 1.0), (300, 400, 20, 3, -1.0, 0.25),
    (400, 300, 3, 20, 0.5, 0.0), (1024, 1024, 128, 128, 1.0, 0.0)])
def test_gbmv_n_vs_openblas(oracle, dtype, m, n, kl, ku, alpha, beta):
    rng = np.random.default_rng(m + kl)
    A = band(rng, kl, ku, n, dtype, dist="n")
    x = rng.standard_normal(n).astype(dtype)
    y = rng.standard_normal(m).astype(dtype)
    f = sblas.dgbmv if dtype == np.float64 else sblas.sgbmv
    ref = f(m, n, kl, ku, alpha, A, x, beta=beta, y=y.copy())
    got = oracle.gbmv("N", m, n, kl, ku, dtype(alpha), A, x, dtype(beta), y)
    assert bits_equal(got, ref)


@pytest.mark.parametrize("m,n,kl,ku", [(2000, 2000, 8, 8), (300, 400, 20, 3), (400, 300, 3, 20)])
def test_gbmv_t_vs_openblas(oracle, m, n, kl, ku):
    rng = np.random.default_rng(m + ku)
    A = band(rng, kl, ku, n, dist="n")
    x = rng.standard_normal(m)
    y = rng.standard_normal(n)
    ref = sblas.dgbmv(m, n, kl, ku, 1.5, A, x, beta=0.5, y=y.copy(), trans=1)
    got = oracle.gbmv("T", m, n, kl, ku, 1.5, A, x, 0.5, y)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_gbmv_dense_tiny(oracle):
    rng = np.random.default_rng(2)
    for (m, n, kl, ku) in [(1, 1, 0, 0), (3, 5, 1, 2), (6, 4, 3, 1), (5, 5, 4, 4)]:
        A = band(rng, kl, ku, n, dist="n", m=m)
        M = oracle.band_to_dense(A, m, n, kl, ku)
        x, xt = rng.standard_normal(n), rng.standard_normal(m)
        np.testing.assert_allclose(oracle.gbmv("N", m, n, kl, ku, 1.0, A, x), M @ x, rtol=1e-14, atol=1e-14)
        np.testing.assert_allclose(oracle.gbmv("T", m, n, kl, ku, 1.0, A, xt), M.T @ xt, rtol=1e-14, atol=1e-14)
        y = np.full(m, np.nan)
        assert not np.isnan(oracle.gbmv("N", m, n, kl, ku, 1.0, A, x, 0.0, y)).any()


LU_PIN = [(300, 4, 3, "u"), (10000, 4, 3, "u"), (4096, 16, 16, "n"), (2000, 64, 64, "n"), (777, 1, 1, "u"),
          (500, 8, 8, "u"), (500, 0, 5, "n"), (500, 6, 0, "n"), (400, 40, 24, "n")]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,kl,ku,dist", LU_PIN)
def test_gbtf2_vs_openblas_dgbtrf(oracle, dtype, n, kl, ku, dist):
    rng = np.random.default_rng(n + kl * 3 + ku)
    A = band(rng, kl, ku, n, dtype, dist=dist)
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    f, piv, info = oracle.gbtf2(n, n, kl, ku, ab)
    trf = slap.dgbtrf if dtype == np.float64 else slap.sgbtrf
    fr, pr, ir = trf(np.asfortranarray(ab.copy()), kl, ku)
    assert info == ir
    assert np.array_equal(piv, pr + 1)
    assert bits_equal(f, fr)


@pytest.mark.parametrize("n,kl,ku", [(2000, 65, 65), (1000, 128, 128)])
def test_gbtf2_vs_blocked_dgbtrf(oracle, n, kl, ku):
    rng = np.random.default_rng(n)
    A = band(rng, kl, ku, n, dist="n")
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    f, piv, info = oracle.gbtf2(n, n, kl, ku, ab)
    fr, pr, ir = slap.dgbtrf(np.asfortranarray(ab.copy()), kl, ku)
    assert info == ir == 0
    assert np.array_equal(piv, pr + 1)
    assert np.abs(f - fr).max() <= 1e-11 * np.abs(fr).max()


def _unpack_lu(f, piv, n, kl, ku):
    """P, L, U from LAPACK band factors (for the P*L*U = A invariant)."""
    kv = kl + ku
    U = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - kv), j + 1):
            U[i, j] = f[kv + i - j, j]
    # apply the row interchanges and multipliers as gbtrs does: L^-1 P A = U
    Linv = np.eye(n)
    for j in range(n):
        p = piv[j] - 1
        if p != j:
            Linv[[j, p], :] = Linv[[p, j], :]
        lm = min(kl, n - 1 - j)
        for t in range(1, lm + 1):
            Linv[j + t, :] -= f[kv + t, j] * Linv[j, :]
    return Linv, U


@pytest.mark.parametrize("n,kl,ku", [(40, 3, 2), (30, 1, 4), (25, 6, 6)])
def test_gbtf2_invariants(oracle, n, kl, ku):
    rng = np.random.default_rng(n * kl)
    A = band(rng, kl, ku, n, dist="n")
    M = oracle.band_to_dense(A, n, n, kl, ku)
    f, piv, info = oracle.gbtf2(n, n, kl, ku, oracle.pad_for_lu(A, n, n, kl, ku))
    Linv, U = _unpack_lu(f, piv, n, kl, ku)
    np.testing.assert_allclose(Linv @ M, U, atol=1e-12 * np.abs(M).max())
    # pivots equal those of dense partial-pivot LU (test_bandedlu.jl:34-38)
    _, dpiv, _ = slap.dgetrf(M.copy())
    assert np.array_equal(piv, dpiv + 1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,kl,ku,dist", LU_PIN + [(1000, 128, 128, "n")])
def test_gbtrs_vs_openblas(oracle, dtype, n, kl, ku, dist):
    rng = np.random.default_rng(n + 5 * kl + ku)
    A = band(rng, kl, ku, n, dtype, dist=dist)
    A[ku, :] += kl + ku + 2  # keep triangular solves finite in FP32
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    trf, trs =(slap.dgbtrf, slap.dgbtrs) if dtype == np.float64 else (slap.sgbtrf, slap.sgbtrs)
    fr, pr, _ = trf(np.asfortranarray(ab.copy()), kl, ku)
    B = np.asfortranarray(rng.standard_normal((n, 3)).astype(dtype))
    xr, _ = trs(fr, kl, ku, B.copy(), pr)
    x = oracle.gbtrs("N", n, kl, ku, fr, pr + 1, B)
    assert bits_equal(x, xr)
    xr, _ = trs(fr, kl, ku, B.copy(), pr, trans=1)
    x = oracle.gbtrs("T", n, kl, ku, fr, pr + 1, B)
    tol = 1e-12 if dtype == np.float64 else 1e-3
    assert np.abs(x - xr).max() <= tol * (np.abs(xr).max() + 1)


def test_tbsv_vs_openblas(oracle):
    rng = np.random.default_rng(8)
    n, k = 500, 7
    a = band(rng, 0, k, n, dist="n")
    a[k, :] += 4.0
    x = rng.standard_normal(n)
    ref = sblas.dtbsv(k, a, x.copy(), lower=0, trans=0, diag=0)
    assert bits_equal(oracle.tbsv_u("N", n, k, a, x), ref)
    ref = sblas.dtbsv(k, a, x.copy(), lower=0, trans=1, diag=0)
    assert np.abs(oracle.tbsv_u("T", n, k, a, x) - ref).max() <= 1e-13 * np.abs(ref).max()


# ------------------------------------------------------------------ gbmm
def _dense_gbmm(oracle, n, k, m, A, Al, Au, B, Bl, Bu):
    return oracle.band_to_dense(A, n, k, Al, Au) @ oracle.band_to_dense(B, k, m, Bl, Bu)


@pytest.mark.parametrize("n,k,m", [(1, 1, 1), (5, 5, 5), (50, 50, 50), (5, 50, 1), (50, 5, 50), (1, 50, 5)])
@pytest.mark.parametrize("Al,Au,Bl,Bu", [(0, 0, 0, 0), (1, 2, 0, 1), (2, 30, 30, 1), (30, 30, 2, 2)])
def test_gbmm_vs_dense_integer(oracle, n, k, m, Al, Au, Bl, Bu):
    """Integer-valued entries: every product and sum is exact, so the band
    result must equal the dense product exactly (the n,nu,m x bands sweep of
    test_linalg.jl:138-223, with C = alpha*A*B + beta*C)."""
    rng = np.random.default_rng(n * 100 + k * 10 + m + Al)
    A = np.asfortranarray(rng.integers(-5, 6, size=(Al + Au + 1, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-5, 6, size=(Bl + Bu + 1, m)).astype(np.float64))
    Cl, Cu = min(n - 1, Al + Bl), min(m - 1, Au + Bu)
    C0 = np.asfortranarray(rng.integers(-5, 6, size=(Cl + Cu + 1, m)).astype(np.float64))
    C, st = oracle.gbmm(n, k, m, 2.0, A, Al, Au, B, Bl, Bu, 3.0, C0, Cl, Cu)
    assert st == 0
    want = 2.0 * _dense_gbmm(oracle, n, k, m, A, Al, Au, B, Bl, Bu) + 3.0 * oracle.band_to_dense(C0, n, m, Cl, Cu)
    assert np.array_equal(oracle.band_to_dense(C, n, m, Cl, Cu), want)


def test_gbmm_beta_zero_overwrites_nan(oracle):
    """NaN garbage in C is overwritten when beta == 0 (test_linalg.jl:314-325)."""
    rng = np.random.default_rng(1)
    A = band(rng, 0, 1, 1, m=1)
    B = band(rng, 0, 2, 2, m=1)
    C0 = np.full((3, 2), np.nan, order="F")
    C, st = oracle.gbmm(1, 1, 2, 1.0, A, 0, 1, B, 0, 2, 0.0, C0, 0, 2)
    got = oracle.band_to_dense(C, 1, 2, 0, 2)
    assert not np.isnan(got).any()
    assert np.array_equal(got, _dense_gbmm(oracle, 1, 1, 2, A, 0, 1, B, 0, 2))


def test_gbmm_not_enough_bands(oracle):
    """test_linalg.jl:272-295: a diagonal A times tridiagonal B fits a
    tridiagonal C; a nonzero super-diagonal in A is a BandError."""
    rng = np.random.default_rng(3)
    A = np.zeros((3, 10), order="F")
    A[1, :] = rng.standard_normal(10)
    B = band(rng, 1, 1, 10, dist="n")
    C, st = oracle.gbmm(10, 10, 10, 1.0, A, 1, 1, B, 1, 1, 0.0, np.zeros((3, 10), order="F"), 1, 1)
    assert st == 0
    np.testing.assert_array_equal(oracle.band_to_dense(C, 10, 10, 1, 1), _dense_gbmm(oracle, 10, 10, 10, A, 1, 1, B, 1, 1))
    A[0, 4] = 1.0
    _, st = oracle.gbmm(10, 10, 10, 1.0, A, 1, 1, B, 1, 1, 0.0, np.zeros((3, 10), order="F"), 1, 1)
    assert st == -1
    # the 2x1 case of test_linalg.jl:289-294 with A[1,1] = 0
    A2 = np.asfortranarray(np.array([[0.0], [0.7], [1.3]]))  # bands (2,0): rows 0..2 = diag, sub1, sub2
    A2[0, 0] = 0.0
    B2 = np.asfortranarray(np.array([[0.0], [2.0], [0.0]]))  # 1x1, bands (1,1)
    C2 = np.asfortranarray(rng.standard_normal((3, 1)))
    got, st = oracle.gbmm(2, 1, 1, 1.0, A2, 2, 0, B2, 1, 1, 1.0, C2, 2, 0)
    assert st == 0
    Ad = oracle.band_to_dense(A2, 2, 1, 2, 0)
    want = Ad @ np.array([[2.0]]) + oracle.band_to_dense(C2, 2, 1, 2, 0)
    np.testing.assert_allclose(oracle.band_to_dense(got, 2, 1, 2, 0), want, rtol=1e-15)


def test_gbmm_bits_match_columnwise_gbmv(oracle):
    """gbmm!'s interior column is one gbmv over A's band window (gbmm.jl:99-127):
    the oracle's entry recipe equals a direct gbmv on that window."""
    rng = np.random.default_rng(12)
    n = 200
    Al, Au, Bl, Bu = 3, 2, 2, 4
    A = band(rng, Al, Au, n, dist="n")
    B = band(rng, Bl, Bu, n, dist="n")
    Cl, Cu = Al + Bl, Au + Bu
    C, _ = oracle.gbmm(n, n, n, 1.5, A, Al, Au, B, Bl, Bu, 0.0, np.zeros((Cl + Cu + 1, n), order="F"), Cl, Cu)
    for j in (Cu + 1, 57, 120):
        p = j - Bu  # 0-based first contraction index
        win = np.asfortranarray(A[:, p:p + Bl + Bu + 1])  # (Al+Au+1) x (Bl+Bu+1), kl=Al+Au, ku=0
        x = B[:, j]
        y = oracle.gbmv("N", Cl + Cu + 1, Bl + Bu + 1, Al + Au, 0, 1.5, win, x)
        assert bits_equal(y, C[:, j])


# -------------------------------------------------------------------- QR
def _householder_dense(oracle, M):
    m, n = M.shape
    R = M.copy()
    tau = np.zeros(min(m, n))
    for k in range(min(m - 1, n)):
        x = R[k:, k].copy()
        nrm = np.sqrt(np.sum(x * x))
        if nrm == 0:
            continue
        nu = np.copysign(nrm, x[0])
        xi = x[0] + nu
        v = x / xi
        v[0] = 1.0
        tau[k] = xi / nu
        R[k:, k:] -= tau[k] * np.outer(v, v @ R[k:, k:])
        R[k + 1:, k] = v[1:]
    return R, tau


@pytest.mark.parametrize("m,n,l,u", [(100, 100, 3, 4), (102, 100, 3, 4), (100, 102, 3, 4), (10, 10, 3, 2),
                                     (14, 10, 3, 2), (10, 14, 3, 2), (100, 100, 1, 1)])
def test_qr_vs_lapack_dgeqrf(oracle, m, n, l, u):
    """qr(A).factors ≈ qrfactUnblocked!(Matrix(A)).factors and τ
    (test_bandedqr.jl:34-56), here against LAPACK dgeqrf (same Householder
    convention: R_kk = -sign(x1)||x||, tau = (beta - alpha)/beta)."""
    rng = np.random.default_rng(m + n + l)
    A = band(rng, l, u, n, m=m)
    M = oracle.band_to_dense(A, m, n, l, u)
    D = oracle.pad_for_lu(A, m, n, l, u)  # bands (l, l+u)
    F, tau = oracle.gbqrf(m, n, l, l + u, D)
    qr, taur, _, _ = slap.dgeqrf(M.copy())
    kmax = min(m - 1, n)
    np.testing.assert_allclose(tau[:kmax], taur[:kmax], rtol=1e-12, atol=1e-14)
    Fd = oracle.band_to_dense(F, m, n, l, l + u)
    # compare R (upper) and the reflectors (lower band) where the band holds them
    mask = np.zeros((m, n), bool)
    for j in range(n):
        for i in range(max(0, j - l - u), min(m, j + l + 1)):
            mask[i, j] = True
    np.testing.assert_allclose(Fd[mask], qr[mask], rtol=1e-10, atol=1e-12)
    if m == n:
        assert tau[-1] == 0.0  # last column not reflected for real square (bandedqr.jl:10)


def test_qr_q_orthogonal_and_solve(oracle):
    rng = np.random.default_rng(21)
    m = n = 60
    l, u = 3, 4
    A = band(rng, l, u, n)
    M = oracle.band_to_dense(A, m, n, l, u)
    F, tau = oracle.gbqrf(m, n, l, l + u, oracle.pad_for_lu(A, m, n, l, u))
    Q = oracle.gbqr_qmul(m, n, l, l + u, F, tau, np.eye(m))
    np.testing.assert_allclose(Q.T @ Q, np.eye(m), atol=1e-13)
    R = np.triu(oracle.band_to_dense(F, m, n, l, l + u))
    np.testing.assert_allclose(Q @ R, M, atol=1e-13)
    QtM = oracle.gbqr_qtmul(m, n, l, l + u, F, tau, M)
    np.testing.assert_allclose(QtM, R, atol=1e-13)
    b = rng.standard_normal((n, 2))
    x = oracle.gbqrs(m, n, l, l + u, F, tau, b)
    np.testing.assert_allclose(M @ x, b, atol=1e-11)


@pytest.mark.parametrize("m,n", [(60, 60), (70, 60), (60, 70)])
def test_qr_rmul_matches_dense_q(oracle, m, n):
    """banded_qr_rmul! restated: A Q and A Q' equal the dense products with the
    orthogonal factor built column by column from Q e_j (itself pinned by
    QᵀQ = I and QR = A above); the two directions differ, so an inverted loop
    order or a transposed application fails."""
    rng = np.random.default_rng(m + 3 * n)
    l, u = 3, 4
    A = band(rng, l, u, n, m=m)
    F, tau = oracle.gbqrf(m, n, l, l + u, oracle.pad_for_lu(A, m, n, l, u))
    Q = oracle.gbqr_qmul(m, n, l, l + u, F, tau, np.eye(m))
    X = rng.standard_normal((9, m))
    XQ = oracle.gbqr_rmul("N", m, n, l, l + u, F, tau, X)
    XQt = oracle.gbqr_rmul("T", m, n, l, l + u, F, tau, X)
    np.testing.assert_allclose(XQ, X @ Q, atol=1e-13)
    np.testing.assert_allclose(XQt, X @ Q.T, atol=1e-13)
    assert np.abs(XQ - XQt).max() > 1e-3


def test_qr_reflector_convention_on_triangular_column(oracle):
    """reflector! on x = [x1; 0...] returns tau = 2 and x1 -> -x1 (its
    convention; LAPACK's dlarfg would return tau = 0 here)."""
    n = 6
    A = np.zeros((2, n), order="F")  # bands (1, 0)
    A[0, :] = np.arange(1, n + 1)  # diagonal, sub-diagonal zero
    F, tau = oracle.gbqrf(n, n, 1, 1, oracle.pad_for_lu(A, n, n, 1, 0))
    np.testing.assert_array_equal(tau[:-1], 2.0)
    assert tau[-1] == 0.0
    diag = F[1, :]
    np.testing.assert_array_equal(diag[:-1], -np.arange(1, n))
    assert diag[-1] == n


def test_qr_closed_form_lower_banded_q(oracle):
    """test_bandedqr.jl:146-152: reflectors v_k = [1, 0, -s_k/(c_k+1)] with
    tau_k = c_k + 1 give an orthogonal Q."""
    n = 10
    k = np.arange(1, n + 1)
    c = np.sqrt(2 * (2 * k + 3) / ((k + 2) * (k + 3)))
    s = np.sqrt((k * (k + 1)) / ((k + 2) * (k + 3)))
    # BandedMatrix(-2 => v) with length(v) = 10 is 12 x 10; store it with
    # bands (2, 0) so the -2 band (row 2) holds -(s/(c+1))
    D = np.zeros((3, n), order="F")
    D[2, :] = -(s / (c + 1))
    tau = c + 1
    Q = oracle.gbqr_qmul(n + 2, n, 2, 0, D, tau, np.eye(n + 2))
    np.testing.assert_allclose(Q.T @ Q, np.eye(n + 2), atol=1e-13)


# ------------------------------------------------------------- gradient
def test_band_grad_finite_differences(oracle):
    rng = np.random.default_rng(5)
    n, kl, ku = 40, 2, 3
    A = band(rng, kl, ku, n, dist="n")
    A[ku, :] += 2 * (kl + ku)
    b, w, x0 = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)

    def loss(Ab, bb):
        M = oracle.band_to_dense(Ab, n, n, kl, ku)
        return w @ np.linalg.solve(M, bb) + 0.5 * np.sum((M @ x0) ** 2)

    M = oracle.band_to_dense(A, n, n, kl, ku)
    x = np.linalg.solve(M, b)
    lam = np.linalg.solve(M.T, w)
    G = oracle.band_grad(n, kl, ku, lam, x, M @ x0, x0)
    h = 1e-6
    for (t, j) in [(ku, 3), (0, 10), (kl + ku, 7), (1, 30)]:
        Ap, Am = A.copy(), A.copy()
        Ap[t, j] += h
        Am[t, j] -= h
        fd = (loss(Ap, b) - loss(Am, b)) / (2 * h)
        assert abs(fd - G[t, j]) <= 1e-6 * max(1, abs(fd))


# ------------------------------------------------- planted mistakes fail pins
def test_planted_unfused_gbmv_fails_pin(oracle):
    """An unfused y += (alpha*x)*a (a plausible restatement mistake) breaks the
    OpenBLAS bit pin, so the pin above is discriminating."""
    rng = np.random.default_rng(0)
    n, kl, ku = 2000, 8, 8
    A = band(rng, kl, ku, n, dist="n")
    x = rng.standard_normal(n)
    ref = sblas.dgbmv(n, n, kl, ku, 1.0, A, x)
    y = np.zeros(n)
    for j in range(n):
        t = x[j]
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            y[i] = y[i] + t * A[ku + i - j, j]
    assert not bits_equal(y, ref)


@pytest.mark.parametrize("variant", [1, 2])
def test_planted_gbtf2_mistakes_fail_pin(oracle, variant):
    """Dividing by the pivot (1) or an unfused rank-1 update (2), planted in
    the oracle's own gbtf2, breaks the bit pin against LAPACK dgbtrf while the
    unplanted restatement passes it."""
    rng = np.random.default_rng(1)
    n, kl, ku = 300, 4, 3
    A = band(rng, kl, ku, n)
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    fr, pr, _ = slap.dgbtrf(np.asfortranarray(ab.copy()), kl, ku)
    f0, _, _ = oracle.gbtf2(n, n, kl, ku, ab)
    assert bits_equal(f0, fr)
    f, _, _ = oracle.gbtf2_planted(variant, n, n, kl, ku, ab)
    assert not bits_equal(f, fr)
