"""GPU parity: band LU (gbtrf) and solves (gbtrs) against the oracle.

For ku <= 64 the oracle is the restated LAPACK dgbtf2 (pinned bit for bit
against scipy's OpenBLAS dgbtrf in tests/test_oracle.py): factors and pivots
must be bit-identical.  gbtrs 'N' must be bit-identical; 'T' within 1e-13
relative."""
import numpy as np
import pytest

from _util import band, bits_equal

pytestmark = pytest.mark.gpu

LU_CASES = [
    # n, kl, ku, dist, diag shift
    (1, 0, 0, "u", 0.0),
    (5, 1, 1, "u", 0.0),
    (300, 4, 3, "u", 0.0),
    (10000, 4, 3, "u", 0.0),          # the plain CPU-runnable config
    (4096, 16, 16, "n", 0.0),
    (4096, 16, 16, "n", 64.0),
    (2000, 2, 2, "u", 5.0),
    (1500, 8, 8, "u", 17.0),
    (777, 1, 1, "u", 0.0),
    (500, 0, 5, "n", 0.0),
    (500, 6, 0, "n", 0.0),
    (600, 24, 24, "n", 0.0),          # kl above 16 with a window under 64 columns: blocked kernel
    (400, 20, 10, "u", 0.0),
    (300, 40, 24, "n", 0.0),
    (260, 64, 64, "n", 0.0),
]


def _make(rng, n, kl, ku, dist, shift, dtype=np.float64):
    A = band(rng, kl, ku, n, np.float64, dist=dist)
    A[ku, :] += shift
    return A.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,kl,ku,dist,shift", LU_CASES)
def test_gbtrf_bitexact(bh, oracle, dtype, n, kl, ku, dist, shift):
    rng = np.random.default_rng(n + 31 * kl + ku)
    A = _make(rng, n, kl, ku, dist, shift, dtype)
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    fref, pref, iref = oracle.gbtf2(n, n, kl, ku, ab)
    BA = bh.BandedMatrix.from_host(A, n, kl, ku)
    F = bh.lu(BA, check=False)
    assert F.info == iref
    assert np.array_equal(F.ipiv, pref)
    assert bits_equal(F.factors.host_data(), fref)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,kl,ku,dist,shift", LU_CASES)
@pytest.mark.parametrize("nrhs", [1, 3])
def test_gbtrs(bh, oracle, dtype, n, kl, ku, dist, shift, nrhs):
    rng = np.random.default_rng(n + 7 * kl + 3 * ku + nrhs)
    # a diagonal shift keeps the triangular solves finite in FP32
    A = _make(rng, n, kl, ku, dist, max(shift, kl + ku + 2.0), dtype)
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    f, piv, info = oracle.gbtf2(n, n, kl, ku, ab)
    if info != 0:
        pytest.skip("singular draw")
    B = np.asfortranarray(rng.standard_normal((n, nrhs)).astype(dtype))
    BA = bh.BandedMatrix.from_host(A, n, kl, ku)
    F = bh.lu(BA)
    for trans in ("N", "T"):
        ref = oracle.gbtrs(trans, n, kl, ku, f, piv, B)
        dB = bh.DeviceArray.from_numpy(B)
        bh.ldiv_(F if trans == "N" else F.transpose(), dB)
        out = dB.numpy()
        if trans == "N":
            assert bits_equal(out, ref), np.abs(out - ref).max()
        else:
            tol = 1e-11 if dtype == np.float64 else 1e-3
            assert np.abs(out - ref).max() <= tol * (np.abs(ref).max() + 1)


def test_gbtrf_singular_column(bh, oracle):
    """A zero column gives LAPACK info = that column (1-based), no exception
    with check=False, SingularException with check=True."""
    rng = np.random.default_rng(3)
    n, kl, ku = 50, 2, 3
    A = band(rng, kl, ku, n)
    A[:, 17] = 0.0
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    fref, pref, iref = oracle.gbtf2(n, n, kl, ku, ab)
    assert iref > 0
    F = bh.lu(bh.BandedMatrix.from_host(A, n, kl, ku), check=False)
    assert F.info == iref and not F.issuccess()
    assert np.array_equal(F.ipiv, pref)
    assert bits_equal(F.factors.host_data(), fref)
    with pytest.raises(bh.SingularException):
        bh.lu(bh.BandedMatrix.from_host(A, n, kl, ku))


def test_gbtrf_ties_pick_first(bh, oracle):
    """Exact ties in |a| pick the first row (IDAMAX); integer-valued entries."""
    n, kl, ku = 64, 3, 2
    rng = np.random.default_rng(9)
    A = np.asfortranarray(rng.integers(-3, 4, size=(kl + ku + 1, n)).astype(np.float64))
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    fref, pref, iref = oracle.gbtf2(n, n, kl, ku, ab)
    F = bh.lu(bh.BandedMatrix.from_host(A, n, kl, ku), check=False)
    assert F.info == iref
    assert np.array_equal(F.ipiv, pref)
    assert bits_equal(F.factors.host_data(), fref)


def test_rectangular_lu(bh, oracle):
    """lu of a 5x4 and 4x5 band (test_bandedlu.jl:147-155 shapes)."""
    rng = np.random.default_rng(4)
    for (m, n) in [(5, 4), (4, 5), (40, 30), (30, 40)]:
        kl, ku = 1, 1
        A = band(rng, kl, ku, n, m=m)
        ab = oracle.pad_for_lu(A, m, n, kl, ku)
        fref, pref, iref = oracle.gbtf2(m, n, kl, ku, ab)
        F = bh.lu(bh.BandedMatrix.from_host(A, m, kl, ku), check=False)
        assert F.shape == (m, n)
        assert np.array_equal(F.ipiv, pref)
        assert bits_equal(F.factors.host_data(), fref)


def test_solve_backward_error_and_no_alias(bh, oracle):
    """A\\b leaves b untouched (test_bandedlu.jl:22-23) and reaches FP64
    backward error <= 1e-12 (test_bandedlu.jl:11-24 with b = 1:5)."""
    rng = np.random.default_rng(0)
    for (n, kl, ku) in [(5, 1, 1), (10000, 4, 3), (2000, 16, 16)]:
        A = band(rng, kl, ku, n)
        b = np.arange(1, n + 1, dtype=np.float64)
        db = bh.DeviceArray.from_numpy(b)
        x = bh.solve(bh.BandedMatrix.from_host(A, n, kl, ku), db).numpy()
        assert np.array_equal(db.numpy(), b)
        M = oracle.band_to_dense(A, n, n, kl, ku) if n <= 2000 else None
        Ax = oracle.gbmv("N", n, n, kl, ku, 1.0, A, x)
        normA = np.abs(A).sum(axis=0).max() * 2
        berr = np.abs(b - Ax).max() / (normA * np.abs(x).max() + np.abs(b).max())
        assert berr <= 1e-12
        if M is not None:
            np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=1e-8, atol=1e-10)
        # from n = 4096 on A\b solves partitioned; method="exact" keeps LAPACK's
        # solve, and both agree to rounding
        xe = bh.solve(bh.BandedMatrix.from_host(A, n, kl, ku), db, method="exact").numpy()
        assert np.abs(x - xe).max() <= 1e-10 * np.abs(xe).max()


def test_solve_errors(bh):
    rng = np.random.default_rng(1)
    A = bh.BandedMatrix.from_host(band(rng, 1, 1, 5, m=7), 7, 1, 1)
    with pytest.raises(bh.DimensionMismatch):
        bh.solve(A, bh.DeviceArray.from_numpy(np.ones(7)))
    Asq = bh.BandedMatrix.from_host(band(rng, 1, 1, 5), 5, 1, 1)
    with pytest.raises(bh.DimensionMismatch):
        bh.ldiv_(bh.lu(Asq), bh.DeviceArray.from_numpy(np.ones(3)))


def test_zero_size(bh):
    A = bh.BandedMatrix.zeros((0, 0), (1, 1))
    F = bh.lu(A)
    assert F.ipiv.size == 0 and F.info == 0
    assert bh.ldiv_(F, bh.DeviceArray((0,))).shape == (0,)


PIV_CASES = [
    # n, kl, ku: no diagonal shift, N(0,1) entries -- row interchanges on most steps
    (300, 4, 3),
    (2000, 8, 8),
    (4096, 16, 16),
    (600, 24, 24),
    (260, 64, 64),
    (1000, 128, 128),  # blocked MFMA factorisation (ku >= 65)
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,kl,ku", PIV_CASES)
@pytest.mark.parametrize("nrhs", [1, 3])
def test_gbtrs_with_row_interchanges(bh, oracle, dtype, n, kl, ku, nrhs):
    """gbtrs 'N' and 'T' on factors with pivoting active.  Both solves run on the
    GPU's own factors, and so does the oracle's xGBTRS restatement: 'N' must be
    bit-identical (the same recipe as dgbtrs), 'T' within 1e-11 (FP64) / 1e-3
    (FP32) of the oracle relative to the solution's scale."""
    rng = np.random.default_rng(n + 13 * kl + nrhs)
    A = band(rng, kl, ku, n, np.float64, dist="n").astype(dtype)
    F = bh.lu(bh.BandedMatrix.from_host(A, n, kl, ku), check=False)
    assert F.info == 0
    swaps = int(np.sum(F.ipiv != np.arange(1, n + 1)))
    assert swaps > n // 4, swaps
    f = F.factors.host_data()
    B = np.asfortranarray(rng.standard_normal((n, nrhs)).astype(dtype))
    for trans in ("N", "T"):
        ref = oracle.gbtrs(trans, n, kl, ku, f, F.ipiv, B)
        dB = bh.DeviceArray.from_numpy(B)
        bh.ldiv_(F if trans == "N" else F.transpose(), dB)
        out = dB.numpy()
        if trans == "N":
            assert bits_equal(out, ref), np.abs(out - ref).max()
        else:
            tol = 1e-11 if dtype == np.float64 else 1e-3
            assert np.abs(out - ref).max() <= tol * (np.abs(ref).max() + 1)


def test_ldiv_into_a_column_view(bh, oracle):
    """ldiv! on a 3-column window of a larger array solves exactly those columns
    in place and leaves the rest untouched (the leading dimension is the parent's)."""
    rng = np.random.default_rng(21)
    n, kl, ku = 300, 5, 4
    A = band(rng, kl, ku, n, dist="n")
    F = bh.lu(bh.BandedMatrix.from_host(A, n, kl, ku))
    big = np.asfortranarray(rng.standard_normal((n + 7, 6)))
    dbig = bh.DeviceArray.from_numpy(big)
    view = dbig.view(2, 1, (n, 3))
    bh.ldiv_(F, view)
    out = dbig.numpy()
    f = F.factors.host_data()
    ref = oracle.gbtrs("N", n, kl, ku, f, F.ipiv, np.asfortranarray(big[2:2 + n, 1:4]))
    assert bits_equal(np.ascontiguousarray(out[2:2 + n, 1:4]), np.ascontiguousarray(ref))
    mask = np.ones(big.shape, bool)
    mask[2:2 + n, 1:4] = False
    assert np.array_equal(out[mask], big[mask])
