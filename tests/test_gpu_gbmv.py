"""GPU parity: gbmv / multi-RHS gbmv / gbmm against the oracle.

The oracle restates the single-threaded OpenBLAS gbmv recipe and gbmm!'s
per-column decomposition (oracle/bandoracle.c); 'N' products must match it
bit for bit, 'T' products within 1e-13 relative (the BLAS dot order is not
specified)."""
import numpy as np
import pytest

from _util import band, bits_equal

pytestmark = pytest.mark.gpu

GBMV_CASES = [
    # m, n, kl, ku, alpha, beta
    (1, 1, 0, 0, 1.0, 0.0),
    (5, 5, 1, 1, 1.0, 0.0),
    (10000, 10000, 4, 3, 1.0, 0.0),
    (2000, 2000, 8, 8, 2.0, 3.0),
    (3001, 2999, 8, 8, -1.5, 0.5),
    (700, 1000, 2, 30, 1.0, 1.0),
    (1000, 700, 30, 2, 0.5, 0.0),
    (4096, 4096, 0, 0, 1.0, 0.0),
    (2048, 2048, 128, 128, 1.0, 0.0),   # wide: direct kernel
    (1500, 1800, 100, 40, 1.25, -2.0),  # wide, rectangular
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,kl,ku,alpha,beta", GBMV_CASES)
def test_gbmv_n_bitexact(bh, oracle, dtype, m, n, kl, ku, alpha, beta):
    rng = np.random.default_rng(m * 7 + n + kl * 13 + ku)
    A = band(rng, kl, ku, n, dtype, dist="n", nan_corners=True, m=m)
    x = rng.standard_normal(n).astype(dtype)
    y = rng.standard_normal(m).astype(dtype)
    if beta == 0.0:
        y[:] = np.nan  # beta == 0 must overwrite garbage
    ref = oracle.gbmv("N", m, n, kl, ku, dtype(alpha), A, x, dtype(beta), y)
    BA = bh.BandedMatrix.from_host(A, m, kl, ku)
    dx, dy = bh.DeviceArray.from_numpy(x), bh.DeviceArray.from_numpy(y)
    bh.mul(dy, BA, dx, alpha, beta)
    out = dy.numpy()
    assert not np.isnan(out).any()
    assert bits_equal(out, ref), np.abs(out - ref).max()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,kl,ku,alpha,beta", GBMV_CASES)
def test_gbmv_t(bh, oracle, dtype, m, n, kl, ku, alpha, beta):
    rng = np.random.default_rng(m + n * 3 + kl + ku * 5)
    A = band(rng, kl, ku, n, dtype, dist="n", nan_corners=True, m=m)
    x = rng.standard_normal(m).astype(dtype)
    y = rng.standard_normal(n).astype(dtype)
    ref = oracle.gbmv("T", m, n, kl, ku, dtype(alpha), A, x, dtype(beta), y)
    BA = bh.BandedMatrix.from_host(A, m, kl, ku)
    dx, dy = bh.DeviceArray.from_numpy(x), bh.DeviceArray.from_numpy(y)
    bh.mul(dy, BA, dx, alpha, beta, trans="T")
    out = dy.numpy()
    tol = 1e-13 if dtype == np.float64 else 2e-5
    scale = np.abs(ref).max() + 1e-30
    assert np.abs(out - ref).max() / scale < tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,kl,ku,nrhs", [(1000, 1000, 8, 8, 64), (4099, 4099, 4, 3, 5), (777, 900, 20, 3, 16),
                                            (2048, 2048, 128, 128, 16), (300, 300, 0, 0, 3),
                                            (1500, 1700, 140, 90, 20), (700, 500, 90, 200, 2)])
def test_gbmv_multi_bitexact(bh, oracle, dtype, m, n, kl, ku, nrhs):
    rng = np.random.default_rng(nrhs + m)
    A = band(rng, kl, ku, n, dtype, dist="n", nan_corners=True, m=m)
    X = np.asfortranarray(rng.standard_normal((n, nrhs)).astype(dtype))
    Y = np.asfortranarray(rng.standard_normal((m, nrhs)).astype(dtype))
    alpha, beta = 1.5, 0.25
    BA = bh.BandedMatrix.from_host(A, m, kl, ku)
    dX, dY = bh.DeviceArray.from_numpy(X), bh.DeviceArray.from_numpy(Y)
    bh.mul(dY, BA, dX, alpha, beta)
    out = dY.numpy()
    for r in range(nrhs):
        ref = oracle.gbmv("N", m, n, kl, ku, dtype(alpha), A, X[:, r], dtype(beta), Y[:, r])
        assert bits_equal(out[:, r], ref), (r, np.abs(out[:, r] - ref).max())


def test_gbmv_closed_form_int_symmetric(bh):
    """test_linalg.jl:52-73: B = Symmetric band (0=>1:10, 1=>11:19, 2=>21:28),
    v = 1:10 -- integer-valued, so every product is exact."""
    n = 10
    M = np.zeros((n, n))
    for i in range(n):
        M[i, i] = i + 1
    for i in range(n - 1):
        M[i, i + 1] = M[i + 1, i] = 11 + i
    for i in range(n - 2):
        M[i, i + 2] = M[i + 2, i] = 21 + i
    from oracle.oracle import dense_to_band

    D = dense_to_band(M, 2, 2)
    BA = bh.BandedMatrix.from_host(D, n, 2, 2)
    v = np.arange(1, n + 1, dtype=np.float64)
    dv = bh.DeviceArray.from_numpy(v)
    dw = bh.DeviceArray.from_numpy(np.ones(n))
    bh.mul(dw, BA, dv, 1.0, 1.0)
    assert np.array_equal(dw.numpy(), M @ v + 1)
    X = np.asfortranarray(np.arange(1, 101, dtype=np.float64).reshape(10, 10, order="F"))
    dX = bh.DeviceArray.from_numpy(X)
    dY = bh.DeviceArray.from_numpy(np.full((10, 10), 2.0))
    bh.mul(dY, BA, dX, 0.0, 1.0)
    assert np.array_equal(dY.numpy(), np.full((10, 10), 2.0))
    dY = bh.DeviceArray.from_numpy(np.ones((10, 10)))
    bh.mul(dY, BA, dX, 1.0, 1.0)
    assert np.array_equal(dY.numpy(), M @ X + 1)


def test_gbmv_negative_bands(bh):
    """_banded_muladd! with bands entirely above / below the diagonal
    (matmul.jl:41-59) against the dense product."""
    rng = np.random.default_rng(5)
    for (l, u) in [(-1, 2), (2, -1), (-2, 3), (3, -2)]:
        n = 40
        A = rng.standard_normal((l + u + 1, n))
        M = np.zeros((n, n))
        for j in range(n):
            for t in range(l + u + 1):
                i = j - u + t
                if 0 <= i < n:
                    M[i, j] = A[t, j]
        BA = bh.BandedMatrix.from_host(np.asfortranarray(A), n, l, u)
        x = rng.standard_normal(n)
        y = np.full(n, np.nan)
        dy = bh.DeviceArray.from_numpy(y)
        bh.mul(dy, BA, bh.DeviceArray.from_numpy(x), 1.0, 0.0)
        np.testing.assert_allclose(dy.numpy(), M @ x, rtol=1e-13, atol=1e-13)


GBMM_CASES = [
    # n, k, m, (Al,Au), (Bl,Bu)
    (10000, 10000, 10000, (4, 3), (4, 3)),
    (5000, 5000, 5000, (8, 8), (4, 4)),
    (50, 50, 50, (2, 30), (30, 1)),
    (50, 5, 50, (1, 2), (0, 30)),
    (5, 50, 1, (2, 1), (30, 30)),
    (1, 1, 1, (0, 0), (0, 0)),
    (300, 200, 250, (5, 0), (0, 7)),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,k,m,ab,bb", GBMM_CASES)
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (2.0, -0.5)])
def test_gbmm_bitexact(bh, oracle, dtype, n, k, m, ab, bb, alpha, beta):
    rng = np.random.default_rng(n + k + m)
    (Al, Au), (Bl, Bu) = ab, bb
    A = band(rng, Al, Au, k, dtype, dist="n", m=n)
    B = band(rng, Bl, Bu, m, dtype, dist="n", m=k)
    Cl, Cu = min(n - 1, Al + Bl), min(m - 1, Au + Bu)
    Cd = band(rng, Cl, Cu, m, dtype, dist="n", m=n)
    ref, st = oracle.gbmm(n, k, m, alpha, A.astype(np.float64), Al, Au, B.astype(np.float64), Bl, Bu, beta,
                          Cd.astype(np.float64), Cl, Cu)
    assert st == 0
    dA = bh.BandedMatrix.from_host(A, n, Al, Au)
    dB = bh.BandedMatrix.from_host(B, k, Bl, Bu)
    dC = bh.BandedMatrix.from_host(Cd, n, Cl, Cu)
    bh.gbmm(dA, dB, dC, alpha, beta)
    out = dC.host_data()
    if dtype == np.float64:
        assert bits_equal(out, ref), np.nanmax(np.abs(out - ref))
    else:
        # the oracle computes in FP64; check the FP32 result against the dense product
        M = oracle.band_to_dense(A, n, k, Al, Au).astype(np.float64) @ oracle.band_to_dense(B, k, m, Bl, Bu)
        Cdn = alpha * M + beta * oracle.band_to_dense(Cd, n, m, Cl, Cu)
        got = oracle.band_to_dense(out, n, m, Cl, Cu)
        assert np.abs(got - Cdn).max() <= 1e-4 * (np.abs(Cdn).max() + 1)


def test_gbmm_wider_c_and_banderror(bh, oracle):
    """C wider than the product (extra bands beta-scaled, gbmm.jl:286-288) and
    the BandError / zero-band stripping rules (test_linalg.jl:272-295)."""
    rng = np.random.default_rng(11)
    n = 60
    A = band(rng, 2, 1, n, dist="n")
    B = band(rng, 1, 2, n, dist="n")
    Cd = band(rng, 5, 6, n, dist="n")
    ref, st = oracle.gbmm(n, n, n, 1.0, A, 2, 1, B, 1, 2, 0.5, Cd, 5, 6)
    assert st == 0
    dC = bh.BandedMatrix.from_host(Cd, n, 5, 6)
    bh.gbmm(bh.BandedMatrix.from_host(A, n, 2, 1), bh.BandedMatrix.from_host(B, n, 1, 2), dC, 1.0, 0.5)
    assert bits_equal(dC.host_data(), ref)
    # diagonal-only A times tridiagonal B into tridiagonal C: zero bands stripped
    A2 = np.zeros((3, 10))
    A2[1, :] = rng.standard_normal(10)
    B2 = band(rng, 1, 1, 10, dist="n")
    C2 = np.zeros((3, 10))
    ref2, st2 = oracle.gbmm(10, 10, 10, 1.0, A2, 1, 1, B2, 1, 1, 0.0, C2, 1, 1)
    assert st2 == 0
    dC2 = bh.BandedMatrix.from_host(C2, 10, 1, 1)
    bh.gbmm(bh.BandedMatrix.from_host(A2, 10, 1, 1), bh.BandedMatrix.from_host(B2, 10, 1, 1), dC2)
    assert bits_equal(dC2.host_data(), ref2)
    A2[0, 3] = 1.0  # a nonzero on band 1 -> BandError
    _, st3 = oracle.gbmm(10, 10, 10, 1.0, A2, 1, 1, B2, 1, 1, 0.0, C2, 1, 1)
    assert st3 == -1
    with pytest.raises(bh.BandError):
        bh.gbmm(bh.BandedMatrix.from_host(A2, 10, 1, 1), bh.BandedMatrix.from_host(B2, 10, 1, 1),
                bh.BandedMatrix.from_host(C2, 10, 1, 1))


# The persistent wave-pipelined single-vector kernel (beta = 0, m <= n,
# kl + ku <= 64, contiguous aligned storage): several tiles per wave, the
# XCD split with fewer tiles than waves, the widest band it takes, and each
# pipeline depth.  BANDEDHIP_GBMV_WAVE picks the depth (1 -> 2 stages).
WAVE_CASES = [
    (1 << 18, 1 << 18, 8, 8),
    (200003, 200003, 40, 24),
    (100000, 100001, 32, 32),
    (64 * 8 * 3 + 5, 64 * 8 * 3 + 5, 1, 0),
    (99999, 150000, 5, 7),
    (70, 70, 3, 3),
]


@pytest.mark.parametrize("stages", ["1", "3", "4"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,kl,ku", WAVE_CASES)
def test_gbmv_wave_pipeline_bitexact(bh, oracle, monkeypatch, stages, dtype, m, n, kl, ku):
    monkeypatch.setenv("BANDEDHIP_GBMV_WAVE", stages)
    rng = np.random.default_rng(m + 31 * kl + ku)
    A = band(rng, kl, ku, n, dtype, dist="n", nan_corners=True, m=m)
    x = rng.standard_normal(n).astype(dtype)
    y = np.full(m, np.nan, dtype)
    ref = oracle.gbmv("N", m, n, kl, ku, dtype(-0.75), A, x, dtype(0.0), y)
    BA = bh.BandedMatrix.from_host(A, m, kl, ku)
    dx, dy = bh.DeviceArray.from_numpy(x), bh.DeviceArray.from_numpy(y)
    bh.mul(dy, BA, dx, -0.75, 0.0)
    out = dy.numpy()
    assert bits_equal(out, ref), np.abs(out - ref).max()
