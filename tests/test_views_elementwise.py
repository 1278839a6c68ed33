"""Banded views and negative bandwidths (src/banded/BandedMatrix.jl:918-954,
src/generic/matmul.jl:41-59, src/banded/gbmm.jl:231-279), and the
element-wise fast paths on band data: lmul!/rmul!, axpy!, copyto! and the
+/- built on them (src/generic/broadcast.jl:942-1020, BandedMatrix.jl:124-165).

CPU: the view geometry (bandwidths and band rows of view(A, kr, jr)) against
brute force on dense matrices: every entry of the dense submatrix must sit
in the view's bands at the slot the geometry names, and nothing outside them.
GPU (C ABI through the mirror): views against dense submatrices and against
the oracle's gbmv / gbtf2 on the same band rows (bit for bit); banded products
with negative bandwidths against exact integer dense products; scale, axpy and
copy bit for bit against numpy on the dense entries (a*x rounded, then added:
the reference's a*X + Y), BandError raised before the destination changes.
"""
import itertools

import numpy as np
import pytest

from _util import bits_equal


def dense_band(rng, m, n, l, u, integer=False):
    M = np.zeros((m, n))
    for i in range(m):
        for j in range(n):
            if -l <= j - i <= u:
                M[i, j] = rng.integers(-4, 5) if integer else rng.standard_normal()
    return M


def band_of(M, l, u):
    m, n = M.shape
    D = np.zeros((max(0, l + u + 1), n), order="F")
    for j in range(n):
        for i in range(m):
            if -l <= j - i <= u:
                D[u + i - j, j] = M[i, j]
    return D


# ------------------------------------------------------------------ CPU
def test_view_geometry_brute_force():
    from banded_hip.bandedmatrix import view_geometry

    rng = np.random.default_rng(0)
    for m, n, l, u in [(12, 15, 2, 3), (20, 9, 4, 0), (9, 20, 0, 5), (14, 14, 3, 3), (10, 10, 1, 1)]:
        M = dense_band(rng, m, n, l, u)
        D = band_of(M, l, u)
        for k0, k1, j0, j1 in itertools.product(range(0, m, 3), range(1, m + 1, 4), range(0, n, 3),
                                                range(1, n + 1, 4)):
            if k1 <= k0 or j1 <= j0:
                continue
            kr, jr = range(k0, k1), range(j0, j1)
            lv, uv, row0, h = view_geometry(l, u, kr, jr)
            S = M[k0:k1, j0:j1]
            assert h == max(0, lv + uv + 1)
            for i in range(S.shape[0]):
                for j in range(S.shape[1]):
                    inside = -lv <= j - i <= uv
                    if S[i, j] != 0:
                        assert inside, (kr, jr, i, j)
                    if inside:  # the view's slot (uv + i - j, j) is A's slot (row0 + uv + i - j, j0 + j)
                        assert D[row0 + uv + i - j, j0 + j] == S[i, j]


def test_view_geometry_rejects_steps():
    from banded_hip.bandedmatrix import view_geometry

    with pytest.raises(ValueError):
        view_geometry(1, 1, range(0, 10, 2), range(0, 10))


def test_argument_errors():
    from banded_hip import _lib

    L = _lib.load()
    assert L.bh_dgbscal(-1, 4, 1, 1, 2.0, None, 3, 0, 0, 0, 4, 0) == -1
    assert L.bh_dgbscal(4, 4, 1, 1, 2.0, None, 2, 0, 4, 0, 4, 0) == -7
    assert L.bh_dgbscal(4, 4, 1, 1, 2.0, None, 3, 0, 5, 0, 4, 0) == -8
    assert L.bh_dgbscal(4, 4, 1, 1, 2.0, None, 3, 0, 4, 3, 2, 0) == -10
    assert L.bh_dgbaxpy(4, 4, 1.0, None, 2, 1, 1, None, 3, 1, 1) == -5
    assert L.bh_dgbcopy(4, 4, None, 3, 1, 1, None, 1, 1, 1) == -8


# ------------------------------------------------------------------ GPU
VIEW_CASES = [(range(0, 30), range(0, 40)), (range(5, 25), range(0, 30)), (range(0, 20), range(7, 37)),
              (range(10, 30), range(10, 30)), (range(3, 4), range(0, 40)), (range(0, 30), range(33, 40))]


@pytest.mark.gpu
@pytest.mark.parametrize("kr,jr", VIEW_CASES)
def test_gpu_view_mul_and_dense(bh, oracle, kr, jr):
    rng = np.random.default_rng(1)
    m, n, l, u = 30, 40, 3, 4
    M = dense_band(rng, m, n, l, u)
    A = bh.BandedMatrix.from_host(band_of(M, l, u), m, l, u)
    V = A.view(kr, jr)
    assert np.array_equal(V.to_dense(), M[kr.start:kr.stop, jr.start:jr.stop])
    x = rng.standard_normal(len(jr))
    y = bh.mul(bh.DeviceArray((len(kr),), np.float64), V, bh.DeviceArray.from_numpy(x)).numpy()
    S = M[kr.start:kr.stop, jr.start:jr.stop]
    assert np.allclose(y, S @ x, rtol=1e-13, atol=1e-13)
    if V.l >= 0 and V.u >= 0:
        ref = oracle.gbmv("N", len(kr), len(jr), V.l, V.u, 1.0, np.asfortranarray(V.host_data()), x)
        assert bits_equal(y, ref)


@pytest.mark.gpu
def test_gpu_lu_of_a_view(bh, oracle):
    rng = np.random.default_rng(2)
    n, l, u = 50, 3, 2
    M = dense_band(rng, n, n, l, u) + np.eye(n) * 0.1
    A = bh.BandedMatrix.from_host(band_of(M, l, u), n, l, u)
    V = A.view(range(10, 40), range(10, 40))  # square, same bands
    assert (V.l, V.u) == (l, u)
    F = bh.lu(V)
    Dv = np.asfortranarray(V.host_data())
    f, piv, info = oracle.gbtf2(30, 30, l, u, oracle.pad_for_lu(Dv, 30, 30, l, u))
    assert F.info == info and np.array_equal(F.ipiv, piv)
    assert bits_equal(F.factors.host_data(), f)


NEG_CASES = [((2, -1), (2, 2)), ((-1, 3), (1, 1)), ((1, 1), (-1, 2)), ((1, 2), (3, -1)), ((-2, 3), (3, -1)),
             ((3, -2), (-1, 3)), ((-3, 1), (1, 1)), ((1, 1), (1, 1))]


@pytest.mark.gpu
@pytest.mark.parametrize("ab,bb", NEG_CASES)
def test_gpu_gbmm_negative_bands_exact(bh, ab, bb):
    rng = np.random.default_rng(3)
    n = 23
    (al, au), (bl, bu) = ab, bb
    MA = dense_band(rng, n, n, al, au, integer=True)
    MB = dense_band(rng, n, n, bl, bu, integer=True)
    A = bh.BandedMatrix.from_host(band_of(MA, al, au), n, al, au)
    B = bh.BandedMatrix.from_host(band_of(MB, bl, bu), n, bl, bu)
    C = bh.gbmm(A, B)
    assert np.array_equal(C.to_dense(), MA @ MB), (ab, bb)
    # with an existing C: C = 3 A B + 2 C_old on C's bands
    cl, cu = max(0, al + bl) + 1, max(0, au + bu) + 1
    MC = dense_band(rng, n, n, cl, cu, integer=True)
    C2 = bh.BandedMatrix.from_host(band_of(MC, cl, cu), n, cl, cu)
    bh.gbmm(A, B, C2, 3.0, 2.0)
    assert np.array_equal(C2.to_dense(), 3 * (MA @ MB) + 2 * MC), (ab, bb)


@pytest.mark.gpu
def test_gpu_lmul_rmul_bitexact(bh):
    rng = np.random.default_rng(4)
    m, n, l, u = 300, 280, 5, 7
    M = dense_band(rng, m, n, l, u)
    A = bh.BandedMatrix.from_host(band_of(M, l, u), m, l, u)
    A.lmul_(1.7)
    assert bits_equal(A.to_dense(), 1.7 * M)
    B = -A * 0.5
    # values equal (outside the band numpy's dense -0.0 compares equal to 0)
    assert np.array_equal(B.to_dense(), (1.7 * M * -1.0) * 0.5)


@pytest.mark.gpu
def test_gpu_scale_block_strong_zero(bh):
    from banded_hip.bandedmatrix import _scale_block

    rng = np.random.default_rng(5)
    m, n, l, u = 40, 40, 3, 3
    D = band_of(dense_band(rng, m, n, l, u), l, u)
    D[:, 10:20] = np.nan  # inside the block: beta == 0 must overwrite NaN
    A = bh.BandedMatrix.from_host(D, m, l, u)
    _scale_block(A, 0, m, 10, 20, 0.0)
    got = A.host_data()
    assert not np.isnan(got[:, 10:20]).any() and np.all(got[:, 10:20] == 0)
    assert bits_equal(got[:, :10], D[:, :10]) and bits_equal(got[:, 20:], D[:, 20:])


AXPY_CASES = [((2, 3), (2, 3)), ((1, 1), (3, 4)), ((0, 2), (2, 0)), ((3, 0), (1, 5)), ((-1, 2), (1, 3))]


@pytest.mark.gpu
@pytest.mark.parametrize("xb,yb", AXPY_CASES)
def test_gpu_axpy_bitexact(bh, xb, yb):
    rng = np.random.default_rng(6)
    m, n = 200, 230
    (xl, xu), (yl, yu) = xb, yb
    # X lives only on the bands it shares with Y, so no BandError
    sl, su = min(xl, yl), min(xu, yu)
    MX = dense_band(rng, m, n, sl, su)
    MY = dense_band(rng, m, n, yl, yu)
    X = bh.BandedMatrix.from_host(band_of(MX, xl, xu), m, xl, xu)
    Y = bh.BandedMatrix.from_host(band_of(MY, yl, yu), m, yl, yu)
    bh.axpy_(-0.3, X, Y)
    assert bits_equal(Y.to_dense(), -0.3 * MX + MY)


@pytest.mark.gpu
def test_gpu_axpy_band_error_leaves_y(bh):
    rng = np.random.default_rng(7)
    m = n = 50
    MX = dense_band(rng, m, n, 2, 2)
    MY = dense_band(rng, m, n, 1, 1)
    X = bh.BandedMatrix.from_host(band_of(MX, 2, 2), m, 2, 2)
    Y = bh.BandedMatrix.from_host(band_of(MY, 1, 1), m, 1, 1)
    before = Y.host_data().copy()
    with pytest.raises(bh.BandError):
        bh.axpy_(1.0, X, Y)
    assert bits_equal(Y.host_data(), before)


@pytest.mark.gpu
def test_gpu_copy_convert_and_plus(bh):
    rng = np.random.default_rng(8)
    m, n = 120, 100
    MA = dense_band(rng, m, n, 2, 3)
    MB = dense_band(rng, m, n, 4, 1)
    A = bh.BandedMatrix.from_host(band_of(MA, 2, 3), m, 2, 3)
    B = bh.BandedMatrix.from_host(band_of(MB, 4, 1), m, 4, 1)
    W = bh.convert(A, (5, 6))
    assert (W.l, W.u) == (5, 6) and bits_equal(W.to_dense(), MA)
    assert bits_equal(W.host_data()[3:9][:, 10:-10], band_of(MA, 2, 3)[:, 10:-10])  # A's rows sit at u'-u..
    with pytest.raises(bh.BandError):
        bh.convert(A, (2, 2))
    S = A + B
    assert (S.l, S.u) == (4, 3) and bits_equal(S.to_dense(), MA + MB)
    Dm = A - B
    assert bits_equal(Dm.to_dense(), MA - MB)


@pytest.mark.gpu
def test_gpu_elementwise_large(bh):
    """n = 2^20, bands (8, 8): axpy and scale keep sums exact on integer data."""
    n, l, u = 1 << 20, 8, 8
    rng = np.random.default_rng(9)
    D = np.asfortranarray(rng.integers(-8, 9, size=(l + u + 1, n)).astype(np.float64))
    X = bh.BandedMatrix.from_host(D, n, l, u)
    Y = bh.BandedMatrix.from_host(D, n, l, u)
    bh.axpy_(2.0, X, Y)
    Y.lmul_(0.5)
    got = Y.host_data()
    inside = np.ones_like(D, bool)
    for t in range(l + u + 1):  # corner slots are outside the matrix and untouched
        j = np.arange(n)
        i = j - u + t
        inside[t] = (i >= 0) & (i < n)
    assert np.array_equal(got[inside], 1.5 * D[inside])
    assert np.array_equal(got[~inside], D[~inside])
