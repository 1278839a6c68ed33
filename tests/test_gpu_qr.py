"""GPU parity: banded Householder QR (bandedqr.jl) against the oracle.

The factorisation follows _banded_qr! / ArrayLayouts.reflector! in the
reference's operation order (sequential norms and dot products, divisions by
x1 + copysign(||x||, x1)), so FP64 factors and tau must equal the oracle's
restatement bit for bit.  Q'B and QB use a tree sum for each reflector's dot
product: within 1e-13 of the oracle (relative, max norm).  The QR solve is
checked against the oracle and through the normal equations."""
import numpy as np
import pytest

from _util import band, bits_equal

pytestmark = pytest.mark.gpu

QR_CASES = [
    # m, n, l, u
    (1, 1, 0, 0),
    (5, 5, 1, 1),
    (10, 10, 2, 1),
    (300, 300, 4, 3),
    (10000, 10000, 4, 3),     # the plain CPU-runnable config
    (400, 300, 3, 5),         # tall
    (300, 400, 3, 5),         # wide (factor and Q only)
    (500, 500, 0, 4),
    (500, 500, 6, 0),
    (600, 600, 15, 10),       # 16-row register window
    (700, 700, 31, 20),       # 32-row window, u' = 51
    (300, 300, 40, 10),       # l+1 > 32: one-workgroup path
    (300, 300, 5, 60),        # u' + 1 > 64: one-workgroup path
]


def _make(rng, m, n, l, u, dtype=np.float64):
    A = band(rng, l, u, n, np.float64, dist="n", m=m)
    return A.astype(dtype)


def _padded(oracle, A, m, n, l, u):
    return oracle.pad_for_lu(A, m, n, l, u)


def _wave_path(l, u):
    """The one-wave factorisation (sequential sums in the reference's order)
    covers l+1 <= 32 and padded u' = l+u with u'+1 <= 64."""
    return l + 1 <= 32 and l + u + 1 <= 64


WIDE_CASES = [(500, 500, 100, 100), (400, 300, 70, 30), (300, 400, 70, 30), (600, 600, 300, 50), (257, 257, 64, 1)]


@pytest.mark.parametrize("m,n,l,u", QR_CASES + WIDE_CASES)
def test_gbqrf_factors(bh, oracle, m, n, l, u):
    """Bit-exact on the one-wave path.  On the wide path (wave-parallel norms
    and dot products) the reflectors differ in rounding, which the later
    columns of an ill-conditioned random band amplify: there Q R must equal A
    to 1e-13 (relative, max norm) and |diag R| the oracle's to 1e-10."""
    rng = np.random.default_rng(m * 7 + n + l * 31 + u)
    A = _make(rng, m, n, l, u)
    F = bh.qr(bh.BandedMatrix.from_host(A, m, l, u))
    D, tau = oracle.gbqrf(m, n, l, l + u, _padded(oracle, A, m, n, l, u))
    g = F.factors.host_data()
    t = F.tau.numpy()[: min(m, n)]
    if _wave_path(l, u):
        assert bits_equal(g, D), np.nanmax(np.abs(g - D))
        assert bits_equal(t, tau)
    else:
        # Q R = A to backward-error accuracy, with Q applied from the GPU's
        # own reflectors by the oracle; R's |diagonal| as the oracle's
        up = l + u
        R = np.zeros((m, n))
        for j in range(n):
            for i in range(max(0, j - up), min(j, m - 1) + 1):
                R[i, j] = g[up + i - j, j]
        QR = oracle.gbqr_qmul(m, n, l, up, g, t, R)
        Ad = oracle.band_to_dense(A, m, n, l, u)
        err = np.abs(QR - Ad).max() / np.abs(Ad).max()
        assert err <= 1e-13, err
        Ro = np.array([D[up, j] for j in range(min(m, n))])
        assert np.abs(np.abs(np.diag(R)[: min(m, n)]) - np.abs(Ro)).max() <= 1e-10 * np.abs(Ro).max()


@pytest.mark.parametrize("m,n,l,u", [c for c in QR_CASES + WIDE_CASES if c[0] <= 2000])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_q_apply(bh, oracle, m, n, l, u, nrhs):
    rng = np.random.default_rng(m + 3 * n + 5 * l + u + nrhs)
    A = _make(rng, m, n, l, u)
    F = bh.qr(bh.BandedMatrix.from_host(A, m, l, u))
    # the oracle applies the GPU's own reflectors (on the wide path they differ
    # from the oracle's factorisation in rounding)
    D, tau = F.factors.host_data(), F.tau.numpy()[: min(m, n)]
    B = np.asfortranarray(rng.standard_normal((m, nrhs)))
    for adj, ref in ((True, oracle.gbqr_qtmul), (False, oracle.gbqr_qmul)):
        Q = F.Q.T if adj else F.Q
        got = (Q @ bh.DeviceArray.from_numpy(B)).numpy()
        want = ref(m, n, l, l + u, D, tau, B)
        assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max())
    # Q'(Q B) = B
    dB = bh.DeviceArray.from_numpy(B)
    F.Q.lmul_(dB)
    F.Q.T.lmul_(dB)
    assert np.abs(dB.numpy() - B).max() <= 1e-13 * max(1.0, np.abs(B).max())


@pytest.mark.parametrize("m,n,l,u", [c for c in QR_CASES + WIDE_CASES if c[0] >= c[1]])
def test_qr_solve(bh, oracle, m, n, l, u):
    rng = np.random.default_rng(11 * m + n + l + 2 * u)
    A = _make(rng, m, n, l, u)
    A[u, :] += 3.0 * (l + u + 1)  # well conditioned
    B = np.asfortranarray(rng.standard_normal((m, 2)))
    F = bh.qr(bh.BandedMatrix.from_host(A, m, l, u))
    X = bh.qr_ldiv_(F, bh.DeviceArray.from_numpy(B)).numpy()[:n]
    D, tau = F.factors.host_data(), F.tau.numpy()[: min(m, n)]  # the GPU's factors, solved by the oracle
    Xo = oracle.gbqrs(m, n, l, l + u, D, tau, B)[:n]
    assert np.abs(X - Xo).max() <= 1e-12 * max(1.0, np.abs(Xo).max())
    # normal equations: A'(A x - b) = 0 (least squares; exact solve when square)
    Ad = oracle.band_to_dense(A, m, n, l, u)
    r = Ad @ X - B
    assert np.abs(Ad.T @ r).max() <= 1e-9 * np.abs(Ad).sum(axis=0).max() * max(1.0, np.abs(X).max())


@pytest.mark.parametrize("m,n,l,u", [(20000, 20000, 3, 4), (400, 300, 3, 5), (700, 700, 31, 20), (500, 500, 6, 0),
                                     (300, 300, 40, 30)])
def test_qr_solve_partitioned(bh, oracle, m, n, l, u):
    """qr_ldiv_(F, B, method="partitioned"): Q'B by the reflectors, then R's
    solve by the partitioned band solve (kl = 0, R's bandwidth l+u <= 63);
    within 1e-12 of the oracle's gbqrs on the GPU's factors.  R wider than 63
    (the last case, l+u = 70) falls back to the exact solve, bit for bit."""
    rng = np.random.default_rng(5 * m + n + l + u)
    A = _make(rng, m, n, l, u)
    A[u, :] += 3.0 * (l + u + 1)
    B = np.asfortranarray(rng.standard_normal((m, 2)))
    F = bh.qr(bh.BandedMatrix.from_host(A, m, l, u))
    X = bh.qr_ldiv_(F, bh.DeviceArray.from_numpy(B), method="partitioned").numpy()[:n]
    D, tau = F.factors.host_data(), F.tau.numpy()[: min(m, n)]
    Xo = oracle.gbqrs(m, n, l, l + u, D, tau, B)[:n]
    assert np.abs(X - Xo).max() <= 1e-12 * max(1.0, np.abs(Xo).max())
    if l + u > 63:
        Xe = bh.qr_ldiv_(F, bh.DeviceArray.from_numpy(B)).numpy()[:n]
        assert bits_equal(X, Xe)
    with pytest.raises(ValueError):
        bh.qr_ldiv_(F, bh.DeviceArray.from_numpy(B), method="nope")


def test_qr_fp32_close_to_fp64(bh, oracle):
    m = n = 2000
    l, u = 4, 3
    rng = np.random.default_rng(5)
    A = _make(rng, m, n, l, u)
    F = bh.qr(bh.BandedMatrix.from_host(A.astype(np.float32), m, l, u))
    D, tau = oracle.gbqrf(m, n, l, l + u, _padded(oracle, A, m, n, l, u))
    g = F.factors.host_data().astype(np.float64)
    assert np.abs(g - D).max() <= 1e-4 * np.abs(D).max()


def test_qr_wide_solve_is_an_error(bh):
    rng = np.random.default_rng(1)
    A = _make(rng, 30, 40, 2, 2)
    F = bh.qr(bh.BandedMatrix.from_host(A, 30, 2, 2))
    with pytest.raises(ValueError):
        bh.qr_ldiv_(F, bh.DeviceArray.from_numpy(np.ones(30)))


RMUL_CASES = [
    # mA, m, n, l, u
    (7, 300, 300, 4, 3),
    (50, 400, 300, 3, 5),
    (33, 300, 400, 3, 5),
    (20, 300, 300, 40, 10),
    (9, 200, 200, 70, 10),  # l + 1 > 64
]


@pytest.mark.parametrize("mA,m,n,l,u", RMUL_CASES)
@pytest.mark.parametrize("trans", ["N", "T"])
def test_q_rmul(bh, oracle, mA, m, n, l, u, trans):
    """A Q and A Q' (banded_qr_rmul!): bit-identical to the oracle's restatement
    (the same sequential sums per row), and equal to the dense product."""
    rng = np.random.default_rng(mA + m + 7 * l + u)
    A = _make(rng, m, n, l, u)
    F = bh.qr(bh.BandedMatrix.from_host(A, m, l, u))
    D = F.factors.host_data()
    tau = F.tau.numpy()[: min(m, n)]
    X = np.asfortranarray(rng.standard_normal((mA, m)))
    ref = oracle.gbqr_rmul(trans, m, n, l, l + u, D, tau, X)
    dX = bh.DeviceArray.from_numpy(X)
    Q = F.Q if trans == "N" else F.Q.T
    bh.qr_rmul_(dX, Q)
    out = dX.numpy()
    assert bits_equal(out, ref), np.abs(out - ref).max()
    Qd = oracle.gbqr_qmul(m, n, l, l + u, D, tau, np.eye(m))
    dense = X @ (Qd if trans == "N" else Qd.T)
    assert np.abs(out - dense).max() <= 1e-12 * (np.abs(dense).max() + 1)


@pytest.mark.parametrize("n,l,u", [(2000, 400, 400), (4096, 128, 128), (1500, 200, 150)])
def test_qr_solve_wide_bands(bh, oracle, n, l, u):
    """qr(A) \\ b where the R band (l+u super-diagonals) is beyond the lane-ring
    solve: the reference's own benchmark shape (2000, (400,400)) among them."""
    rng = np.random.default_rng(n + l)
    A = _make(rng, n, n, l, u)
    B = np.asfortranarray(rng.standard_normal((n, 2)))
    F = bh.qr(bh.BandedMatrix.from_host(A, n, l, u))
    dB = bh.DeviceArray.from_numpy(B)
    bh.qr_ldiv_(F, dB)
    x = dB.numpy()
    D, tau = F.factors.host_data(), F.tau.numpy()[:n]  # the GPU's factors, solved by the oracle
    xref = oracle.gbqrs(n, n, l, l + u, D, tau, B)
    assert np.abs(x - xref).max() <= 1e-10 * (np.abs(xref).max() + 1)
    for c in range(2):
        r = B[:, c] - oracle.gbmv("N", n, n, l, u, 1.0, A, x[:, c])
        berr = np.abs(r).max() / (np.abs(A).sum(axis=0).max() * np.abs(x[:, c]).max() + np.abs(B[:, c]).max())
        assert berr <= 1e-13, berr
