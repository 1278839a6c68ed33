"""GPU: gradients of L = w' A^{-1} b + 1/2 ||A x0||^2 (the differentiation form).

The fused band kernel must equal the oracle's band_grad bit for bit on the
same vectors; the full pullback (LU, 'N' solve, gbmv, 'T' solve with the same
factors, band kernel) must agree with the CPU gradient assembled from the
oracle (1e-11 relative; the 'T' solve's dot products are summed in a
different order) and with central finite differences on random band entries
and right-hand-side entries."""
import numpy as np
import pytest

from _util import band, bits_equal

pytestmark = pytest.mark.gpu


def _problem(seed, n, l, u, shift):
    rng = np.random.default_rng(seed)
    A = band(rng, l, u, n, dist="n")
    A[u, :] += shift
    b, w, x0 = (rng.standard_normal(n) for _ in range(3))
    return A, b, w, x0


def _cpu_grad(oracle, A, b, w, x0, n, l, u):
    f, piv, info = oracle.gbtf2(n, n, l, u, oracle.pad_for_lu(A, n, n, l, u))
    x = oracle.gbtrs("N", n, l, u, f, piv, b)
    lam = oracle.gbtrs("T", n, l, u, f, piv, w)
    y0 = oracle.gbmv("N", n, n, l, u, 1.0, A, x0)
    G = oracle.band_grad(n, l, u, lam, x, y0, x0)
    return float(w @ x + 0.5 * y0 @ y0), G, lam


def test_band_grad_kernel_bits(bh, oracle):
    n, l, u = 3000, 5, 7
    rng = np.random.default_rng(1)
    lam, x, y0, x0 = (rng.standard_normal(n) for _ in range(4))
    G = bh.BandedMatrix.zeros((n, n), (l, u))
    from banded_hip import _lib

    dv = [bh.DeviceArray.from_numpy(v) for v in (lam, x, y0, x0)]
    _lib.call("bh_dgb_band_grad", n, n, l, u, *[d.ptr for d in dv], G.data.ptr, G.data.ld)
    assert bits_equal(G.host_data(), oracle.band_grad(n, l, u, lam, x, y0, x0))


def test_band_outer_terms(bh, oracle):
    m, n, l, u = 700, 500, 3, 4
    rng = np.random.default_rng(2)
    u1, u2 = rng.standard_normal(m), rng.standard_normal(m)
    v1, v2 = rng.standard_normal(n), rng.standard_normal(n)
    G0 = band(rng, l, u, n, dist="n", m=m)
    G = bh.BandedMatrix.from_host(G0, m, l, u)
    d = [bh.DeviceArray.from_numpy(v) for v in (u1, v1, u2, v2)]
    bh.band_outer_(G, -1.5, d[0], d[1], 0.5, d[2], d[3], beta=2.0)
    want = G0.copy()
    for j in range(n):
        for t in range(l + u + 1):
            i = j - u + t
            if 0 <= i < m:
                want[t, j] = 2.0 * G0[t, j] + (-1.5 * u1[i]) * v1[j] + (0.5 * u2[i]) * v2[j]
    got = G.host_data()
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()


@pytest.mark.parametrize("n,l,u,shift", [(3000, 16, 16, 64.0), (2000, 4, 3, 10.0), (262144, 16, 16, 64.0)])
def test_loss_grad_matches_cpu_gradient(bh, oracle, n, l, u, shift):
    A, b, w, x0 = _problem(n + l, n, l, u, shift)
    dA_ = bh.BandedMatrix.from_host(A, n, l, u)
    L, dA, db, x = bh.loss_and_grad(dA_, *(bh.DeviceArray.from_numpy(v) for v in (b, w, x0)))
    Lc, Gc, lamc = _cpu_grad(oracle, A, b, w, x0, n, l, u)
    assert abs(L - Lc) <= 1e-11 * abs(Lc)
    g = dA.host_data()
    mask = np.isfinite(Gc)
    assert np.abs(g - Gc)[mask].max() <= 1e-11 * np.abs(Gc)[mask].max()
    assert np.abs(db.numpy() - lamc).max() <= 1e-11 * np.abs(lamc).max()


def test_loss_grad_finite_differences(bh, oracle):
    """Central differences (L(a+h) - L(a-h)) / 2h on 64 band entries and 8
    entries of b.  The difference of the two losses is formed from the
    elementwise differences of x = A\\b and y0 = A x0 (w'(x+ - x-) and
    (y+ - y-)'(y+ + y-)/2), which avoids the cancellation of subtracting two
    O(10^6) loss values; it is still a pure finite difference of the forward map."""
    n, l, u = 2000, 16, 16
    A, b, w, x0 = _problem(7, n, l, u, 64.0)
    dev = lambda v: bh.DeviceArray.from_numpy(v)

    def fwd(Ah, bh_):
        F = bh.lu(bh.BandedMatrix.from_host(Ah, n, l, u))
        x = bh.ldiv_(F, dev(bh_)).numpy()
        y0 = oracle.gbmv("N", n, n, l, u, 1.0, Ah, x0)
        return x, y0

    def fd(p_, m_, h):
        (xp, yp), (xm, ym) = p_, m_
        return (w @ (xp - xm) + 0.5 * (yp - ym) @ (yp + ym)) / (2 * h)

    L, dA, db, _ = bh.loss_and_grad(bh.BandedMatrix.from_host(A, n, l, u), dev(b), dev(w), dev(x0))
    g, gb = dA.host_data(), db.numpy()
    rng = np.random.default_rng(11)
    for _ in range(64):
        j = int(rng.integers(0, n))
        t = int(rng.integers(max(0, u - j), min(l + u, n - 1 - j + u) + 1))
        h = 1e-6 * max(1.0, abs(A[t, j]))
        Ap, Am = A.copy(), A.copy()
        Ap[t, j] += h
        Am[t, j] -= h
        d = fd(fwd(Ap, b), fwd(Am, b), h)
        assert abs(d - g[t, j]) <= 1e-6 * max(1.0, abs(g[t, j])), (t, j, d, g[t, j])
    for i in rng.integers(0, n, 8):
        h = 1e-6 * max(1.0, abs(b[i]))
        bp, bm = b.copy(), b.copy()
        bp[i] += h
        bm[i] -= h
        d = fd(fwd(A, bp), fwd(A, bm), h)
        assert abs(d - gb[i]) <= 1e-6 * max(1.0, abs(gb[i])), (i, d, gb[i])
