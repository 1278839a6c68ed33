"""GPU: the blocked band solve (bh_?gbtrs_blocked) with the factors of
bh_?gbtrf, against the oracle's xGBTRS restatement on the same factors.

The blocked solve applies each 32-row panel through the inverses of its
diagonal blocks and sums a panel's updates at once, so it equals xGBTRS to
rounding, not bit for bit.  Bars: the solution within 1e-11 (FP64) / 2e-3
(FP32) of the oracle's solve with the same factors (max norm, relative), and
the FP64 normwise backward error ||b - A x|| / (||A|| ||x|| + ||b||) within
1e-14.  Cases cover the MFMA, blocked and lane LU kernels' factors, pivoting
(no diagonal shift), partial last panels, n below one window, more than 16
right-hand sides (two workgroups, the second ragged), and the edges."""
import ctypes as C

import numpy as np
import pytest

from _util import band

pytestmark = pytest.mark.gpu

CASES = [
    (4096, 128, 128, 16, True),
    (4096, 128, 128, 16, False),
    (3000, 40, 30, 5, False),
    (777, 17, 3, 20, False),
    (1000, 64, 200, 3, True),
    (50, 16, 16, 3, False),
    (33, 1, 0, 1, False),
    (129, 5, 70, 2, False),
    (2048, 384, 32, 4, True),
    (2048, 100, 150, 4, False),
]


def _solve(bh, dt, A, B, n, kl, ku):
    """(GPU factors, pivots, blocked solution)"""
    _lib = bh._lib
    p = "d" if dt == np.float64 else "s"
    dA = bh.BandedMatrix.from_host(A.astype(dt), n, kl, ku)
    F = bh.lu(dA, check=False)
    assert F.info == 0
    nb = C.c_size_t()
    _lib.call("bh_gbtrs_blocked_workspace", p.encode(), n, kl, ku, C.byref(nb))
    assert nb.value > 0
    work = bh.DeviceArray((nb.value,), np.uint8)
    X = bh.DeviceArray.from_numpy(np.asfortranarray(B.astype(dt)))
    ipiv = bh.DeviceArray.from_numpy(F.ipiv.astype(np.int64))
    _lib.call(f"bh_{p}gbtrs_blocked", n, kl, ku, B.shape[1], F.factors.data.ptr, F.factors.data.ld, ipiv.ptr, X.ptr,
              X.ld, work.ptr, nb.value)
    return F.factors.host_data(), F.ipiv, X.numpy()


@pytest.mark.parametrize("n,kl,ku,nrhs,shift", CASES)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_blocked_solve_matches_oracle(bh, oracle, dt, n, kl, ku, nrhs, shift):
    print(f"blocked gbtrs: {np.dtype(dt).name} n={n} kl={kl} ku={ku} nrhs={nrhs} shift={shift}", flush=True)
    if dt == np.float64 and (kl > 128 or kl + ku > 256):
        # outside the FP64 range (register budget of the update tiles)
        nb = C.c_size_t()
        bh._lib.call("bh_gbtrs_blocked_workspace", b"d", n, kl, ku, C.byref(nb))
        assert nb.value == 0
        return
    rng = np.random.default_rng(n + kl + 7 * ku + nrhs)
    A = band(rng, kl, ku, n, dist="n")
    if shift:
        A[ku, :] += 2.0 * (kl + ku)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    f, piv, x = _solve(bh, dt, A, B, n, kl, ku)
    xref = oracle.gbtrs("N", n, kl, ku, f.astype(np.float64), piv, B.astype(dt).astype(np.float64))
    err = np.abs(x - xref).max() / np.abs(xref).max()
    assert err <= (1e-11 if dt == np.float64 else 2e-3), err
    if dt == np.float64:
        anrm = np.abs(A).sum(axis=0).max()  # ||A||_1 >= row sums of the band's columns
        anrm = max(anrm, np.abs(A).max())
        r = B - np.stack([oracle.gbmv("N", n, n, kl, ku, 1.0, A, x[:, c]) for c in range(nrhs)], axis=1)
        eta = np.abs(r).max() / (anrm * np.abs(x).max() + np.abs(B).max())
        assert eta <= 1e-14, eta


def test_blocked_solve_edges(bh):
    _lib = bh._lib
    nb = C.c_size_t()
    # kl = 0: nothing to block (bh_?gbtrs applies)
    _lib.call("bh_gbtrs_blocked_workspace", b"d", 100, 0, 5, C.byref(nb))
    assert nb.value == 0
    st = bh.load_library().bh_dgbtrs_blocked(100, 0, 5, 1, None, 6, None, None, 100, None, 0)
    assert st == -1004
    # bands past the supported range
    _lib.call("bh_gbtrs_blocked_workspace", b"s", 1000, 385, 0, C.byref(nb))
    assert nb.value == 0
    # n = 0 and nrhs = 0 return at once
    L = bh.load_library()
    assert L.bh_dgbtrs_blocked(0, 3, 3, 1, None, 10, None, None, 1, None, 0) == 0
    assert L.bh_dgbtrs_blocked(10, 3, 3, 0, None, 10, None, None, 10, None, 0) == 0
    # a short workspace is refused before any device work
    _lib.call("bh_gbtrs_blocked_workspace", b"d", 100, 3, 3, C.byref(nb))
    assert L.bh_dgbtrs_blocked(100, 3, 3, 1, None, 10, None, None, 100, None, nb.value - 1) == -1004


def test_blocked_solve_through_ldiv(bh):
    """ldiv_(F, B, method="blocked") routes 'N' to the blocked solve (equal to
    rounding), 'T' to the exact one; "auto" takes the blocked solve for wide
    bands from n = 4096 on."""
    rng = np.random.default_rng(9)
    n, kl, ku = 5000, 64, 64
    A = band(rng, kl, ku, n, dist="n")
    A[ku, :] += 2.0 * (kl + ku)
    F = bh.lu(bh.BandedMatrix.from_host(A, n, kl, ku))
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    xe = bh.ldiv_(F, bh.DeviceArray.from_numpy(B)).numpy()
    for method in ("blocked", "auto"):
        xb = bh.ldiv_(F, bh.DeviceArray.from_numpy(B), method=method).numpy()
        assert np.abs(xb - xe).max() <= 1e-11 * np.abs(xe).max()
    te = bh.ldiv_(F.transpose(), bh.DeviceArray.from_numpy(B)).numpy()
    tb = bh.ldiv_(F.transpose(), bh.DeviceArray.from_numpy(B), method="blocked").numpy()
    assert np.array_equal(te, tb)
