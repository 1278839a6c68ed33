"""GPU parity: the MFMA-blocked band LU -- for ku >= 65, kl >= 32 (the
regime in which the reference's LAPACK.gbtrf! runs blocked xGBTRF) and for
long systems (min(m, n) >= 16384) with kl, ku >= 8.  A planted first column
checks the pivot search's tie rules: magnitudes equal in their high 32 bits
(the FP64 search's second wave maximum decides) and exactly equal ones (the smallest
position wins).

Pivots (ipiv) and info must equal the oracle's (xGBTF2 restated, which picks
the same pivots as xGBTRF); factors must agree to rounding (FP64: 1e-12 of
the factor scale; FP32: 1e-4), and the factors must solve A x = b with a
small backward error."""
import ctypes as C

import numpy as np
import pytest

from _util import band

pytestmark = pytest.mark.gpu

CASES = [
    # m, n, kl, ku, dist, diagonal shift
    (100, 100, 32, 65, "n", 0.0),
    (200, 200, 32, 65, "u", 0.0),
    (333, 333, 45, 70, "n", 0.0),
    (1000, 1000, 64, 100, "n", 0.0),
    (517, 517, 100, 90, "n", 0.0),
    (1000, 1000, 128, 128, "n", 0.0),
    (1000, 1000, 128, 128, "n", 512.0),
    (4096, 4096, 128, 128, "n", 0.0),
    (300, 400, 40, 70, "n", 0.0),
    (400, 300, 40, 70, "n", 0.0),
    (40, 40, 32, 65, "n", 0.0),
    (16384, 16384, 8, 8, "n", 0.0),      # long narrow systems
    (20000, 20000, 16, 16, "n", 0.0),
    (20000, 20000, 16, 16, "n", 64.0),
    (17000, 16500, 12, 9, "u", 0.0),
]


def _factor(bh, oracle, dtype, m, n, kl, ku, dist, shift):
    rng = np.random.default_rng(m + 3 * n + 31 * kl + ku)
    A = band(rng, kl, ku, n, np.float64, dist=dist, m=m)
    A[ku, :] += shift
    A = A.astype(dtype)
    ab = oracle.pad_for_lu(A, m, n, kl, ku)
    fref, pref, iref = oracle.gbtf2(m, n, kl, ku, ab.copy())
    dab = bh.DeviceArray.from_numpy(ab)
    ipiv = bh.DeviceArray((max(min(m, n), 1),), np.int64)
    info = C.c_int64(-7)
    p = "d" if dtype == np.float64 else "s"
    bh._lib.call(f"bh_{p}gbtrf", m, n, kl, ku, dab.ptr, dab.ld, ipiv.ptr, C.byref(info))
    return A, ab, fref, pref, iref, dab.numpy(), ipiv.numpy()[: min(m, n)], info.value


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,kl,ku,dist,shift", CASES)
def test_gbtrf_mfma(bh, oracle, dtype, m, n, kl, ku, dist, shift):
    print(f"case {dtype.__name__} m={m} n={n} kl={kl} ku={ku} shift={shift}", flush=True)
    A, ab, fref, pref, iref, f, piv, info = _factor(bh, oracle, dtype, m, n, kl, ku, dist, shift)
    assert info == iref
    same = int(np.sum(piv == pref))
    if dtype == np.float64:
        assert same == len(pref), f"pivots {same}/{len(pref)}"
    else:
        assert same >= len(pref) - 2, f"pivots {same}/{len(pref)}"
    if dtype == np.float64 or same == len(pref):
        scale = np.abs(fref).max()
        tol = 1e-12 if dtype == np.float64 else 1e-4
        err = np.abs(f.astype(np.float64) - fref.astype(np.float64)).max() / scale
        assert err <= tol, err


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,kl,ku", [(1000, 128, 128), (777, 45, 70)])
def test_gbtrf_mfma_solve(bh, oracle, dtype, n, kl, ku):
    A, ab, fref, pref, iref, f, piv, info = _factor(bh, oracle, dtype, n, n, kl, ku, "n", 0.0)
    assert info == 0
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n).astype(dtype)
    x = oracle.gbtrs("N", n, kl, ku, f, piv, b)
    Ad = oracle.band_to_dense(A.astype(np.float64), n, n, kl, ku)
    r = b.astype(np.float64) - Ad @ x.astype(np.float64)
    berr = np.abs(r).max() / (np.abs(Ad).sum(1).max() * np.abs(x).max() + np.abs(b).max())
    assert berr <= (1e-14 if dtype == np.float64 else 1e-5), berr


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("bump,expect", [(2.0 ** -40, 5), (0.0, 2)])
def test_gbtrf_mfma_pivot_ties(bh, oracle, dtype, bump, expect):
    """Column 0 holds 7 at rows 2 and 5, the second one larger by a relative
    2^-40 (equal high words in FP64) or equal; every other entry is below 1."""
    n, kl, ku = 16384, 8, 8
    rng = np.random.default_rng(11)
    A = band(rng, kl, ku, n, np.float64, dist="u") - 0.5
    A[ku + 2, 0] = 7.0
    A[ku + 5, 0] = 7.0 * (1.0 + bump)
    A = A.astype(dtype)
    ab = oracle.pad_for_lu(A, n, n, kl, ku)
    fref, pref, iref = oracle.gbtf2(n, n, kl, ku, ab.copy())
    dab = bh.DeviceArray.from_numpy(ab)
    ipiv = bh.DeviceArray((n,), np.int64)
    info = C.c_int64(-7)
    p = "d" if dtype == np.float64 else "s"
    bh._lib.call(f"bh_{p}gbtrf", n, n, kl, ku, dab.ptr, dab.ld, ipiv.ptr, C.byref(info))
    piv = ipiv.numpy()
    exp_row = expect if (dtype == np.float64 or bump == 0.0) else 2  # in FP32 the bump rounds away
    assert pref[0] == exp_row + 1
    assert piv[0] == pref[0]
    assert info.value == iref
    same = int(np.sum(piv == pref))
    assert same >= (n if dtype == np.float64 else n - 2), f"pivots {same}/{n}"
