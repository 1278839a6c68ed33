"""GPU: FP32 band LU + FP64 iterative refinement (bh_dsgbsv).

The FP32 factorisation itself is pinned bit for bit against the oracle's FP32
gbtf2; the refined FP64 solution must meet LAPACK dsgesv's backward-error
test and agree with the FP64 oracle solve to 1e-12 (forward error, max norm)
on diagonally dominant systems; the pivoting variant must either converge or
report the FP64 fallback."""
import numpy as np
import pytest

from _util import band, bits_equal

pytestmark = pytest.mark.gpu


def _problem(seed, n, kl, ku, nrhs, shift):
    rng = np.random.default_rng(seed)
    A = band(rng, kl, ku, n, dist="n")
    A[ku, :] += shift
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    return A, B


@pytest.mark.parametrize("n,kl,ku,nrhs", [(4096, 128, 128, 16), (2000, 16, 16, 3), (10000, 4, 3, 1), (500, 2, 7, 5)])
def test_mixed_shifted_meets_1e12(bh, oracle, n, kl, ku, nrhs):
    A, B = _problem(n + kl, n, kl, ku, nrhs, 2.0 * (kl + ku))
    dA = bh.BandedMatrix.from_host(A, n, kl, ku)
    res = bh.solve_mixed(dA, bh.DeviceArray.from_numpy(B))
    assert res.iters >= 0, "refinement should converge on a diagonally dominant system"
    x = res.x.numpy()
    f, piv, info = oracle.gbtf2(n, n, kl, ku, oracle.pad_for_lu(A, n, n, kl, ku))
    xref = oracle.gbtrs("N", n, kl, ku, f, piv, B)
    ferr = np.abs(x - xref).max() / np.abs(xref).max()
    assert ferr <= 1e-12, ferr
    # dsgesv test on every column
    anrm = max(np.abs(A[:, :]).sum(axis=0).max(), 1.0)
    for c in range(nrhs):
        r = B[:, c] - oracle.gbmv("N", n, n, kl, ku, 1.0, A, x[:, c])
        assert np.abs(r).max() <= np.abs(x[:, c]).max() * 2 * anrm * 1.1102230246251565e-16 * np.sqrt(n) * 4
    # FP32 pivots equal the FP64 oracle's on a diagonally dominant matrix
    assert np.array_equal(res.ipiv32, piv)


def test_mixed_fp32_factor_bits(bh, oracle):
    """The FP32 factorisation inside the mixed solve is the oracle's sgbtf2."""
    n, kl, ku = 1000, 8, 8
    A, _ = _problem(3, n, kl, ku, 1, 0.0)
    A32 = A.astype(np.float32)
    f, piv, info = oracle.gbtf2(n, n, kl, ku, oracle.pad_for_lu(A32, n, n, kl, ku))
    F = bh.lu(bh.BandedMatrix.from_host(A32, n, kl, ku), check=False)
    assert np.array_equal(F.ipiv, piv) and bits_equal(F.factors.host_data(), f)


@pytest.mark.parametrize("n,kl,ku", [(2048, 64, 64), (3000, 16, 16), (4096, 128, 128)])
def test_mixed_unshifted_converges_or_falls_back(bh, oracle, n, kl, ku):
    A, B = _problem(7 + n, n, kl, ku, 4, 0.0)
    dA = bh.BandedMatrix.from_host(A, n, kl, ku)
    res = bh.solve_mixed(dA, bh.DeviceArray.from_numpy(B))
    x = res.x.numpy()
    # both branches end with an FP64-accurate answer: small backward error
    for c in range(B.shape[1]):
        r = B[:, c] - oracle.gbmv("N", n, n, kl, ku, 1.0, A, x[:, c])
        normA = np.abs(A).sum(axis=0).max()
        berr = np.abs(r).max() / (normA * np.abs(x[:, c]).max() + np.abs(B[:, c]).max())
        assert berr <= 1e-12, (res.iters, berr)


def test_mixed_workspace_too_small_is_an_error(bh):
    import ctypes as C

    from banded_hip import _lib

    n, kl, ku = 64, 2, 2
    A, B = _problem(1, n, kl, ku, 1, 10.0)
    dA = bh.BandedMatrix.from_host(A, n, kl, ku)
    dB, dX = bh.DeviceArray.from_numpy(B), bh.DeviceArray((n, 1))
    it, info = C.c_int64(), C.c_int64()
    L = _lib.load()
    st = L.bh_dsgbsv(n, kl, ku, 1, dA.data.ptr, dA.data.ld, dB.ptr, n, dX.ptr, n, None, 0, None, C.byref(it),
                     C.byref(info))
    assert st == -1004


def test_mixed_unshifted_128_forward_error(bh, oracle):
    """The precision workload's pivoting variant at reduced n: l = u = 128, no
    diagonal shift, 16 right-hand sides.  The refined FP64 solution agrees with
    the FP64 oracle's solve to 1e-12 (forward error relative to the solution's
    max norm), whether refinement converged (iters >= 0) or fell back to the
    FP64 factorisation (iters < 0); row interchanges are active."""
    n, kl, ku, nrhs = 4096, 128, 128, 16
    A, B = _problem(99, n, kl, ku, nrhs, 0.0)
    dA = bh.BandedMatrix.from_host(A, n, kl, ku)
    res = bh.solve_mixed(dA, bh.DeviceArray.from_numpy(B))
    f, piv, info = oracle.gbtf2(n, n, kl, ku, oracle.pad_for_lu(A, n, n, kl, ku))
    assert info == 0
    assert int(np.sum(piv != np.arange(1, n + 1))) > n // 2
    xref = oracle.gbtrs("N", n, kl, ku, f, piv, B)
    x = res.x.numpy()
    ferr = np.abs(x - xref).max() / np.abs(xref).max()
    assert ferr <= 1e-12, (res.iters, ferr)
