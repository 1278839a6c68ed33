"""GPU: the partitioned band solve (bh_?gbtrs_part, 'N' and 'T') with the
factors of bh_?gbtrf, against the oracle's xGBTRS restatement on the same
factors.

The partitioned solve splits the rows into blocks solved at once and joins
them through the linear map each block applies to the state it takes over, so
it equals xGBTRS to rounding, not bit for bit.  Bars: the normwise backward
error ||b - op(A) x|| / (||A|| ||x|| + ||b||) within 1e-13 (FP64) / 2e-5
(FP32, against the FP32 matrix), and, for diagonally dominant systems (where
the condition number is small), the solution within 1e-11 (FP64) / 1e-3
(FP32) of the oracle's solve (max norm, relative).  Cases: the adjoint and
plain workloads' shapes (n = 262144, (16,16); n = 10^4, (4,3)), pivoting (no
diagonal shift; the triangular cases kl = 0 and ku = 0 are shifted, a random
triangular matrix's condition growing exponentially with n, as does that of a
random band with very unequal bandwidths, so the unshifted cases keep kl and
ku close), the widest bands
the solve takes (kl = 31, kl + ku = 63), more than 16 right-hand sides, a single block, n below the window, and zero
bandwidths."""
import ctypes as C

import numpy as np
import pytest

from _util import band

pytestmark = pytest.mark.gpu

CASES = [
    (262144, 16, 16, 1, True),
    (65536, 16, 16, 2, False),
    (10000, 4, 3, 1, False),
    (5000, 31, 32, 3, False),
    (3001, 8, 8, 20, False),
    (4000, 12, 5, 1, True),
    (100, 3, 3, 2, False),
    (40, 20, 10, 1, False),
    (2000, 0, 5, 1, True),
    (2000, 5, 0, 1, True),
    (1, 0, 0, 1, False),
]


def _factor(bh, dt, A, n, kl, ku):
    dA = bh.BandedMatrix.from_host(A.astype(dt), n, kl, ku)
    F = bh.lu(dA, check=False)
    assert F.info == 0
    return F


def _part(bh, dt, F, B, trans):
    _lib = bh._lib
    p = "d" if dt == np.float64 else "s"
    n, nrhs = B.shape
    nb = C.c_size_t()
    _lib.call("bh_gbtrs_part_workspace", p.encode(), n, F.kl, F.ku, nrhs, C.byref(nb))
    assert nb.value > 0
    work = bh.DeviceArray((nb.value,), np.uint8)
    X = bh.DeviceArray.from_numpy(np.asfortranarray(B.astype(dt)))
    _lib.call(f"bh_{p}gbtrs_part", trans.encode(), n, F.kl, F.ku, nrhs, F.factors.data.ptr, F.factors.data.ld,
              F.ipiv_dev.ptr, X.ptr, X.ld, work.ptr, nb.value)
    return X.numpy()


@pytest.mark.parametrize("n,kl,ku,nrhs,shift", CASES)
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_partitioned_solve_matches_oracle(bh, oracle, dt, trans, n, kl, ku, nrhs, shift):
    print(f"partitioned gbtrs: {np.dtype(dt).name} {trans} n={n} kl={kl} ku={ku} nrhs={nrhs} shift={shift}",
          flush=True)
    rng = np.random.default_rng(n + 3 * kl + 7 * ku + nrhs + (trans == "T"))
    A = band(rng, kl, ku, n, dist="n")
    if shift or n < 64:
        A[ku, :] += 2.0 * (kl + ku) + 1.0
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    F = _factor(bh, dt, A, n, kl, ku)
    x = _part(bh, dt, F, B, trans)
    Ad = A.astype(dt).astype(np.float64)
    Bd = B.astype(dt).astype(np.float64)
    anrm = max(np.abs(Ad).sum(axis=0).max(), np.abs(Ad).max())
    r = Bd - np.stack([oracle.gbmv(trans, n, n, kl, ku, 1.0, Ad, x[:, c].astype(np.float64)) for c in range(nrhs)],
                      axis=1)
    eta = np.abs(r).max() / (anrm * np.abs(x).max() + np.abs(Bd).max())
    assert eta <= (1e-13 if dt == np.float64 else 2e-5), eta
    if shift or n < 64:
        f = F.factors.host_data().astype(np.float64)
        xref = oracle.gbtrs(trans, n, kl, ku, f, F.ipiv, Bd)
        err = np.abs(x - xref).max() / np.abs(xref).max()
        assert err <= (1e-11 if dt == np.float64 else 1e-3), err


def test_partitioned_solve_through_ldiv(bh):
    """ldiv_(F, B, method="partitioned") routes to the partitioned solve and
    agrees with the exact solve to rounding; bands outside its range fall
    back to the exact solve."""
    rng = np.random.default_rng(5)
    for (n, kl, ku) in ((20000, 6, 9), (3000, 40, 10)):
        A = band(rng, kl, ku, n, dist="n")
        A[ku, :] += 2.0 * (kl + ku)
        F = bh.lu(bh.BandedMatrix.from_host(A, n, kl, ku))
        b = rng.standard_normal(n)
        for fact in (F, F.transpose()):
            xe = bh.ldiv_(fact, bh.DeviceArray.from_numpy(b)).numpy()
            xp = bh.ldiv_(fact, bh.DeviceArray.from_numpy(b), method="partitioned").numpy()
            assert np.abs(xp - xe).max() <= 1e-12 * np.abs(xe).max()
    with pytest.raises(ValueError):
        bh.ldiv_(F, bh.DeviceArray.from_numpy(b), method="nope")


def test_partitioned_solve_edges(bh):
    _lib = bh._lib
    L = bh.load_library()
    nb = C.c_size_t()
    # bands past the range: kl <= 31 and kl + ku <= 63
    _lib.call("bh_gbtrs_part_workspace", b"d", 1000, 32, 0, 1, C.byref(nb))
    assert nb.value == 0
    _lib.call("bh_gbtrs_part_workspace", b"s", 1000, 20, 44, 1, C.byref(nb))
    assert nb.value == 0
    assert L.bh_dgbtrs_part(b"N", 1000, 32, 0, 1, None, 65, None, None, 1000, None, 0) == -1004
    # n = 0 and nrhs = 0 return at once
    assert L.bh_dgbtrs_part(b"N", 0, 3, 3, 1, None, 10, None, None, 1, None, 0) == 0
    assert L.bh_dgbtrs_part(b"T", 10, 3, 3, 0, None, 10, None, None, 10, None, 0) == 0
    # a short workspace is refused before any device work
    _lib.call("bh_gbtrs_part_workspace", b"d", 100, 3, 3, 1, C.byref(nb))
    assert nb.value > 0
    assert L.bh_dgbtrs_part(b"N", 100, 3, 3, 1, None, 10, None, None, 100, None, nb.value - 1) == -1004
    # bad arguments
    assert L.bh_dgbtrs_part(b"X", 100, 3, 3, 1, None, 10, None, None, 100, None, 0) == -1
    assert L.bh_dgbtrs_part(b"N", 100, 3, 3, 1, None, 9, None, None, 100, None, 0) == -7
