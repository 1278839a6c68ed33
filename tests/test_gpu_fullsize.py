"""GPU parity at the full sizes of the bandwidth-bound and batching workloads.

Bandwidth-bound workload: n = 2^20, A with (l,u) = (8,8), B with (4,4), N(0,1)
entries; gbmv with 1 and 64 right-hand sides and A*B (giving (12,12) bands),
bit for bit against the oracle (its gbmv recipe, and its restatement of the
reference's gbmm! column decomposition).  Batching workload: 4096 systems, n
log-uniform in [256, 8192], (l,u) in {(1,1),(2,2),(4,3),(8,8)}; every
system's factors, pivots, info and solution against the oracle's dgbtf2 /
dgbtrs, bit for bit."""
import numpy as np
import pytest

from _util import band, bits_equal, in_band_mask

pytestmark = pytest.mark.gpu

N = 1 << 20


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gbmv_full_size(bh, oracle, dtype):
    rng = np.random.default_rng(0)
    A = band(rng, 8, 8, N, dtype, dist="n")
    x = rng.standard_normal(N).astype(dtype)
    BA = bh.BandedMatrix.from_host(A, N, 8, 8)
    y = (BA @ bh.DeviceArray.from_numpy(x)).numpy()
    assert bits_equal(y, oracle.gbmv("N", N, N, 8, 8, dtype(1.0), A, x))


def test_gbmv_64rhs_full_size(bh, oracle):
    rng = np.random.default_rng(1)
    A = band(rng, 8, 8, N, np.float64, dist="n")
    X = np.asfortranarray(rng.standard_normal((N, 64)))
    dY = bh.DeviceArray((N, 64), np.float64)
    bh.mul(dY, bh.BandedMatrix.from_host(A, N, 8, 8), bh.DeviceArray.from_numpy(X), 1.0, 0.0)
    Y = dY.numpy()
    for r in range(64):
        assert bits_equal(Y[:, r], oracle.gbmv("N", N, N, 8, 8, 1.0, A, X[:, r])), r


def test_gbmm_full_size(bh, oracle):
    rng = np.random.default_rng(2)
    A = band(rng, 8, 8, N, np.float64, dist="n")
    B = band(rng, 4, 4, N, np.float64, dist="n")
    C = bh.gbmm(bh.BandedMatrix.from_host(A, N, 8, 8), bh.BandedMatrix.from_host(B, N, 4, 4))
    assert C.bandwidths() == (12, 12)
    Cd = np.zeros((25, N), order="F")
    ref, st = oracle.gbmm(N, N, N, 1.0, A, 8, 8, B, 4, 4, 0.0, Cd, 12, 12)
    assert st == 0
    out = C.host_data()
    mask = in_band_mask(N, N, 12, 12)  # slots outside the matrix are not part of C
    assert bits_equal(np.where(mask, out, 0.0), np.where(mask, ref, 0.0))


@pytest.mark.parametrize("dtype", [np.float64])
def test_batched_full_workload(bh, oracle, dtype):
    n, kl, ku, bands, rhs = bh.c4_batch(4096, 0)
    bt = bh.BatchedGBSV(n, kl, ku, dtype)
    AB, B = bt.pack(bands, rhs)
    bt.upload(AB, B)
    bt.solve()
    res = bt.results()
    assert len(res) == 4096
    for s, (x, f, p, info) in enumerate(res):
        nn, l, u = int(n[s]), int(kl[s]), int(ku[s])
        ld = 2 * l + u + 1
        ab = AB[bt.ab_off[s]:bt.ab_off[s] + ld * nn].reshape(nn, ld).T.copy(order="F")
        fo, po, io = oracle.gbtf2(nn, nn, l, u, ab)
        assert info == io == 0, s
        assert np.array_equal(p, po), s
        assert bits_equal(f, fo), s
        xo = oracle.gbtrs("N", nn, l, u, fo, po, B[bt.vec_off[s]:bt.vec_off[s] + nn])
        assert bits_equal(x, xo), s
    bt.close()
