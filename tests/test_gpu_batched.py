"""GPU: batched band solves (one persistent launch) against the oracle.

Every system's factors, pivots and info must equal the oracle's dgbtf2
(sgbtf2) bit for bit and its solution the oracle's dgbtrs 'N'; graph replays
and autotuned schedules must reproduce the eager result."""
import ctypes as C

import numpy as np
import pytest

from _util import bits_equal

pytestmark = pytest.mark.gpu


def _batch(bh, dtype, nsys=96, seed=3, nmin=16, nmax=700,
           classes=((1, 1), (2, 2), (4, 3), (8, 8), (0, 5), (3, 9), (8, 40), (1, 0))):
    n, kl, ku, bands, rhs = bh.c4_batch(nsys, seed, nmin, nmax, classes)
    bands = [b.astype(dtype) for b in bands]
    rhs = [r.astype(dtype) for r in rhs]
    return n, kl, ku, bands, rhs


def _check(oracle, bt, n, kl, ku, AB, B):
    res = bt.results()
    for s, (x, f, p, info) in enumerate(res):
        nn, l, u = int(n[s]), int(kl[s]), int(ku[s])
        ld = 2 * l + u + 1
        ab = AB[bt.ab_off[s]:bt.ab_off[s] + ld * nn].reshape(nn, ld).T.copy(order="F")
        fo, po, io = oracle.gbtf2(nn, nn, l, u, ab)
        assert bits_equal(f, fo), (s, nn, l, u)
        assert np.array_equal(p, po) and info == io, (s, nn, l, u)
        xo = oracle.gbtrs("N", nn, l, u, fo, po, B[bt.vec_off[s]:bt.vec_off[s] + nn])
        assert bits_equal(x, xo), (s, nn, l, u)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_matches_oracle(bh, oracle, dtype):
    n, kl, ku, bands, rhs = _batch(bh, dtype)
    bt = bh.BatchedGBSV(n, kl, ku, dtype)
    AB, B = bt.pack(bands, rhs)
    bt.upload(AB, B)
    bt.solve()
    _check(oracle, bt, n, kl, ku, AB, B)
    bt.close()


def test_batched_graph_replays(bh, oracle):
    n, kl, ku, bands, rhs = _batch(bh, np.float64, nsys=64, seed=9)
    bt = bh.BatchedGBSV(n, kl, ku)
    AB, B = bt.pack(bands, rhs)
    g = bt.graph()
    for rep in range(3):
        bt.upload(AB, B)  # the solve is in place: fresh inputs before every replay
        g.launch()
        bh.sync()
        _check(oracle, bt, n, kl, ku, AB, B)
    g.close()
    bt.close()


def test_batched_autotune_keeps_results(bh, oracle):
    n, kl, ku, bands, rhs = _batch(bh, np.float64, nsys=128, seed=5)
    bt = bh.BatchedGBSV(n, kl, ku)
    AB, B = bt.pack(bands, rhs)
    bt.upload(AB, B)
    ms = bt.autotune()
    assert ms > 0.0
    order, wpb, blocks = bt.schedule()
    assert order in (0, 2) and wpb in (1, 2, 4) and blocks >= 1
    # the autotuner left the inputs untouched
    assert bits_equal(bt.AB.numpy(), AB) and bits_equal(bt.B.numpy(), B)
    bt.solve()
    _check(oracle, bt, n, kl, ku, AB, B)
    for o in (0, 1, 2):
        bt.schedule(order=o, waves_per_block=2, blocks=64)
        bt.upload(AB, B)
        bt.solve()
        _check(oracle, bt, n, kl, ku, AB, B)
    bt.close()


def test_batched_singular_system_reports_its_info(bh, oracle):
    n, kl, ku, bands, rhs = _batch(bh, np.float64, nsys=8, seed=2, classes=((2, 2),))
    bands[3][:, 5] = 0.0  # column 5 of system 3 is zero
    bands[3][2, 5] = 0.0
    bt = bh.BatchedGBSV(n, kl, ku)
    AB, B = bt.pack(bands, rhs)
    bt.upload(AB, B)
    bt.solve()
    infos = [r[3] for r in bt.results()]
    assert infos[3] > 0 and all(i == 0 for k, i in enumerate(infos) if k != 3)
    bt.close()


def test_batched_plan_rejects_wide_kl(bh):
    from banded_hip import BandedHIPError

    with pytest.raises(BandedHIPError):
        bh.BatchedGBSV([100, 100], [2, 9], [2, 2])


def test_batched_schedule_is_fixed_while_a_graph_exists(bh, oracle):
    """A captured graph bakes in the schedule: changing it (or autotuning)
    while the graph exists is an error (BH_ERR_BUSY), and works again once
    the graph is destroyed; the graph keeps its plan and buffers alive."""
    from banded_hip import BandedHIPError

    n, kl, ku, bands, rhs = _batch(bh, np.float64, nsys=16, seed=4)
    bt = bh.BatchedGBSV(n, kl, ku)
    AB, B = bt.pack(bands, rhs)
    bt.upload(AB, B)
    g = bt.graph()
    with pytest.raises(BandedHIPError) as ei:
        bt.schedule(order=1, waves_per_block=1, blocks=8)
    assert ei.value.status == -1006
    with pytest.raises(BandedHIPError):
        bt.autotune()
    g.launch()
    bh.sync()
    _check(oracle, bt, n, kl, ku, AB, B)
    g.close()
    bt.schedule(order=1, waves_per_block=1, blocks=8)
    bt.upload(AB, B)
    bt.solve()
    _check(oracle, bt, n, kl, ku, AB, B)
    bt.close()
