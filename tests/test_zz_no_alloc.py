"""GPU: bh_dsgbsv_noalloc, the mixed solve whose workspace is the library's
statically allocated device scratch, against bh_dsgbsv with a caller
workspace on the same inputs.  Both run the same kernels in the same order,
so X, the refinement count, info and the FP32 pivots must be identical bit
for bit.  Two calls in a row, the second smaller (the scratch keeps the first
call's contents), one call that takes the FP64 fallback (which overlays the
refinement's workspace), and the size limit."""
import ctypes as C

import numpy as np
import pytest

from _util import band, bits_equal

pytestmark = pytest.mark.gpu

BH_ERR_WORKSPACE = -1004


def _problem(seed, n, kl, ku, nrhs, shift, scale=1.0):
    rng = np.random.default_rng(seed)
    A = band(rng, kl, ku, n, dist="n")
    A[ku, :] += shift
    A *= scale
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    return A, B


def _both(bh, A_h, B_h, n, kl, ku, nrhs):
    """(X, iters, info, ipiv32) from bh_dsgbsv and from bh_dsgbsv_noalloc."""
    _lib = bh._lib
    A = bh.BandedMatrix.from_host(A_h, n, kl, ku)
    B = bh.DeviceArray.from_numpy(B_h)
    out = []
    for name in ("bh_dsgbsv", "bh_dsgbsv_noalloc"):
        X = bh.DeviceArray((n, nrhs), np.float64)
        piv = bh.DeviceArray((max(1, n),), np.int64)
        iters, info = C.c_int64(), C.c_int64()
        if name == "bh_dsgbsv":
            nb = C.c_size_t()
            _lib.call("bh_dsgbsv_workspace", n, kl, ku, nrhs, C.byref(nb))
            work = bh.DeviceArray((max(1, nb.value),), np.uint8)
            _lib.call(name, n, kl, ku, nrhs, A.data.ptr, A.data.ld, B.ptr, n, X.ptr, n, work.ptr, nb.value, piv.ptr,
                      C.byref(iters), C.byref(info))
        else:
            _lib.call(name, n, kl, ku, nrhs, A.data.ptr, A.data.ld, B.ptr, n, X.ptr, n, piv.ptr, C.byref(iters),
                      C.byref(info))
        out.append((X.numpy(), int(iters.value), int(info.value), piv.numpy()[:n]))
    return out


def _check_same(ref, got):
    assert got[1] == ref[1] and got[2] == ref[2], (ref[1:3], got[1:3])
    assert bits_equal(got[3], ref[3])
    assert bits_equal(got[0], ref[0])


def test_noalloc_matches_workspace_entry_two_calls(bh):
    # first call: MFMA band LU (kl >= 32, ku >= 65), 16 right-hand sides
    n, kl, ku, nrhs = 4096, 80, 80, 16
    print(f"noalloc call 1: n={n} kl={kl} ku={ku} nrhs={nrhs}", flush=True)
    A, B = _problem(1, n, kl, ku, nrhs, 2.0 * (kl + ku))
    ref, got = _both(bh, A, B, n, kl, ku, nrhs)
    assert ref[1] >= 0
    _check_same(ref, got)
    # second call, smaller, on the scratch the first call left behind
    n, kl, ku, nrhs = 1000, 8, 8, 2
    print(f"noalloc call 2: n={n} kl={kl} ku={ku} nrhs={nrhs}", flush=True)
    A, B = _problem(2, n, kl, ku, nrhs, 2.0 * (kl + ku))
    ref, got = _both(bh, A, B, n, kl, ku, nrhs)
    assert ref[1] >= 0
    _check_same(ref, got)


def test_noalloc_fallback_matches(bh):
    # every entry below the FP32 range: the FP32 copy is all zeros, its LU
    # reports a zero pivot and the FP64 fallback runs (iters < 0), its factors
    # placed over the refinement's workspace in the static scratch
    n, kl, ku, nrhs = 3000, 2, 3, 3
    print(f"noalloc fallback: n={n} kl={kl} ku={ku} nrhs={nrhs}", flush=True)
    A, B = _problem(3, n, kl, ku, nrhs, 2.0 * (kl + ku), scale=1e-50)
    ref, got = _both(bh, A, B, n, kl, ku, nrhs)
    assert ref[1] < 0 and ref[2] == 0
    _check_same(ref, got)


def test_noalloc_rejects_sizes_beyond_the_scratch(bh):
    # C3's shape needs ~202 MB for the FP64 fallback: over the 64 MiB limit;
    # the check is on the host, before any device work
    iters, info = C.c_int64(), C.c_int64()
    L = bh.load_library()
    st = L.bh_dsgbsv_noalloc(65536, 128, 128, 16, None, 385, None, 65536, None, 65536, None, C.byref(iters),
                             C.byref(info))
    assert st == BH_ERR_WORKSPACE
    assert b"static scratch" in L.bh_last_error()
