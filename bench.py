"""Benchmark of the banded hot path on one MI355X.

Headline (value): mixed-precision A\\b solves/s on the precision workload
  n = 65536, l = u = 128, band entries iid N(0,1) with the diagonal shifted by
  +2(l+u), 16 right-hand sides iid N(0,1); one step = one complete A\\b (FP32
  band LU + FP64 iterative refinement to LAPACK dsgesv's backward-error test),
  inputs resident in HBM.  The plain FP64 GPU A\\b (band LU + solve) is timed
  beside it; both solutions are compared with the FP64 CPU solution (forward
  error, max norm) and the FP64 pivots with the CPU's.
Second half of the metric: FP64 gbmv (8,8) at n = 2^20 in HBM GB/s
  (algorithmic bytes 152 n = A + x + y), buffers rotated so that every launch
  reads HBM rather than the 256 MB last-level cache.
Beside them (extra fields): gbmv FP32 / 64 right-hand sides and A*B (8,8)x(4,4)
  at n = 2^20; the batch of 4096 systems (one persistent launch, autotuned);
  the gradient of dot(w, A\\b) + 0.5||A x0||^2 at n = 262144, l = u = 16; the
  plain n = 10^4, (4,3) case; qr and qr \\ b at the reference's benchmark shapes.
cpu_baseline: the reference's CPU path for the headline solve -- OpenBLAS
  dgbtrf + dgbtrs (the LAPACK routines BandedLU.jl:98 and linalg.jl:28 call),
  median of 10 after 2 warm-ups at 1 thread and at every host thread (the
  faster is the value; CPU model and core count in the line); per-workload CPU
  timings sit in each extra field.

Prints one JSON line.  Only one GPU is supported: --gpus N > 1 exits.
`--traffic-probe NAME` runs only a few launches of one kernel and a 256 MiB
device copy (the scale for rocprofv3 FETCH_SIZE / WRITE_SIZE counters).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "bandedmatrices.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBPS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _host_threads():
    """Host cores this process may run on (the affinity mask, not the machine's count)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_precision_problem(n=65536, l=128, u=128, nrhs=16, shift=True, seed=0):
    rng = np.random.default_rng(seed)
    A = np.asfortranarray(rng.standard_normal((l + u + 1, n)))
    if shift:
        A[u, :] += 2 * (l + u)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    return A, B


def lu_flops(n, l, u):
    """Multiply-adds x2 of a band LU without row exchanges: scale l + update 2 l u per column."""
    return float(n) * (l + 2.0 * l * u)


def solve_flops(n, l, u, nrhs):
    return float(nrhs) * 2.0 * n * (2 * l + u)


def cpu_reference_solve(A, B, l, u, threads):
    """FP64 band LU + solve with OpenBLAS (scipy), the reference's LAPACK path."""
    import scipy.linalg.lapack as la
    from threadpoolctl import threadpool_limits

    n = A.shape[1]
    ab = np.zeros((2 * l + u + 1, n), order="F")
    ab[l:, :] = A
    with threadpool_limits(threads):
        t0 = time.perf_counter()
        lub, piv, info = la.dgbtrf(ab, l, u)
        x, info2 = la.dgbtrs(lub, l, u, B, piv)
        t1 = time.perf_counter()
    return x, piv + 1, t1 - t0


def backward_error(A, l, u, X, B):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), FP64 on the host."""
    import scipy.linalg.blas as sb

    n = A.shape[1]
    anrm = 0.0
    rows = np.zeros(n)
    for t in range(l + u + 1):  # row sums of |A| from the band rows
        j = np.arange(n)
        i = j - u + t
        ok = (i >= 0) & (i < n)
        np.add.at(rows, i[ok], np.abs(A[t, ok]))
    anrm = rows.max()
    worst = 0.0
    for c in range(X.shape[1]):
        r = B[:, c] - sb.dgbmv(n, n, l, u, 1.0, A, X[:, c])
        worst = max(worst, np.abs(r).max() / (anrm * np.abs(X[:, c]).max() + np.abs(B[:, c]).max()))
    return float(worst)


def gpu_time(bh, fn, reps):
    t = bh.Timer()
    t.start()
    for _ in range(reps):
        fn()
    t.stop()
    return t.elapsed_ms() / reps


def warm(bh, fn, ms=100.0):
    """Run fn back to back for about ms of GPU time, so the timed loop after
    it starts at steady clocks (the host-side set-up of a workload's operands
    leaves the GPU idle long enough to clock down)."""
    one = gpu_time(bh, fn, 1)
    for _ in range(int(min(2000, max(1, ms / max(one, 1e-3))))):
        fn()
    bh.sync()


# ----------------------------------------------------------------- precision (headline)
def bench_precision(bh, steps, warmup, shift=True):
    import ctypes as C

    from banded_hip import _lib

    n, l, u, nrhs = 65536, 128, 128, 16
    A_h, B_h = make_precision_problem(n, l, u, nrhs, shift=shift)
    A = bh.BandedMatrix.from_host(A_h, n, l, u)
    B = bh.DeviceArray.from_numpy(B_h)
    X = bh.DeviceArray((n, nrhs), np.float64)
    solver = bh.MixedSolver(n, l, u, nrhs)
    for _ in range(warmup):
        res = solver.solve(A, B, X)
    bh.sync()
    iters = []
    t = bh.Timer()
    t.start()
    for _ in range(steps):
        res = solver.solve(A, B, X)
        iters.append(res.iters)
    t.stop()
    ms_mixed = t.elapsed_ms() / steps
    x_mixed = X.numpy()

    def fp64_solve():
        # A\b in FP64: the MFMA LU and, for these wide bands, the MFMA-blocked solve
        F = bh.lu(A)
        Xd = B.copy()
        bh.ldiv_(F, Xd, method="auto")
        return F, Xd

    for _ in range(warmup):
        F64, Xd = fp64_solve()
    bh.sync()
    t.start()
    for _ in range(steps):
        F64, Xd = fp64_solve()
    t.stop()
    ms_fp64 = t.elapsed_ms() / steps
    x_fp64 = Xd.numpy()

    # the pieces of the mixed step, each timed alone on the library stream
    ab32 = bh.pad_for_lu(A, np.float32)
    keep = ab32.data.copy()
    ipiv = bh.DeviceArray((n,), np.int64)
    info = C.c_int64()
    lu_ms = []
    for _ in range(max(3, steps)):
        _lib.call("bh_memcpy_dtod", ab32.data.ptr, keep.ptr, keep.nbytes)
        t.start()
        _lib.call("bh_sgbtrf", n, n, l, u, ab32.data.ptr, ab32.data.ld, ipiv.ptr, C.byref(info))
        t.stop()
        lu_ms.append(t.elapsed_ms())
    B32 = bh.DeviceArray.from_numpy(np.asfortranarray(B_h.astype(np.float32)))
    # the FP32 solve the refinement uses: the blocked MFMA solve (its operand
    # pre-pass included, which the mixed solver runs once per factorisation)
    nb = C.c_size_t()
    _lib.call("bh_gbtrs_blocked_workspace", b"s", n, l, u, C.byref(nb))
    work = bh.DeviceArray((max(1, nb.value),), np.uint8)
    trs_ms = []
    for _ in range(max(3, steps)):
        t.start()
        _lib.call("bh_sgbtrs_blocked", n, l, u, nrhs, ab32.data.ptr, ab32.data.ld, ipiv.ptr, B32.ptr, n, work.ptr,
                  nb.value)
        t.stop()
        trs_ms.append(t.elapsed_ms())
    ring_ms = []
    for _ in range(3):
        t.start()
        _lib.call("bh_sgbtrs", b"N", n, l, u, nrhs, ab32.data.ptr, ab32.data.ld, ipiv.ptr, B32.ptr, n)
        t.stop()
        ring_ms.append(t.elapsed_ms())
    return dict(n=n, l=l, u=u, nrhs=nrhs, A_h=A_h, B_h=B_h, ms_mixed=ms_mixed, ms_fp64=ms_fp64, iters=iters,
                x_mixed=x_mixed, x_fp64=x_fp64, ipiv64=F64.ipiv, lu32_ms=float(np.median(lu_ms[1:])),
                trs32_ms=float(np.median(trs_ms[1:])), trs32_ring_ms=float(np.median(ring_ms[1:])))


# ----------------------------------------------------------------- bandwidth-bound kernels
def bench_bandwidth(bh, steps, skip_cpu=True):
    n = 1 << 20
    rng = np.random.default_rng(1)
    out = {}
    for dt, tag in ((np.float64, "fp64"), (np.float32, "fp32")):
        kl = ku = 8
        sets = []
        # enough matrices that the rotation spans > 512 MB: no launch finds
        # its operands in the 256 MB last-level cache
        for _ in range(4 if dt == np.float64 else 7):
            A = bh.BandedMatrix.from_host(np.asfortranarray(rng.standard_normal((17, n)).astype(dt)), n, kl, ku)
            x = bh.DeviceArray.from_numpy(rng.standard_normal(n).astype(dt))
            y = bh.DeviceArray((n,), dt)
            sets.append((A, x, y))
        k = [0]

        def one():
            A, x, y = sets[k[0] % len(sets)]
            k[0] += 1
            bh.mul(y, A, x)

        warm(bh, one)
        ms = gpu_time(bh, one, max(30, steps * 6))
        nbytes = (17 + 2) * np.dtype(dt).itemsize * float(n)
        out[f"gbmv_{tag}_8x8_n2p20"] = {"GBps": nbytes / ms / 1e6, "frac_of_hbm_peak": nbytes / ms / 1e6 / HBM_PEAK_GBPS,
                                        "us": ms * 1e3, "bytes": nbytes}
        if dt == np.float64:
            # A'x (the transposed product of BandedRows / adjoint views), same rotation
            def one_t():
                A, x, y = sets[k[0] % len(sets)]
                k[0] += 1
                bh.mul(y, A, x, trans="T")

            warm(bh, one_t)
            ms = gpu_time(bh, one_t, max(30, steps * 6))
            out["gbmv_fp64_T_8x8_n2p20"] = {"GBps": nbytes / ms / 1e6, "frac_of_hbm_peak": nbytes / ms / 1e6 /
                                            HBM_PEAK_GBPS, "us": ms * 1e3, "bytes": nbytes}
        es = np.dtype(dt).itemsize
        # 64 right-hand sides: A read once, X and Y streamed
        X = bh.DeviceArray.from_numpy(np.asfortranarray(rng.standard_normal((n, 64)).astype(dt)))
        Y = bh.DeviceArray((n, 64), dt)
        A0 = sets[0][0]
        f = lambda: bh.mul(Y, A0, X)
        warm(bh, f)
        ms = gpu_time(bh, f, max(20, steps * 4))
        nbytes = (17 + 64 + 64) * float(es) * n
        out[f"gbmv_{tag}_8x8_64rhs_n2p20"] = {"GBps": nbytes / ms / 1e6, "frac_of_hbm_peak": nbytes / ms / 1e6 /
                                              HBM_PEAK_GBPS, "us": ms * 1e3, "bytes": nbytes}
        del X, Y
        # A*B: (8,8) x (4,4) -> (12,12)
        Bm = bh.BandedMatrix.from_host(np.asfortranarray(rng.standard_normal((9, n)).astype(dt)), n, 4, 4)
        Cm = bh.BandedMatrix.zeros((n, n), (12, 12), dt)
        g = lambda: bh.gbmm(A0, Bm, Cm)
        warm(bh, g)
        ms = gpu_time(bh, g, max(40, steps * 8))
        nbytes = (17 + 9 + 25) * float(es) * n
        out[f"gbmm_{tag}_8x8_4x4_n2p20"] = {"GBps": nbytes / ms / 1e6, "frac_of_hbm_peak": nbytes / ms / 1e6 /
                                            HBM_PEAK_GBPS, "us": ms * 1e3, "bytes": nbytes}
        del Bm, Cm, sets
    if not skip_cpu:
        # the reference's CPU path: one dgbmv per right-hand side, gbmm! per column
        import scipy.linalg.blas as sb
        from threadpoolctl import threadpool_limits

        from oracle import oracle as O

        Ah = np.asfortranarray(rng.standard_normal((17, n)))
        xh = rng.standard_normal(n)
        cpu = {}
        for thr, tag in ((1, "1core"), (os.cpu_count() or 1, "allcores")):
            with threadpool_limits(thr):
                sb.dgbmv(n, n, 8, 8, 1.0, Ah, xh)
                t0 = time.perf_counter()
                for _ in range(10):
                    sb.dgbmv(n, n, 8, 8, 1.0, Ah, xh)
                us = (time.perf_counter() - t0) / 10 * 1e6
                cpu[f"gbmv_fp64_us_{tag}"] = us
                cpu[f"gbmv_fp64_GBps_{tag}"] = 152.0 * n / us / 1e3
                Xh = np.asfortranarray(rng.standard_normal((n, 64)))
                t0 = time.perf_counter()
                for c in range(64):  # the reference's mul! loop: one gbmv per column
                    sb.dgbmv(n, n, 8, 8, 1.0, Ah, Xh[:, c])
                cpu[f"gbmv_fp64_64rhs_us_{tag}"] = (time.perf_counter() - t0) * 1e6
        Bh = np.asfortranarray(rng.standard_normal((9, n)))
        t0 = time.perf_counter()
        O.gbmm(n, n, n, 1.0, Ah, 8, 8, Bh, 4, 4, 0.0, np.zeros((25, n), order="F"), 12, 12)
        cpu["gbmm_fp64_us_1core"] = (time.perf_counter() - t0) * 1e6
        cpu["cores"] = os.cpu_count()
        cpu["sample"] = ("OpenBLAS 0.3.28 (scipy) dgbmv, mean of 10 calls per thread count; 64 RHS timed as the 64 "
                         "per-column calls the reference's mul! loop makes; A*B by the C restatement of gbmm!")
        out["cpu_bandwidth_n2p20"] = cpu
    return out


# ----------------------------------------------------------------- batching
def bench_batched(bh, skip_cpu):
    """4096 systems in one graph launch.  Algorithmic bytes per system: the
    padded band (2l+u+1) x n read and written once, the right-hand side read
    and the solution written: itemsize * n * (2(2l+u+1) + 2)."""
    n, kl, ku, bands, rhs = bh.c4_batch(4096, 0)
    slots = float(np.sum(n * (2 * (2 * kl + ku + 1) + 2)))
    out = {"systems": 4096, "rows_total": int(n.sum())}
    for dt, tag in ((np.float64, "fp64"), (np.float32, "fp32")):
        bt = bh.BatchedGBSV(n, kl, ku, dt)
        AB, B = bt.pack([b.astype(dt) for b in bands], [r.astype(dt) for r in rhs])
        bt.upload(AB, B)
        tuned_ms = bt.autotune()
        g = bt.graph()
        ms = []
        t = bh.Timer()
        for _ in range(5):
            bt.upload(AB, B)
            bh.sync()
            t.start()
            g.launch()
            t.stop()
            ms.append(t.elapsed_ms())
        g.close()
        med = float(np.median(ms))
        nbytes = slots * np.dtype(dt).itemsize
        out[tag] = {"ms_per_batch": med, "systems_per_s": 4096 / (med * 1e-3), "schedule": list(bt.schedule()),
                    "autotune_best_ms": tuned_ms,
                    "roofline": {"bound": "hbm", "achieved": nbytes / (med * 1e-3) / 1e9, "peak": HBM_PEAK_GBPS,
                                 "unit": "GB/s", "frac": nbytes / (med * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                                 "bytes_per_launch": nbytes}}
        bt.close()
    if not skip_cpu:
        import scipy.linalg.lapack as la
        from threadpoolctl import threadpool_limits

        with threadpool_limits(1):
            t0 = time.perf_counter()
            for s in range(4096):
                l, u = int(kl[s]), int(ku[s])
                ab = np.zeros((2 * l + u + 1, n[s]), order="F")
                ab[l:, :] = bands[s]
                lub, piv, info = la.dgbtrf(ab, l, u)
                la.dgbtrs(lub, l, u, rhs[s], piv)
            out["cpu_1core_ms_per_batch"] = (time.perf_counter() - t0) * 1e3
        out["cpu_sample"] = "all 4096 systems, scipy OpenBLAS dgbtrf+dgbtrs one after another on 1 thread, FP64"
    return out


# ----------------------------------------------------------------- differentiation
def bench_adjoint(bh, steps, skip_cpu):
    """Gradient of dot(w, A\\b) + 0.5||A x0||^2 at n = 262144, l = u = 16.
    Each piece is timed on its own beside the whole; algorithmic bytes per
    piece (FP64): LU reads the band and writes the padded factors
    ((l+u+1) + (2l+u+1)) 8n; each solve reads the factors and one vector and
    writes one ((2l+u+1) + 2) 8n; gbmv (l+u+1+2) 8n; the band-projected update
    reads four vectors and writes the band ((l+u+1) + 4) 8n."""
    import ctypes as C

    from banded_hip import _lib

    n, l, u = 262144, 16, 16
    rng = np.random.default_rng(4)
    A = np.asfortranarray(rng.standard_normal((l + u + 1, n)))
    A[u, :] += 2 * (l + u)
    b, w, x0 = (rng.standard_normal(n) for _ in range(3))
    dA = bh.BandedMatrix.from_host(A, n, l, u)
    db, dw, dx0 = (bh.DeviceArray.from_numpy(v) for v in (b, w, x0))
    bh.loss_and_grad(dA, db, dw, dx0, method="exact")
    bh.loss_and_grad(dA, db, dw, dx0, method="partitioned")
    bh.sync()
    # the two solves partitioned across the GPU (bh_dgbtrs_part) -- the
    # headline figure -- and with the bit-exact sequential chain beside it
    ms = gpu_time(bh, lambda: bh.loss_and_grad(dA, db, dw, dx0, method="partitioned"), max(2, steps // 2))
    ms_exact = gpu_time(bh, lambda: bh.loss_and_grad(dA, db, dw, dx0, method="exact"), max(2, steps // 2))
    Lp, gp, lp_, xp = bh.loss_and_grad(dA, db, dw, dx0, method="partitioned")
    Le, ge, le_, xe = bh.loss_and_grad(dA, db, dw, dx0, method="exact")
    gph, geh = gp.data.numpy(), ge.data.numpy()
    out = {"n": n, "l": l, "u": u, "ms_loss_and_grad": ms, "solves": "partitioned (bh_dgbtrs_part)",
           "ms_loss_and_grad_exact_solves": ms_exact,
           "grad_rel_diff_partitioned_vs_exact": float(np.abs(gph - geh).max() / np.abs(geh).max()),
           "loss_rel_diff_partitioned_vs_exact": abs(Lp - Le) / abs(Le)}
    # the pieces
    F = bh.lu(dA)
    x = db.copy()
    lam = dw.copy()
    y0 = bh.DeviceArray((n,), np.float64)
    G = bh.BandedMatrix.zeros((n, n), (l, u))
    hp = 2 * l + u + 1
    pieces = {
        "lu": (lambda: bh.lu(dA), ((l + u + 1) + hp) * 8.0 * n),
        "ldiv_N": (lambda: bh.ldiv_(F, x, method="partitioned"), (hp + 2) * 8.0 * n),
        "ldiv_N_exact": (lambda: bh.ldiv_(F, x), (hp + 2) * 8.0 * n),
        "gbmv": (lambda: bh.mul(y0, dA, dx0), (l + u + 3) * 8.0 * n),
        "ldiv_T": (lambda: bh.ldiv_(F.transpose(), lam, method="partitioned"), (hp + 2) * 8.0 * n),
        "ldiv_T_exact": (lambda: bh.ldiv_(F.transpose(), lam), (hp + 2) * 8.0 * n),
        "band_grad": (lambda: _lib.call("bh_dgb_band_grad", n, n, l, u, lam.ptr, x.ptr, y0.ptr, dx0.ptr, G.data.ptr,
                                        G.data.ld), (l + u + 1 + 4) * 8.0 * n),
    }
    per = {}
    for name, (fn, nbytes) in pieces.items():
        fn()
        bh.sync()
        pm = gpu_time(bh, fn, 3)
        per[name] = {"ms": pm, "bytes": nbytes, "GBps": nbytes / (pm * 1e-3) / 1e9,
                     "frac_of_hbm_peak": nbytes / (pm * 1e-3) / 1e9 / HBM_PEAK_GBPS}
    out["pieces"] = per
    big = max((k for k in per if not k.endswith("_exact")), key=lambda k: per[k]["ms"])
    out["roofline"] = {"kernel": f"{big} (the longest piece)", "bound": "hbm", "achieved": per[big]["GBps"],
                       "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": per[big]["frac_of_hbm_peak"],
                       "bytes_per_launch": per[big]["bytes"]}
    if not skip_cpu:
        import scipy.linalg.blas as sb
        import scipy.linalg.lapack as la
        from threadpoolctl import threadpool_limits

        def cpu_grad():
            ab = np.zeros((hp, n), order="F")
            ab[l:, :] = A
            lub, piv, info = la.dgbtrf(ab, l, u)
            xc, _ = la.dgbtrs(lub, l, u, b, piv)
            yc = sb.dgbmv(n, n, l, u, 1.0, A, x0)
            lc, _ = la.dgbtrs(lub, l, u, w, piv, trans=1)
            gA = np.empty_like(A)
            for t in range(l + u + 1):  # row t of the band holds A[j - u + t, j]
                j0, j1 = max(0, u - t), min(n, n + u - t)
                i0 = j0 - u + t
                i1 = i0 + (j1 - j0)
                gA[t, j0:j1] = yc[i0:i1] * x0[j0:j1] - lc[i0:i1] * xc[j0:j1]
            return float(w @ xc + 0.5 * yc @ yc), gA, lc

        for thr, tag in ((1, "1core"), (_host_threads(), "allcores")):
            with threadpool_limits(thr):
                cpu_grad()
                t0 = time.perf_counter()
                for _ in range(3):
                    cpu_grad()
                out[f"cpu_ms_loss_and_grad_{tag}"] = (time.perf_counter() - t0) / 3 * 1e3
        out["cpu_sample"] = ("full reference gradient: OpenBLAS dgbtrf + dgbtrs 'N' + dgbmv + dgbtrs 'T' + the band "
                             "projection (numpy, one vector op per band row); mean of 3 after 1 warm-up")
    return out


# ----------------------------------------------------------------- banded QR
def bench_qr(bh, steps, skip_cpu):
    """qr(A) and qr(A) \\ b at the reference's benchmark shapes
    (test/benchmark.jl:131-148): n = 100000 with (3,4) and n = 2000 with (400,400)."""
    out = {}
    for n, l, u in ((100000, 3, 4), (2000, 400, 400)):
        rng = np.random.default_rng(5)
        A_h = np.asfortranarray(rng.random((l + u + 1, n)))
        b_h = rng.standard_normal(n)
        A = bh.BandedMatrix.from_host(A_h, n, l, u)
        b = bh.DeviceArray.from_numpy(b_h)
        F = bh.qr(A)
        bh.sync()
        r = {"n": n, "l": l, "u": u}
        r["ms_qr"] = gpu_time(bh, lambda: bh.qr(A), max(2, steps // 2))
        r["ms_qr_ldiv"] = gpu_time(bh, lambda: bh.qr_ldiv_(F, b.copy()), max(2, steps // 2))
        x = bh.qr_ldiv_(F, b.copy()).numpy()
        if l + u <= 63:
            # Q'b by the reflectors, then R's solve partitioned across the GPU (to rounding)
            r["ms_qr_ldiv_partitioned"] = gpu_time(bh, lambda: bh.qr_ldiv_(F, b.copy(), method="partitioned"),
                                                   max(2, steps // 2))
            xp = bh.qr_ldiv_(F, b.copy(), method="partitioned").numpy()
            r["rel_diff_qr_ldiv_partitioned_vs_exact"] = float(np.abs(xp - x).max() / np.abs(x).max())
        import scipy.linalg.blas as sb

        res = b_h - sb.dgbmv(n, n, l, u, 1.0, A_h, x)
        anrm = np.abs(A_h).sum(axis=0).max()  # ||A||_1 bound on the band columns
        # normwise backward error: random uniform band matrices are exponentially
        # ill-conditioned in n, so the residual relative to ||b|| says nothing
        r["backward_error"] = float(np.abs(res).max() / (anrm * np.abs(x).max() + np.abs(b_h).max()))
        if not skip_cpu:
            from oracle import oracle as O

            ab = np.zeros((2 * l + u + 1, n), order="F")
            ab[l:, :] = A_h
            t0 = time.perf_counter()
            fac, tau = O.gbqrf(n, n, l, l + u, ab.copy(order="F"))
            r["cpu_ms_qr_1core"] = (time.perf_counter() - t0) * 1e3
            t0 = time.perf_counter()
            O.gbqrs(n, n, l, l + u, fac, tau, b_h.copy())
            r["cpu_ms_qr_ldiv_1core"] = (time.perf_counter() - t0) * 1e3
            r["cpu_sample"] = ("the C restatement of the reference's banded_qr! / lmul! + tbsv (bandedqr.jl), "
                               "1 thread, one call each")
        out[f"qr_n{n}_{l}x{u}"] = r
    return out


# ----------------------------------------------------------------- plain, CPU-runnable case
def bench_plain(bh, steps, skip_cpu):
    n, l, u = 10000, 4, 3
    rng = np.random.default_rng(0)
    A_h = np.asfortranarray(rng.random((l + u + 1, n)))
    b_h = rng.standard_normal(n)
    A = bh.BandedMatrix.from_host(A_h, n, l, u)
    b = bh.DeviceArray.from_numpy(b_h)
    y = bh.DeviceArray((n,), np.float64)
    C = bh.BandedMatrix.zeros((n, n), (8, 6))
    f_mv = lambda: bh.mul(y, A, b)
    f_mm = lambda: bh.gbmm(A, A, C)
    f_sv = lambda: bh.solve(A, b)
    Fp = bh.lu(A)
    xb = b.copy()
    f_lu = lambda: bh.lu(A)
    f_trs = lambda: bh.ldiv_(Fp, xb, method="partitioned")
    f_trs_exact = lambda: bh.ldiv_(Fp, xb)
    for f in (f_mv, f_mm, f_sv, f_lu, f_trs, f_trs_exact):
        f()
    bh.sync()
    out = {"n": n, "l": l, "u": u, "us_gbmv": gpu_time(bh, f_mv, 50) * 1e3, "us_gbmm": gpu_time(bh, f_mm, 50) * 1e3,
           "us_solve": gpu_time(bh, f_sv, 5) * 1e3, "us_lu": gpu_time(bh, f_lu, 5) * 1e3,
           "us_ldiv_partitioned": gpu_time(bh, f_trs, 20) * 1e3, "us_ldiv_exact": gpu_time(bh, f_trs_exact, 5) * 1e3}
    # FP64 backward error of the GPU solve, and its pivots against OpenBLAS's
    import scipy.linalg.blas as sb
    import scipy.linalg.lapack as la

    x = f_sv().numpy()
    r = b_h - sb.dgbmv(n, n, l, u, 1.0, A_h, x)
    arow = np.zeros(n)  # row sums of |A|: ||A||_inf
    for t in range(l + u + 1):
        j = np.arange(n)
        i = j - u + t
        ok = (i >= 0) & (i < n)
        np.add.at(arow, i[ok], np.abs(A_h[t, ok]))
    out["backward_error_fp64"] = float(np.abs(r).max() / (arow.max() * np.abs(x).max() + np.abs(b_h).max()))
    ab = np.zeros((2 * l + u + 1, n), order="F")
    ab[l:, :] = A_h
    F = bh.lu(A)
    _, piv_cpu, _ = la.dgbtrf(ab, l, u)
    out["pivots_equal_cpu"] = f"{int(np.sum(F.ipiv == piv_cpu + 1))}/{n}"
    if not skip_cpu:
        from threadpoolctl import threadpool_limits

        from oracle import oracle as O

        with threadpool_limits(1):
            t0 = time.perf_counter()
            for _ in range(20):
                lub, piv, info = la.dgbtrf(ab, l, u)
                la.dgbtrs(lub, l, u, b_h, piv)
            out["cpu_us_solve_1core"] = (time.perf_counter() - t0) / 20 * 1e6
            t0 = time.perf_counter()
            for _ in range(200):
                sb.dgbmv(n, n, l, u, 1.0, A_h, b_h)
            out["cpu_us_gbmv_1core"] = (time.perf_counter() - t0) / 200 * 1e6
        t0 = time.perf_counter()
        for _ in range(20):
            O.gbmm(n, n, n, 1.0, A_h, l, u, A_h, l, u, 0.0, np.zeros((15, n), order="F"), 8, 6)
        out["cpu_us_gbmm_1core"] = (time.perf_counter() - t0) / 20 * 1e6
        out["cpu_sample"] = ("OpenBLAS 0.3.28 (scipy) dgbmv / dgbtrf+dgbtrs on 1 thread; A*A by the C restatement "
                             "of gbmm! (the reference's generic kernel), 1 thread")
    return out


def bench_triangular(bh, steps, skip_cpu):
    """Triangular band solve and product (tbsv!/tbmv!), FP64, at the headline's
    band (n = 65536, k = 128) and a narrow long one (n = 2^20, k = 8)."""
    import scipy.linalg.blas as sb

    out = {}
    for n, k in ((65536, 128), (1 << 20, 8)):
        rng = np.random.default_rng(1)
        a_h = np.asfortranarray(rng.uniform(-1.0, 1.0, size=(k + 1, n)) * (0.5 / k))
        a_h[k, :] = 1.0 + rng.random(n)  # upper: diagonal in row k
        aL_h = np.asfortranarray(np.roll(a_h, -k, axis=0))  # lower: diagonal in row 0
        x_h = rng.standard_normal(n)
        A, AL = bh.DeviceArray.from_numpy(a_h), bh.DeviceArray.from_numpy(aL_h)
        X = bh.DeviceArray.from_numpy(x_h)
        r = {"n": n, "k": k}
        for tag, uplo, trans, M in (("tbsv_UN", "U", "N", A), ("tbsv_LN", "L", "N", AL), ("tbsv_UT", "U", "T", A)):
            f = lambda: bh.tbsv_(uplo, trans, "N", n, k, M, X)
            f()
            X.upload(x_h)
            r[f"us_{tag}"] = gpu_time(bh, f, max(2, steps)) * 1e3
            X.upload(x_h)
        r["ns_per_step_tbsv_UN"] = r["us_tbsv_UN"] * 1e3 / n
        if k <= 63:
            # the upper triangle through the partitioned band solve (kl = 0), to rounding
            T = bh.UpperTriangular(bh.BandedMatrix.from_host(a_h, n, 0, k))
            for trans in ("N", "T"):
                X.upload(x_h)
                xe = T.ldiv_(X, trans=trans).numpy()
                X.upload(x_h)
                xp = T.ldiv_(X, trans=trans, method="partitioned").numpy()
                r[f"rel_diff_partitioned_vs_exact_U{trans}"] = float(np.abs(xp - xe).max() / np.abs(xe).max())
                fp = lambda: T.ldiv_(X, trans=trans, method="partitioned")
                r[f"us_tbsv_U{trans}_partitioned"] = gpu_time(bh, fp, max(5, steps)) * 1e3
            X.upload(x_h)
            del T
        XB = bh.DeviceArray.from_numpy(np.tile(x_h[:, None], (1, 64)))
        fb = lambda: bh.tbsv_("U", "N", "N", n, k, A, XB)
        fb()
        r["us_tbsv_UN_64rhs"] = gpu_time(bh, fb, 2) * 1e3
        del XB
        fm = lambda: bh.tbmv_("U", "N", "N", n, k, A, X)
        fm()
        r["us_tbmv_UN"] = gpu_time(bh, fm, 20) * 1e3
        r["tbmv_GBps_algorithmic"] = ((k + 1) * n + 2 * n) * 8 / (r["us_tbmv_UN"] * 1e-6) / 1e9
        if not skip_cpu:
            from threadpoolctl import threadpool_limits

            with threadpool_limits(1):
                t0 = time.perf_counter()
                for _ in range(5):
                    sb.dtbsv(k, a_h, x_h)
                r["cpu_us_tbsv_UN_1core"] = (time.perf_counter() - t0) / 5 * 1e6
                r["cpu_us_tbsv_UN_64rhs_1core"] = 64 * r["cpu_us_tbsv_UN_1core"]  # 64 independent calls
                t0 = time.perf_counter()
                for _ in range(5):
                    sb.dtbmv(k, a_h, x_h)
                r["cpu_us_tbmv_UN_1core"] = (time.perf_counter() - t0) / 5 * 1e6
            r["cpu_sample"] = "OpenBLAS 0.3.28 (scipy) dtbsv / dtbmv, upper 'N', 1 thread, 5 calls each"
        out[f"triangular_n{n}_k{k}"] = r
    return out


def bench_elementwise(bh, steps, skip_cpu):
    """Element-wise operations on band data (lmul!, axpy!, copyto! with a band
    change), FP64, n = 2^20, bands (8, 8): HBM-bound, so reported as algorithmic
    GB/s (every in-matrix slot read and written once)."""
    n, l, u = 1 << 20, 8, 8
    h = l + u + 1
    rng = np.random.default_rng(2)
    D = np.asfortranarray(rng.standard_normal((h, n)))
    # three independent sets, used in turn, so no launch finds its operands in
    # the 256 MB last-level cache (each set holds 460 MB)
    sets = [(bh.BandedMatrix.from_host(D, n, l, u), bh.BandedMatrix.from_host(D, n, l, u),
             bh.BandedMatrix.undef((n, n), (l + 2, u + 2))) for _ in range(3)]
    it = {"k": 0}

    def nxt():
        it["k"] += 1
        return sets[it["k"] % 3]

    def f_ax():
        X, Y, _ = nxt()
        bh.axpy_(1e-3, X, Y)

    def f_sc():
        _, Y, _ = nxt()
        Y.lmul_(0.999)

    def f_cp():
        X, _, W = nxt()
        bh.copyto_(W, X)

    for f in (f_ax, f_sc, f_cp):
        f()
    reps = max(21, steps * 3 * 2)
    slots = h * n * 8
    r = {"n": n, "l": l, "u": u}
    r["us_axpy"] = gpu_time(bh, f_ax, reps) * 1e3
    r["axpy_GBps"] = 3 * slots / (r["us_axpy"] * 1e-6) / 1e9
    r["us_lmul"] = gpu_time(bh, f_sc, reps) * 1e3
    r["lmul_GBps"] = 2 * slots / (r["us_lmul"] * 1e-6) / 1e9
    r["us_copy_band_change"] = gpu_time(bh, f_cp, reps) * 1e3
    r["copy_GBps"] = (slots + (h + 4) * n * 8) / (r["us_copy_band_change"] * 1e-6) / 1e9
    if not skip_cpu:
        from threadpoolctl import threadpool_limits

        Yh = D.copy()
        with threadpool_limits(1):
            t0 = time.perf_counter()
            for _ in range(10):
                Yh += 1e-3 * D  # a*X + Y on the band array, as banded_generic_axpy! loops over it
            r["cpu_us_axpy_1core"] = (time.perf_counter() - t0) / 10 * 1e6
            t0 = time.perf_counter()
            for _ in range(10):
                Yh *= 0.999
            r["cpu_us_lmul_1core"] = (time.perf_counter() - t0) / 10 * 1e6
        r["cpu_sample"] = "numpy on the 17 x 2^20 band array, 1 thread, 10 calls each"
    return r


def bench_cholesky(bh, steps, skip_cpu):
    """Symmetric positive definite band Cholesky (pbtrf), its solve (pbtrs) and
    the symmetric product (sbmv), FP64, upper storage, n = 65536 at kd = 8 and
    kd = 128."""
    import ctypes as C

    import scipy.linalg.blas as sb
    import scipy.linalg.lapack as la

    out = {}
    for n, kd in ((65536, 8), (65536, 128)):
        rng = np.random.default_rng(3)
        a = np.asfortranarray(rng.uniform(-1.0, 1.0, size=(kd + 1, n)))
        a[kd, :] = kd + 1 + rng.random(n)
        b = rng.standard_normal(n)
        A = bh.DeviceArray.from_numpy(a)
        info = C.c_int64()
        f_tr = lambda: (A.upload(a), bh._lib.call("bh_dpbtrf", b"U", n, kd, A.ptr, A.ld, C.byref(info)))
        f_tr()
        r = {"n": n, "kd": kd}
        t = bh.Timer()
        ms = []
        for _ in range(max(2, steps // 2)):
            A.upload(a)
            t.start()
            bh._lib.call("bh_dpbtrf", b"U", n, kd, A.ptr, A.ld, C.byref(info))
            t.stop()
            ms.append(t.elapsed_ms())
        r["ms_pbtrf"] = float(np.median(ms))
        X = bh.DeviceArray.from_numpy(b)
        f_rs = lambda: bh._lib.call("bh_dpbtrs", b"U", n, kd, 1, A.ptr, A.ld, X.ptr, n)
        f_rs()
        r["ms_pbtrs"] = gpu_time(bh, f_rs, max(2, steps // 2))
        if kd <= 63:
            # U^T y = b, U x = y by the partitioned band solve (kl = 0, identity pivots), to rounding
            nb = C.c_size_t()
            bh._lib.call("bh_gbtrs_part_workspace", b"d", n, 0, kd, 1, C.byref(nb))
            ws = bh.DeviceArray((nb.value,), np.uint8)
            ip = bh.DeviceArray.from_numpy(np.arange(1, n + 1, dtype=np.int64))

            def f_part():
                for tr in (b"T", b"N"):
                    bh._lib.call("bh_dgbtrs_part", tr, n, 0, kd, 1, A.ptr, A.ld, ip.ptr, X.ptr, n, ws.ptr, nb.value)

            X.upload(b)
            f_rs()
            xe = X.numpy()
            X.upload(b)
            f_part()
            r["rel_diff_pbtrs_partitioned_vs_exact"] = float(np.abs(X.numpy() - xe).max() / np.abs(xe).max())
            r["ms_pbtrs_partitioned"] = gpu_time(bh, f_part, max(5, steps))
            del ws, ip
        Y = bh.DeviceArray((n,), np.float64)
        f_mv = lambda: bh._lib.call("bh_dsbmv", b"U", n, kd, 1.0, A.ptr, A.ld, X.ptr, 1, 0.0, Y.ptr, 1)
        f_mv()
        r["us_sbmv"] = gpu_time(bh, f_mv, 20) * 1e3
        r["sbmv_GBps_algorithmic"] = ((kd + 1) * n + 2 * n) * 8 / (r["us_sbmv"] * 1e-6) / 1e9
        if not skip_cpu:
            from threadpoolctl import threadpool_limits

            with threadpool_limits(1):
                t0 = time.perf_counter()
                for _ in range(3):
                    c, _ = la.dpbtrf(a)
                r["cpu_ms_pbtrf_1core"] = (time.perf_counter() - t0) / 3 * 1e3
                t0 = time.perf_counter()
                for _ in range(3):
                    la.dpbtrs(c, b)
                r["cpu_ms_pbtrs_1core"] = (time.perf_counter() - t0) / 3 * 1e3
                t0 = time.perf_counter()
                for _ in range(10):
                    sb.dsbmv(kd, 1.0, a, b)
                r["cpu_us_sbmv_1core"] = (time.perf_counter() - t0) / 10 * 1e6
            r["cpu_sample"] = "OpenBLAS 0.3.28 (scipy) dpbtrf / dpbtrs / dsbmv, upper, 1 thread"
        out[f"cholesky_n{n}_kd{kd}"] = r
    return out


def traffic_probe(bh, name):
    """A few launches of one kernel plus a 256 MiB device copy, for counter runs."""
    import ctypes as C

    from banded_hip import _lib

    src = bh.DeviceArray((1 << 25,), np.float64)
    dst = bh.DeviceArray((1 << 25,), np.float64)
    for _ in range(3):
        _lib.call("bh_memcpy_dtod", dst.ptr, src.ptr, src.nbytes)
    if name == "gbmv64":
        n = 1 << 20
        A = bh.BandedMatrix.from_host(np.asfortranarray(np.random.default_rng(1).standard_normal((17, n))), n, 8, 8)
        x = bh.DeviceArray.from_numpy(np.ones(n))
        y = bh.DeviceArray((n,), np.float64)
        for _ in range(3):
            bh.mul(y, A, x)
    elif name == "gbmv64x64" or name == "gbmm64":
        n = 1 << 20
        rng = np.random.default_rng(1)
        A = bh.BandedMatrix.from_host(np.asfortranarray(rng.standard_normal((17, n))), n, 8, 8)
        if name == "gbmv64x64":
            X = bh.DeviceArray.from_numpy(np.asfortranarray(rng.standard_normal((n, 64))))
            Y = bh.DeviceArray((n, 64), np.float64)
            for _ in range(3):
                bh.mul(Y, A, X)
        else:
            Bm = bh.BandedMatrix.from_host(np.asfortranarray(rng.standard_normal((9, n))), n, 4, 4)
            Cm = bh.BandedMatrix.zeros((n, n), (12, 12))
            for _ in range(3):
                bh.gbmm(A, Bm, Cm)
    elif name == "lu32":
        n, l, u = 65536, 128, 128
        A_h, _ = make_precision_problem(n, l, u, 1)
        ab = bh.pad_for_lu(bh.BandedMatrix.from_host(A_h, n, l, u), np.float32)
        ipiv = bh.DeviceArray((n,), np.int64)
        info = C.c_int64()
        _lib.call("bh_sgbtrf", n, n, l, u, ab.data.ptr, ab.data.ld, ipiv.ptr, C.byref(info))
    elif name == "trs32":
        n, l, u, nrhs = 65536, 128, 128, 16
        A_h, B_h = make_precision_problem(n, l, u, nrhs)
        ab = bh.pad_for_lu(bh.BandedMatrix.from_host(A_h, n, l, u), np.float32)
        ipiv = bh.DeviceArray((n,), np.int64)
        info = C.c_int64()
        _lib.call("bh_sgbtrf", n, n, l, u, ab.data.ptr, ab.data.ld, ipiv.ptr, C.byref(info))
        nb = C.c_size_t()
        _lib.call("bh_gbtrs_blocked_workspace", b"s", n, l, u, C.byref(nb))
        work = bh.DeviceArray((nb.value,), np.uint8)
        B32 = bh.DeviceArray.from_numpy(np.asfortranarray(B_h.astype(np.float32)))
        _lib.call("bh_sgbtrs_blocked", n, l, u, nrhs, ab.data.ptr, ab.data.ld, ipiv.ptr, B32.ptr, n, work.ptr,
                  nb.value)
    elif name == "part64":
        # the adjoint workload's matrix: one factorisation, then three 'N' and
        # three 'T' partitioned solves
        n, l, u = 262144, 16, 16
        rng = np.random.default_rng(1)
        A = np.asfortranarray(rng.standard_normal((l + u + 1, n)))
        A[u, :] += 2 * (l + u)
        F = bh.lu(bh.BandedMatrix.from_host(A, n, l, u))
        b = bh.DeviceArray.from_numpy(rng.standard_normal(n))
        for _ in range(3):
            bh.ldiv_(F, b.copy(), method="partitioned")
            bh.ldiv_(F.transpose(), b.copy(), method="partitioned")
    bh.sync()
    print(json.dumps({"traffic_probe": name, "copy_bytes": src.nbytes}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--noshift", action="store_true", help="the pivoting variant (no diagonal shift)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--traffic-probe", default=None)
    ap.add_argument("--only", choices=["triangular", "elementwise", "cholesky", "bandwidth"], default=None,
                    help="run one side workload and print its figures as one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU per process (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design, not an implementation.", file=sys.stderr)
        sys.exit(2)

    import banded_hip as bh

    L = bh.load_library()
    if L.bh_init(0) != 0:
        raise RuntimeError(L.bh_last_error().decode())
    if args.traffic_probe:
        traffic_probe(bh, args.traffic_probe)
        return
    if args.only == "triangular":
        print(json.dumps(bench_triangular(bh, args.steps, args.skip_cpu)), flush=True)
        return
    if args.only == "cholesky":
        print(json.dumps(bench_cholesky(bh, args.steps, args.skip_cpu)), flush=True)
        return
    if args.only == "elementwise":
        print(json.dumps(bench_elementwise(bh, args.steps, args.skip_cpu)), flush=True)
        return
    if args.only == "bandwidth":
        print(json.dumps(bench_bandwidth(bh, args.steps, args.skip_cpu)), flush=True)
        return

    log("precision workload")
    P = bench_precision(bh, args.steps, args.warmup, shift=not args.noshift)
    n, l, u, nrhs = P["n"], P["l"], P["u"], P["nrhs"]
    extra = {}
    if not args.headline_only:
        log("bandwidth-bound kernels")
        extra.update(bench_bandwidth(bh, args.steps, args.skip_cpu))
        log("batched systems")
        extra["batched_4096"] = bench_batched(bh, args.skip_cpu)
        log("differentiation")
        extra["adjoint_n262144_16x16"] = bench_adjoint(bh, args.steps, args.skip_cpu)
        log("plain case")
        extra["plain_n1e4_4x3"] = bench_plain(bh, args.steps, args.skip_cpu)
        log("triangular band solve and product")
        extra.update(bench_triangular(bh, args.steps, args.skip_cpu))
        log("element-wise band operations")
        extra["elementwise_n2p20_8x8"] = bench_elementwise(bh, args.steps, args.skip_cpu)
        log("symmetric band Cholesky and product")
        extra.update(bench_cholesky(bh, args.steps, args.skip_cpu))
        log("banded QR")
        extra.update(bench_qr(bh, args.steps, args.skip_cpu))
        log("precision workload, no diagonal shift (pivoting active)")
        P2 = bench_precision(bh, max(1, args.steps // 2), 1, shift=False)
        ns = {"mixed_solves_per_s": 1000.0 / P2["ms_mixed"], "fp64_gpu_solves_per_s": 1000.0 / P2["ms_fp64"],
              "refinement_iters": P2["iters"],
              "backward_error_mixed": backward_error(P2["A_h"], P2["l"], P2["u"], P2["x_mixed"], P2["B_h"]),
              "backward_error_fp64_gpu": backward_error(P2["A_h"], P2["l"], P2["u"], P2["x_fp64"], P2["B_h"])}
        if not args.skip_cpu:
            xc2, pivc2, tc2 = cpu_reference_solve(P2["A_h"], P2["B_h"], P2["l"], P2["u"], _host_threads())
            fe2 = lambda x: float(np.abs(x - xc2).max() / np.abs(xc2).max())
            ns.update({"cpu_solves_per_s": 1.0 / tc2, "forward_error_mixed": fe2(P2["x_mixed"]),
                       "forward_error_fp64_gpu": fe2(P2["x_fp64"]),
                       "pivots_fp64_gpu_equal_cpu": f"{int(np.sum(P2['ipiv64'] == pivc2))}/{P2['n']}"})
        extra["precision_noshift_n65536_128x128"] = ns
    else:
        extra.update(bench_bandwidth(bh, args.steps))

    out = {}
    if not args.skip_cpu:
        log("CPU baseline")
        allt = args.cpu_threads or _host_threads()
        per = {}
        for thr in sorted({1, allt}):
            for _ in range(2):  # warm-up
                cpu_reference_solve(P["A_h"], P["B_h"], l, u, thr)
            ts = []
            for _ in range(10):
                xc, pivc, tc = cpu_reference_solve(P["A_h"], P["B_h"], l, u, thr)
                ts.append(tc)
            per[thr] = float(np.median(ts))
        best = min(per, key=per.get)
        out["cpu_baseline"] = {
            "value": 1.0 / per[best], "unit": "solves/s", "cores": best, "kind": "reference",
            "sample": ("the whole workload (n=65536, l=u=128, 16 RHS), FP64, OpenBLAS 0.3.28 dgbtrf+dgbtrs via "
                       "scipy -- the LAPACK routines BandedLU.jl:98 / linalg.jl:28 call; median of 10 after 2 "
                       "warm-ups, at 1 thread and at every host thread; value = the faster"),
            "solves_per_s_by_threads": {str(k): 1.0 / v for k, v in per.items()},
            "cpu_model": _cpu_model(), "nproc": _host_threads(),
        }
        fe = lambda x: float(np.abs(x - xc).max() / np.abs(xc).max())
        out["forward_error_vs_cpu_fp64"] = {"mixed": fe(P["x_mixed"]), "fp64_gpu": fe(P["x_fp64"])}
        out["backward_error"] = {"mixed": backward_error(P["A_h"], l, u, P["x_mixed"], P["B_h"]),
                                 "fp64_gpu": backward_error(P["A_h"], l, u, P["x_fp64"], P["B_h"]),
                                 "cpu": backward_error(P["A_h"], l, u, xc, P["B_h"])}
        out["pivots_fp64_gpu_equal_cpu"] = f"{int(np.sum(P['ipiv64'] == pivc))}/{n}"

    g = extra["gbmv_fp64_8x8_n2p20"]
    lu_fl = lu_flops(n, l, u)
    lu_tf = lu_fl / (P["lu32_ms"] * 1e-3) / 1e12
    trs_fl = solve_flops(n, l, u, nrhs)
    trs_tf = trs_fl / (P["trs32_ms"] * 1e-3) / 1e12
    traffic = traffic_gbmv = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("lu32_bytes_per_launch")
        traffic_gbmv = tj.get("gbmv64_bytes_per_launch")
    line = {
        "metric": "A\\b solves/s n=65536 l=u=128 mixed vs FP64 at 1e-12 err; gbmv HBM GB/s",
        "value": 1000.0 / P["ms_mixed"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": P["ms_mixed"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor + f64 refine",
        "data": "synthetic (seeded N(0,1) band, diagonal +2(l+u); 16 RHS N(0,1))" if not args.noshift else
                "synthetic (seeded N(0,1) band, no shift; 16 RHS N(0,1))",
        "config": {"workload": "precision: A\\b n=65536 l=u=128 16 RHS" + (" (no shift)" if args.noshift else ""),
                   "n": n, "l": l, "u": u, "nrhs": nrhs},
        "fp64_gpu_solves_per_s": 1000.0 / P["ms_fp64"],
        "ms_per_step_fp64_gpu": P["ms_fp64"],
        "refinement_iters": P["iters"],
        "gbmv_GBps": g["GBps"],
        "roofline": {
            "kernel": "gbtrf_mfma_kernel<float,3> (FP32 blocked band LU, MFMA trailing updates; the largest "
                      "kernel of the mixed A\\b); kernel_ms is the bh_sgbtrf call (zero-fill pre-pass + LU)",
            "bound": "mfma", "achieved": lu_tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": lu_tf / FP32_PEAK_TFLOPS, "traffic": traffic, "kernel_ms": P["lu32_ms"],
            "flops_per_launch": lu_fl,
        },
        "roofline_solve": {"kernel": "trs_prep<float,32> + trs_fwd<float,32,8,1> + trs_bwd<float,32,8,2> "
                                     "(bh_sgbtrs_blocked, 16 RHS, FP32)", "bound": "mfma", "achieved": trs_tf,
                           "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": trs_tf / FP32_PEAK_TFLOPS,
                           "kernel_ms": P["trs32_ms"], "flops_per_launch": trs_fl,
                           "bit_exact_ring_solve_ms": P["trs32_ring_ms"]},
        "roofline_gbmv": {"kernel": "gbmv_n_wave<double,1,11,2> (FP64 gbmv 'N', n=2^20, (8,8))", "bound": "hbm",
                          "achieved": g["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                          "frac": g["GBps"] / HBM_PEAK_GBPS, "traffic": traffic_gbmv,
                          "bytes_per_launch": g["bytes"], "kernel_us": g["us"]},
    }
    line.update(out)
    line["workloads"] = extra
    print(json.dumps(line))


if __name__ == "__main__":
    main()
