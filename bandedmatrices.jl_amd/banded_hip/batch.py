"""Batched band solves: many independent square systems A_s x_s = b_s in one
persistent launch (no reference counterpart: the reference factors one
matrix per call, src/banded/BandedLU.jl:90-111; callers loop lu(A_s) \\ b_s).

Per system the results equal LAPACK dgbtf2 + dgbtrs 'N' bit for bit: the
factors (padded band storage, as lu(A).factors), pivots (1-based) and info,
and x overwriting b.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .bandedmatrix import _prefix
from .device import DeviceArray


class BatchGraph:
    """A captured batched solve bound to its plan's buffers; launch() replays it.
    It keeps the BatchedGBSV (plan and buffers) alive; while it exists the
    plan's schedule cannot change (schedule/autotune raise)."""

    def __init__(self, handle, owner):
        self._h = handle
        self._owner = owner

    def launch(self):
        _lib.call("bh_batch_graph_launch", self._h)

    def close(self):
        if self._h:
            _lib.load().bh_batch_graph_destroy(self._h)
            self._h = None

    __del__ = close


class BatchedGBSV:
    """A batch of square band systems with bands (kl_s, ku_s), kl_s <= 8, kl_s + ku_s <= 63."""

    def __init__(self, n, kl, ku, dtype=np.float64):
        self._graphs = []
        self.n = np.asarray(n, np.int64)
        self.kl = np.asarray(kl, np.int64)
        self.ku = np.asarray(ku, np.int64)
        self.dtype = np.dtype(dtype)
        self.nsys = len(self.n)
        plan = C.c_void_p()
        _lib.call("bh_batch_plan_create", self.nsys, self.n.ctypes.data, self.kl.ctypes.data, self.ku.ctypes.data,
                  C.byref(plan))
        self._plan = plan
        self.ab_off = np.zeros(self.nsys, np.int64)
        self.vec_off = np.zeros(self.nsys, np.int64)
        abe, vce = C.c_int64(), C.c_int64()
        _lib.call("bh_batch_plan_layout", plan, self.ab_off.ctypes.data, self.vec_off.ctypes.data, C.byref(abe),
                  C.byref(vce))
        self.ab_elems, self.vec_elems = abe.value, vce.value
        self.AB = DeviceArray((max(self.ab_elems, 1),), self.dtype)
        self.B = DeviceArray((max(self.vec_elems, 1),), self.dtype)
        self.ipiv = DeviceArray((max(self.vec_elems, 1),), np.int64)
        self.info = DeviceArray((max(self.nsys, 1),), np.int64)

    def ldab(self, s):
        return int(2 * self.kl[s] + self.ku[s] + 1)

    def pack(self, bands, rhs):
        """Host arrays for the packed layout: bands[s] is (kl+ku+1) x n band data
        (BandedMatrix layout), rhs[s] length n.  Returns (AB, B) host arrays."""
        AB = np.zeros(max(self.ab_elems, 1), self.dtype)
        B = np.zeros(max(self.vec_elems, 1), self.dtype)
        for s in range(self.nsys):
            n, kl, ku = int(self.n[s]), int(self.kl[s]), int(self.ku[s])
            ld = 2 * kl + ku + 1
            blk = AB[self.ab_off[s]:self.ab_off[s] + ld * n].reshape(n, ld)  # column-major view
            blk[:, kl:] = np.asarray(bands[s], self.dtype).T
            B[self.vec_off[s]:self.vec_off[s] + n] = rhs[s]
        return AB, B

    def upload(self, AB, B):
        _lib.call("bh_memcpy_htod", self.AB.ptr, np.ascontiguousarray(AB, self.dtype).ctypes.data,
                  self.AB.nbytes)
        _lib.call("bh_memcpy_htod", self.B.ptr, np.ascontiguousarray(B, self.dtype).ctypes.data, self.B.nbytes)

    def solve(self):
        """Factor and solve every system in place (stream-ordered)."""
        _lib.call(f"bh_{_prefix(self.dtype)}gbsv_batched", self._plan, self.AB.ptr, self.B.ptr, self.ipiv.ptr,
                  self.info.ptr)

    def autotune(self) -> float:
        """Pick the fastest schedule on scratch copies of the current inputs."""
        ms = C.c_double()
        _lib.call(f"bh_{_prefix(self.dtype)}gbsv_batched_autotune", self._plan, self.AB.ptr, self.B.ptr,
                  C.byref(ms))
        return ms.value

    def schedule(self, order=None, waves_per_block=None, blocks=None):
        o, w, b = C.c_int(), C.c_int(), C.c_int()
        _lib.call("bh_batch_plan_get_schedule", self._plan, C.byref(o), C.byref(w), C.byref(b))
        if order is None and waves_per_block is None and blocks is None:
            return o.value, w.value, b.value
        _lib.call("bh_batch_plan_schedule", self._plan, o.value if order is None else order,
                  w.value if waves_per_block is None else waves_per_block, b.value if blocks is None else blocks)
        return self.schedule()

    def graph(self) -> BatchGraph:
        g = C.c_void_p()
        _lib.call(f"bh_{_prefix(self.dtype)}gbsv_batched_graph", self._plan, self.AB.ptr, self.B.ptr, self.ipiv.ptr,
                  self.info.ptr, C.byref(g))
        bg = BatchGraph(g, self)
        self._graphs.append(bg)
        return bg

    def results(self):
        """(x, factors, ipiv, info) per system as host arrays."""
        AB, B, P, I = self.AB.numpy(), self.B.numpy(), self.ipiv.numpy(), self.info.numpy()
        out = []
        for s in range(self.nsys):
            n, ld = int(self.n[s]), self.ldab(s)
            f = AB[self.ab_off[s]:self.ab_off[s] + ld * n].reshape(n, ld).T
            out.append((B[self.vec_off[s]:self.vec_off[s] + n].copy(), np.asfortranarray(f),
                        P[self.vec_off[s]:self.vec_off[s] + n].copy(), int(I[s])))
        return out

    def close(self):
        for g in getattr(self, "_graphs", []):
            g.close()
        self._graphs = []
        if getattr(self, "_plan", None):
            _lib.load().bh_batch_plan_destroy(self._plan)
            self._plan = None

    __del__ = close


def c4_batch(nsys=4096, seed=0, nmin=256, nmax=8192, classes=((1, 1), (2, 2), (4, 3), (8, 8))):
    """Synthetic batch of the batching workload: n log-uniform in [nmin, nmax],
    (kl, ku) uniform over the classes, band entries U[0,1) with the diagonal
    shifted by kl+ku+1, right-hand sides N(0,1)."""
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys))).astype(np.int64)
    cls = rng.integers(0, len(classes), nsys)
    kl = np.array([classes[c][0] for c in cls], np.int64)
    ku = np.array([classes[c][1] for c in cls], np.int64)
    bands, rhs = [], []
    for s in range(nsys):
        d = rng.random((kl[s] + ku[s] + 1, n[s]))
        d[ku[s], :] += kl[s] + ku[s] + 1
        bands.append(d)
        rhs.append(rng.standard_normal(n[s]))
    return n, kl, ku, bands, rhs
