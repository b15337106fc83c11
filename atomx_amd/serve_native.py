"""Native serving loop (``ax_server_*``): continuous batching of all-quartet
R^k requests by a C++ scheduler thread, with no Python on the launch path.

Requests are the ``SlaterRequest`` objects of ``atomx_amd.serve`` (grid,
angular momenta, kmax, precision, host orbital samples, and an optional
``state``: a resident state keeps the plan of its first request for the
later ones, which must have the same shape).  The server reads a
request's arrays when its launch set is prepared, so the wrapper keeps them
alive until the request completes.  Results are ``SlaterResult`` objects
whose values are views of the server's pinned result buffer; the buffer
slice is handed back to the server when the result's values are garbage
collected.  The same request gives bitwise the same values here, in the
Python server and in a plan run alone.
"""

from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import device as dv
from ._lib import PRECISIONS, call, check, load
from .serve import SlaterResult


def _addr(a):
    """Address of a contiguous array's data."""
    return a.__array_interface__["data"][0]


def _rank_layout(ls, kmax):
    """(m_k, packed offset of rank k) of the all-pairs tensor, by angular
    momentum class (the counts of ``valid_pairs``)."""
    cnt = {}
    for l in np.asarray(ls).tolist():
        cnt[l] = cnt.get(l, 0) + 1
    cls = sorted(cnt.items())
    m = [0] * (int(kmax) + 1)
    for i, (l1, c1) in enumerate(cls):
        for l2, c2 in cls[i:]:
            n = c1 * (c1 + 1) // 2 if l1 == l2 else c1 * c2
            for k in range(l2 - l1, min(l1 + l2, int(kmax)) + 1, 2):
                m[k] += n
    offs, o = [], 0
    for x in m:
        offs.append(o)
        o += x * (x + 1) // 2
    return m, offs


class _Release:
    """Hands a completed ticket back to the server when its values go away."""

    def __init__(self, server, ticket):
        self._server, self._ticket = server, ticket

    def __del__(self):
        s = self._server
        if s is not None and s._h.value:
            try:
                s._lib.ax_server_release(s._h, ctypes.c_int64(self._ticket))
            except Exception:
                pass


class NativeSlaterServer:
    """``SlaterServer`` for all-pairs R^k requests with the scheduler, plan
    set-up, packing and launches in native code.

    ``budget_bytes`` bounds the device bytes of the requests in flight (plans
    and staged inputs and outputs; default: 80 % of the free device memory),
    ``max_inflight`` launch sets at most are in flight on their own streams,
    each of at most ``max_batch`` requests."""

    def __init__(self, budget_bytes=None, max_inflight=4, max_batch=16):
        dv.require_gpu()
        self._lib = load()
        if budget_bytes is None:
            free, _ = dv.torch().cuda.mem_get_info()
            budget_bytes = int(0.8 * free)
        h = ctypes.c_void_p()
        call("ax_server_create", int(budget_bytes), int(max_inflight), int(max_batch),
             ctypes.byref(h))
        self._h = h
        self._keep = {}  # ticket -> (request, arrays) until the request completes
        self._states = {}  # request state -> the server's state key
        self._geo = {}     # request state -> its validated geometry (``_geometry``)
        self._layouts = {}  # (l bytes, kmax) -> rank layout of the packed values
        self._next_key = 1
        self._submit = self._lib.ax_server_submit
        self._mu = threading.Lock()

    def _geometry(self, req):
        """(grid, weights, l as contiguous arrays, their addresses, state
        key) of a request, validated once per state while the request keeps
        passing the same grid, weights and l objects (the server compares
        the grid values against its state's on every submit)."""
        state = getattr(req, "state", None)
        g = self._geo.get(state) if state is not None else None
        if g is not None and g[0] is req.gr and g[1] is req.grw and g[2] is req.ls:
            return g
        gr = np.ascontiguousarray(req.gr, dtype=np.float64)
        grw = np.ascontiguousarray(req.grw, dtype=np.float64)
        if gr.ndim != 2 or gr.shape != grw.shape:
            raise ValueError("grid and weights must both be (nv, n_gauss)")
        ls = np.ascontiguousarray(req.ls, dtype=np.int32)
        key = 0
        if state is not None:
            with self._mu:
                key = self._states.get(state)
                if key is None:
                    key = self._states[state] = self._next_key
                    self._next_key += 1
        g = (req.gr, req.grw, req.ls, gr, grw, ls, _addr(gr), _addr(grw), _addr(ls),
             gr.shape[0], gr.shape[1], ls.size, key)
        if state is not None:
            self._geo[state] = g
        return g

    def submit(self, req):
        """Queues a ``SlaterRequest``; returns its ticket."""
        if req.kind != "slater":
            raise ValueError("the native server takes all-pairs Slater-integral requests")
        prec = PRECISIONS.get(req.precision)
        if prec is None:
            raise ValueError(f"precision must be one of {tuple(PRECISIONS)}")
        (_, _, _, gr, grw, ls, gr_p, grw_p, ls_p, nv, ng, n_orb, key) = self._geometry(req)
        P, Q = req.P, req.Q
        if type(P) is not np.ndarray or P.dtype != np.float64 or not P.flags.c_contiguous:
            P = np.ascontiguousarray(P, dtype=np.float64)
        if Q is not None and (type(Q) is not np.ndarray or Q.dtype != np.float64
                              or not Q.flags.c_contiguous):
            Q = np.ascontiguousarray(Q, dtype=np.float64)
        shape = (n_orb, nv * ng)
        if P.shape != shape or (Q is not None and Q.shape != shape):
            raise ValueError(f"orbital samples must be {shape}")
        t = ctypes.c_int64()
        rc = self._submit(self._h, gr_p, grw_p, nv, ng, n_orb, ls_p, int(req.kmax), prec,
                          _addr(P), None if Q is None else _addr(Q), key, ctypes.byref(t))
        if rc:
            check(rc, "ax_server_submit")
        with self._mu:
            self._keep[t.value] = (req, (gr, grw, ls, P, Q))
        return t.value

    def done(self, ticket):
        d = ctypes.c_int32()
        call("ax_server_poll", self._h, ctypes.c_int64(ticket), ctypes.byref(d))
        return bool(d.value)

    def result(self, ticket):
        """Waits for a ticket and returns its ``SlaterResult`` (the request's
        error is raised here)."""
        lib = self._lib
        rc = lib.ax_server_wait(self._h, ticket)
        if rc:
            check(rc, "ax_server_wait")
        with self._mu:
            req, _ = self._keep.pop(ticket)
        ptr, size, slices = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int32()
        rc = lib.ax_server_result(self._h, ticket, ctypes.byref(ptr), ctypes.byref(size),
                                  ctypes.byref(slices))
        if rc:
            try:
                check(rc, "ax_server_result")
            finally:
                lib.ax_server_release(self._h, ticket)
        buf = (ctypes.c_double * size.value).from_address(ptr.value)
        buf._hold = _Release(self, ticket)  # lives as long as any view of the values
        values = np.frombuffer(buf, dtype=np.float64)
        lk = (np.asarray(req.ls, dtype=np.int32).tobytes(), int(req.kmax))
        layout = self._layouts.get(lk)
        if layout is None:
            layout = self._layouts[lk] = _rank_layout(req.ls, req.kmax)
            if len(self._layouts) > 4096:
                self._layouts.pop(next(iter(self._layouts)))
        m, offs = layout
        return SlaterResult(values, offs, m, slices=int(slices.value))

    _STATS = ("launches", "plans", "peak_bytes", "setup_ms", "pack_wait_ms", "alloc_ms",
              "queue_ms", "plans_made", "resident_states", "resident_bytes")

    def stats(self):
        """Counters since creation: launch sets, plans run, peak device bytes
        (in flight plus resident), milliseconds of the scheduler thread in
        plan set-up, in waiting for the packing, in allocations and in
        queueing copies and launches, and plans made; and now: resident
        states holding a plan and their bytes."""
        v = (ctypes.c_int64 * len(self._STATS))()
        call("ax_server_stats", self._h, v, len(self._STATS))
        return {k: (x / 1e6 if k.endswith("_ms") else x) for k, x in zip(self._STATS, v)}

    def drop_state(self, state):
        """Frees a resident state's plan once none of its requests is queued
        or in flight; a later request naming it starts a new state."""
        with self._mu:
            key = self._states.pop(state, None)
            self._geo.pop(state, None)
        if key is not None:
            call("ax_server_drop_state", self._h, ctypes.c_int64(key))

    def close(self):
        """Finishes every queued and in-flight request and frees the server;
        values of results still held stay valid (their pinned buffers are
        then left to the process)."""
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ax_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
