"""Request server for Slater-integral and Fock-build requests on one GPU.

Continuous batching.  A scheduler thread owns the device.  Requests wait in a
FIFO queue; the head is admitted as soon as an in-flight slot is free and its
HBM footprint (estimated before anything is allocated) fits the budget beside
everything in flight and resident.  Admitted Slater-integral requests of one
precision mode are launched together (one batched launch set for up to 16
plans, ``ax_rk_tensor_run_batch``) on one of a few streams, so one batch runs
while the next is admitted and launched; each Fock-build request takes a
stream slot of its own.  Results return through futures; latency is measured
from ``submit`` to the moment the result is ready on the host.

Residency.  A request that names a ``state`` keeps its plan (grid tables, pair
densities, Y^k workspaces) and its input and output buffers on the device
after it completes, so the next request of that state (the next SCF
iteration of the same atom) uploads only its orbitals.  Idle resident states
are evicted least recently used when admission needs their memory.

Determinism.  A request's plan, tiles, K splits and summation order do not
depend on what else is in flight, so its result is bitwise the same alone and
in any mix.

The scheduler is independent of the device: it drives an executor object
(``GpuExecutor`` here) through footprint / reserve / launch / done / finish /
evict calls, which is what the CPU tests exercise with a stand-in executor.
"""

from __future__ import annotations

import collections
import itertools
import os
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Any, Hashable, Optional, Sequence

import numpy as np

__all__ = ["SlaterRequest", "SlaterBatchRequest", "FockRequest", "SlaterResult", "FockResult",
           "Scheduler", "GpuExecutor", "SlaterServer", "percentile"]


# ---------------------------------------------------------------------------
# Requests and results
# ---------------------------------------------------------------------------
@dataclass
class SlaterRequest:
    """All R^k(ab;cd), k = 0..kmax, over the valid pair-pairs of the orbitals
    sampled on a Gauss grid (the tensor of ``atomx_amd.tensor``)."""
    gr: np.ndarray                 # (nv, ng) nodes
    grw: np.ndarray                # (nv, ng) weights
    P: np.ndarray                  # (n_orb, nv*ng) large-component samples
    Q: Optional[np.ndarray]        # small-component samples, or None
    ls: Sequence[int]
    kmax: int
    precision: str = "split"
    state: Optional[Hashable] = None

    kind = "slater"


@dataclass
class SlaterBatchRequest:
    """All R^k of several atoms in one request (one SCF iteration of a batch
    of atoms or ions, say); every atom is a ``SlaterRequest`` with its own
    grid, orbitals and resident state.  The atoms are scheduled together and
    the result is the list of their ``SlaterResult``."""
    atoms: Sequence[SlaterRequest]

    kind = "slater_batch"

    @property
    def precision(self):
        return self.atoms[0].precision if self.atoms else "fp64"


@dataclass
class FockRequest:
    """Y^k-form Fock pieces of every target of one atom (``FockYk.build``)."""
    bsp_data: dict
    P: np.ndarray                  # (n_orb, ns) coefficients
    Q: np.ndarray
    occ: Sequence[float]
    exchange_coef: np.ndarray      # (n_orb, n_orb, K): target, source, rank
    targets: Optional[Sequence[int]] = None
    state: Optional[Hashable] = None

    kind = "fock"


@dataclass
class SlaterResult:
    values: np.ndarray             # packed upper triangles, rank by rank
    offsets: list
    m: list
    latency_s: float = 0.0
    # int8 slices per operand the contraction used (0: it ran in FP64, also
    # when split precision was asked for and the grid needs FP64)
    slices: int = 0

    def matrix(self, k):
        m = self.m[k]
        vals = self.values[self.offsets[k]:self.offsets[k] + m * (m + 1) // 2]
        M = np.zeros((m, m))
        i, j = np.triu_indices(m)
        M[i, j] = vals
        M[j, i] = vals
        return M


@dataclass
class FockResult:
    blocks: dict                   # VJ_PP, VJ_QQ, K_PP, K_QQ: (n_targets, ns, ns)
    latency_s: float = 0.0


def percentile(values, q):
    """Nearest-rank percentile (q in [0, 100]) of a non-empty sequence."""
    v = np.sort(np.asarray(values, dtype=float))
    if v.size == 0:
        return float("nan")
    rank = int(np.ceil(q / 100.0 * v.size))
    return float(v[min(max(rank, 1), v.size) - 1])


# ---------------------------------------------------------------------------
# Scheduler (device independent)
# ---------------------------------------------------------------------------
@dataclass
class _Ticket:
    req: Any
    future: Future
    t_submit: float
    seq: int
    nbytes: int = 0


@dataclass
class _Job:
    handle: Any
    tickets: list
    nbytes: int
    t_launch: float = field(default_factory=time.perf_counter)


class Scheduler:
    """FIFO admission by footprint, batched launches, in-flight slots.

    executor interface:
      ready(req) -> bool         False while the request's resident state is in flight
      footprint(req) -> int      device bytes the request holds while in flight
      reserve(req)               pin the request's resident state (if any)
      resident_bytes() -> int    bytes held outside in-flight footprints (idle states, caches)
      evict(nbytes) -> int       drop idle resident states (LRU) and caches, bytes freed
      batch_key(req)             requests with equal keys may share a launch
      batch_weight(req)          (optional) work units toward max_batch, default 1
      launch(reqs) -> handle     start the work of a batch (asynchronous)
      done(handle) -> bool
      finish(handle) -> list     results (or exceptions) in request order
    """

    def __init__(self, executor, budget_bytes, max_inflight=4, max_batch=16, poll_s=2e-4):
        self.ex = executor
        self.budget = int(budget_bytes)
        self.max_inflight = int(max_inflight)
        self.max_batch = int(max_batch)
        self.poll_s = float(poll_s)
        self._cv = threading.Condition()
        self._arrivals = collections.deque()
        self._pending = collections.deque()
        self._inflight = []
        self._inflight_bytes = 0
        self._seq = itertools.count()
        self._stop = False
        self.latencies = []
        self.completed = 0
        self.failed = 0
        self.peak_bytes = 0
        self.peak_inflight = 0
        self.launches = 0
        self.error = None
        self._thread = threading.Thread(target=self._run, name="slater-scheduler", daemon=True)
        self._thread.start()

    # -- client side -----------------------------------------------------------
    def submit(self, req) -> Future:
        fut = Future()
        t = _Ticket(req, fut, time.perf_counter(), next(self._seq))
        with self._cv:
            if self.error is not None:
                raise RuntimeError(f"scheduler stopped after an error: {self.error!r}")
            if self._stop:
                raise RuntimeError("scheduler is closed")
            self._arrivals.append(t)
            self._cv.notify()
        return fut

    def close(self, wait=True):
        with self._cv:
            self._stop = True
            self._cv.notify()
        if wait:
            self._thread.join()

    def reset_stats(self):
        self.latencies = []
        self.completed = self.failed = self.launches = 0
        self.peak_bytes = self.peak_inflight = 0

    def stats(self):
        lat = self.latencies
        return {"completed": self.completed, "failed": self.failed,
                "p50_ms": percentile(lat, 50) * 1e3 if lat else None,
                "p99_ms": percentile(lat, 99) * 1e3 if lat else None,
                "peak_bytes": self.peak_bytes, "peak_inflight": self.peak_inflight,
                "launches": self.launches, "budget_bytes": self.budget}

    def _weight(self, req):
        """Work units of a request toward max_batch (atoms of a batch request)."""
        fn = getattr(self.ex, "batch_weight", None)
        return fn(req) if fn is not None else 1

    # -- scheduler thread --------------------------------------------------------
    def _run(self):
        try:
            while True:
                with self._cv:
                    while self._arrivals:
                        self._pending.append(self._arrivals.popleft())
                    if self._stop and not self._pending and not self._inflight:
                        return
                progressed = self._complete()
                progressed |= self._admit_and_launch()
                if not progressed:
                    with self._cv:
                        if not self._arrivals:
                            busy = bool(self._inflight or self._pending)
                            if busy or not self._stop:
                                self._cv.wait(self.poll_s if busy else 0.05)
        except BaseException as exc:  # fail every outstanding request loudly
            with self._cv:
                # later submits raise instead of queueing work nobody runs
                self.error = exc
                self._stop = True
                rest = list(self._arrivals) + list(self._pending)
                self._arrivals.clear()
                self._pending.clear()
            for job in self._inflight:
                rest.extend(job.tickets)
            self._inflight = []
            for t in rest:
                if not t.future.done():
                    t.future.set_exception(exc)
            raise

    def _complete(self):
        progressed = False
        still = []
        for job in self._inflight:
            if not self.ex.done(job.handle):
                still.append(job)
                continue
            progressed = True
            results = self.ex.finish(job.handle)
            now = time.perf_counter()
            self._inflight_bytes -= job.nbytes
            for t, res in zip(job.tickets, results):
                if isinstance(res, BaseException):
                    self.failed += 1
                    t.future.set_exception(res)
                    continue
                lat = now - t.t_submit
                if hasattr(res, "latency_s"):
                    res.latency_s = lat
                self.latencies.append(lat)
                self.completed += 1
                t.future.set_result(res)
        self._inflight = still
        return progressed

    def _admit_and_launch(self):
        if len(self._inflight) >= self.max_inflight:
            return False
        admitted = []
        nbytes = 0
        weight = 0
        while self._pending:
            t = self._pending[0]
            if not self.ex.ready(t.req):
                break  # its resident state is still in flight (FIFO: no overtaking)
            need = self.ex.footprint(t.req)
            if need > self.budget:
                self._pending.popleft()
                self.failed += 1
                t.future.set_exception(MemoryError(
                    f"request needs {need} device bytes, budget is {self.budget}"))
                continue
            used = self._inflight_bytes + nbytes + self.ex.resident_bytes()
            if used + need > self.budget:
                self.ex.evict(used + need - self.budget)
                used = self._inflight_bytes + nbytes + self.ex.resident_bytes()
                if used + need > self.budget:
                    break  # wait for in-flight work to free memory (FIFO: no overtaking)
            w = self._weight(t.req)
            if admitted and (weight + w > self.max_batch
                             or self.ex.batch_key(t.req) != self.ex.batch_key(admitted[0].req)):
                break
            weight += w
            self._pending.popleft()
            self.ex.reserve(t.req)
            t.nbytes = need
            nbytes += need
            admitted.append(t)
        if not admitted:
            return False
        try:
            handle = self.ex.launch([t.req for t in admitted])
        except BaseException as exc:  # the admitted tickets are no longer queued
            for t in admitted:
                if not t.future.done():
                    t.future.set_exception(exc)
            raise
        self._inflight.append(_Job(handle, admitted, nbytes))
        self._inflight_bytes += nbytes
        self.launches += 1
        self.peak_inflight = max(self.peak_inflight, len(self._inflight))
        self.peak_bytes = max(self.peak_bytes, self._inflight_bytes + self.ex.resident_bytes())
        return True


# ---------------------------------------------------------------------------
# GPU executor
# ---------------------------------------------------------------------------
def _grid_key(gr, grw):
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    for a in (gr, grw):
        a = np.ascontiguousarray(a, dtype=np.float64)
        h.update(np.asarray(a.shape, dtype=np.int64).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()


class _State:
    """Device objects of one resident state (or of one transient request)."""

    def __init__(self, key, shape_key, nbytes):
        self.key = key
        self.shape_key = shape_key
        self.nbytes = nbytes
        self.busy = False
        self.plan = None
        self.fock = None
        self.P = self.Q = self.out = None
        self.host_out = None
        # the grid, grid-weight and l arrays and the scalars of the request
        # that last matched this state: a request passing the same objects
        # is not checked again (the arrays are kept, so their ids stay theirs)
        self.match = None

    def close(self):
        if self.plan is not None:
            self.plan.close()
        self.plan = self.fock = None
        self.P = self.Q = self.out = self.host_out = None


class GpuExecutor:
    """Runs Slater-integral and Fock-build requests on the current GPU."""

    def __init__(self, n_streams=4):
        import torch

        from . import device as dv
        from ._lib import PRECISIONS, load
        dv.require_gpu()
        self._torch = torch
        self._dv = dv
        self._lib = load()
        self._prec = PRECISIONS
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self._free_streams = list(range(n_streams))
        # pinned staging per in-flight slot: the uploads of the slot's launch
        # set (reused once the slot's job has finished)
        self._stage = [None] * n_streams
        # device staging bytes charged to the launch set in flight on each slot
        self._slot_need = [0] * n_streams
        # host copies into the staging, split across threads (numpy copies
        # run without the interpreter lock); the workers use this device
        try:
            cpus = len(os.sched_getaffinity(0))
        except AttributeError:  # pragma: no cover
            cpus = os.cpu_count() or 1
        self._pack_threads = max(1, min(8, cpus // 2))
        device = torch.cuda.current_device()
        self._pool = (ThreadPoolExecutor(self._pack_threads,
                                         initializer=lambda: torch.cuda.set_device(device))
                      if self._pack_threads > 1 else None)
        self.states = collections.OrderedDict()  # resident, LRU order
        self._idle_bytes = 0                     # bytes of idle resident states
        self._estimates = {}                     # request shape -> (bytes, R^k count)
        self.plans_created = 0
        self._fock_cache = {}
        self._groups = collections.OrderedDict()  # batch layouts by their states' keys

    # -- footprints ----------------------------------------------------------------
    def _info(self, req):
        """(device bytes of a fresh state's plan, R^k count, device staging
        bytes of the request's inputs and outputs), computed once per request
        (the estimate once per request shape).  The state's shape key, which
        hashes the grid, is left to state creation (``_new_state``): a
        request of an existing state is matched by the identity of its
        arrays."""
        info = getattr(req, "_ax_info", None)
        if info is None:
            if req.kind == "slater":
                nv, ng = np.shape(req.gr)
                sk = (nv, ng, tuple(int(x) for x in req.ls), int(req.kmax), req.precision,
                      req.Q is not None)
                est = self._estimates.get(sk)
                if est is None:
                    n_rk = _count_rk(req.ls, int(req.kmax))
                    n_in = len(req.ls) * nv * ng * (2 if req.Q is not None else 1)
                    est = (self._slater_bytes(req), n_rk, 8 * (n_in + n_rk))
                    self._estimates[sk] = est
                    if len(self._estimates) > 4096:
                        self._estimates.pop(next(iter(self._estimates)))
                info = est
            else:
                info = (self._fock_bytes(req), 0, 0)
            req._ax_info = info
        return info

    def _slater_shape(self, req):
        """Identity of a resident state's plan (None for a one-off request)."""
        if req.state is None:
            return None
        nv, ng = np.shape(req.gr)
        ls = tuple(int(x) for x in req.ls)
        return ("slater", _grid_key(req.gr, req.grw), nv, ng, ls, int(req.kmax), req.precision,
                req.Q is not None)

    def _slater_bytes(self, req):
        """Device bytes of the request's plan (what a resident state holds);
        its inputs and outputs live in the slot staging, charged per launch."""
        import ctypes
        nv, ng = np.shape(req.gr)
        ls = np.ascontiguousarray(req.ls, dtype=np.int32)
        plan = int(self._lib.ax_rk_tensor_estimate(
            nv, ng, ls.size, ls.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), int(req.kmax),
            self._prec[req.precision]))
        if plan < 0:
            raise ValueError("invalid Slater-integral request")
        return plan

    def _fock_bytes(self, req):
        from .fock_yk import FockYk
        n_orb = np.shape(req.P)[0]
        nt = n_orb if req.targets is None else len(req.targets)
        bd = req.bsp_data
        ns, nv, ng = int(bd["ns"]), int(bd["nv"]), np.shape(bd["gr"])[1]
        return FockYk.footprint(n_orb, nt, ns, nv * ng, nv, ng, int(bd["ks"]))

    def _key(self, req):
        return (req.kind, req.state) if req.state is not None else None

    def footprint(self, req):
        """Device bytes a request holds while in flight: its plans (held by
        the resident state when there is one) and its staged inputs and
        outputs."""
        if req.kind == "slater_batch":
            states, total = self.states, 0
            for a in req.atoms:
                st = states.get((a.kind, a.state)) if a.state is not None else None
                info = self._info(a)
                total += (st.nbytes if st is not None else info[0]) + info[2]
            return total
        info = self._info(req)
        key = self._key(req)
        if key is not None and key in self.states:
            return self.states[key].nbytes + info[2]
        return info[0] + info[2]

    def _staging_need(self, reqs):
        """Device staging bytes a launch set's requests were charged."""
        total = 0
        for r in reqs:
            for a in (r.atoms if r.kind == "slater_batch" else (r,)):
                total += self._info(a)[2]
        return total

    def ready(self, req):
        if req.kind == "slater_batch":
            states = self.states
            for a in req.atoms:
                if a.state is not None:
                    st = states.get((a.kind, a.state))
                    if st is not None and st.busy:
                        return False
            return True
        key = self._key(req)
        return key is None or key not in self.states or not self.states[key].busy

    def _set_busy(self, st, busy):
        """Marks a resident state busy/idle, keeping the idle-bytes total."""
        if st.busy == busy:
            return
        st.busy = busy
        if st.key is not None and self.states.get(st.key) is st:
            self._idle_bytes += -st.nbytes if busy else st.nbytes

    def reserve(self, req):
        if req.kind == "slater_batch":
            states = self.states
            for a in req.atoms:
                st = states.get((a.kind, a.state)) if a.state is not None else None
                if st is not None and not st.busy:  # the common case, inline
                    st.busy = True
                    if st.key is not None:
                        self._idle_bytes -= st.nbytes
                    states.move_to_end(st.key)
                else:
                    self.reserve(a)
            return
        key = self._key(req)
        if key is None:
            return
        st = self.states.get(key)
        if st is None:  # placeholder until the launch builds it
            st = self._new_state(req, key)
            self.states[key] = st
        else:
            self._set_busy(st, True)
        self.states.move_to_end(key)

    def _held_staging(self):
        """Device staging bytes held by the slots beyond what their in-flight
        launch sets were charged (all of it for a free slot)."""
        extra = 0
        for sid, buf in enumerate(self._stage):
            if buf is not None:
                extra += max(0, 8 * buf[1].numel() - self._slot_need[sid])
        return extra

    def resident_bytes(self):
        """Device bytes held outside the in-flight footprints: idle resident
        states and the slots' device staging."""
        return self._idle_bytes + self._held_staging()

    def evict(self, nbytes):
        """Frees at least nbytes if it can: idle resident states (least
        recently used first), then the device staging of free slots."""
        freed = 0
        for key in list(self.states):
            if freed >= nbytes:
                break
            st = self.states[key]
            if st.busy:
                continue
            freed += st.nbytes
            self._idle_bytes -= st.nbytes
            st.close()
            del self.states[key]
        for sid in list(self._free_streams):
            if freed >= nbytes:
                break
            buf = self._stage[sid]
            if buf is not None:
                freed += 8 * buf[1].numel()
                self._stage[sid] = None
        return freed

    def batch_weight(self, req):
        return len(req.atoms) if req.kind == "slater_batch" else 1

    def batch_key(self, req):
        if req.kind in ("slater", "slater_batch"):
            return ("slater", req.precision)
        return ("fock", id(req))

    # -- launches --------------------------------------------------------------------
    @staticmethod
    def _match(req):
        return (req.gr, req.grw, req.ls, req.kmax, req.precision, req.Q is not None)

    @staticmethod
    def _same(match, req):
        return (match is not None and match[0] is req.gr and match[1] is req.grw and
                match[2] is req.ls and match[3] == req.kmax and match[4] == req.precision and
                match[5] == (req.Q is not None))

    def _state_for(self, req):
        key = self._key(req)
        if key is not None and key in self.states:
            st = self.states[key]
            if st.shape_key is not None and req.kind == "slater" and not self._same(st.match, req):
                # hashed and compared again only when other objects are passed
                if self._slater_shape(req) != st.shape_key:
                    raise ValueError(f"state {req.state!r} was created with another shape")
                st.match = self._match(req)
            self._set_busy(st, True)
            return st
        st = self._new_state(req, key)
        if key is not None:
            self.states[key] = st
        return st

    def _new_state(self, req, key):
        """A busy state for ``req``, with its shape key (the grid hashed
        once per state) and the plan bytes of its estimate."""
        slater = req.kind == "slater"
        st = _State(key, self._slater_shape(req) if slater else None, self._info(req)[0])
        st.busy = True
        if slater:
            st.match = self._match(req)
        return st

    def launch(self, reqs):
        torch = self._torch
        sid = self._free_streams.pop(0)
        stream = self.streams[sid]
        job = {"sid": sid, "reqs": reqs, "states": [], "errors": [None] * len(reqs),
               "keep": [], "event": torch.cuda.Event()}
        self._slot_need[sid] = self._staging_need(reqs)
        with torch.cuda.stream(stream):
            if reqs[0].kind in ("slater", "slater_batch"):
                self._launch_slater(reqs, job, stream)
            else:
                self._launch_fock(reqs[0], job, stream)
            job["event"].record(stream)
        return job

    def _staging(self, sid, n_in, n):
        """Pinned host staging of at least n_in input values and device
        staging of at least n values (inputs, then outputs) for slot sid,
        grown as needed and reused once the slot's job is finished."""
        torch, dv = self._torch, self._dv
        buf = self._stage[sid]
        if buf is None or buf[0].numel() < n_in or buf[1].numel() < n:
            host = buf[0] if buf is not None and buf[0].numel() >= n_in else None
            if host is None:
                n_in = max(n_in, 1 << 16, buf[0].numel() if buf is not None else 0)
                host = torch.empty(n_in, dtype=torch.float64, pin_memory=True)
            # the device staging grows to exactly what is asked (admission
            # charged it), never past the largest launch set it served
            dev = buf[1] if buf is not None and buf[1].numel() >= n else None
            if dev is None:
                dev = dv.empty((max(n, buf[1].numel() if buf is not None else 0),))
            buf = (host, dev)
            self._stage[sid] = buf
        return buf

    _PACK_SERIAL = 1 << 19  # values below which one thread packs

    def _pack(self, chunks, dst):
        """Concatenates the 1-D arrays ``chunks`` into ``dst``; large inputs
        are cut into equal ranges of the concatenation, one per thread."""
        self._pack_wait(self._pack_start(chunks, dst))

    def _pack_start(self, chunks, dst):
        """Starts ``_pack``: small inputs are copied at once, large ones by
        the worker threads; returns what ``_pack_wait`` waits for."""
        n = dst.size
        if self._pool is None or n < self._PACK_SERIAL:
            np.concatenate(chunks, out=dst)
            return []
        starts = np.concatenate([[0], np.cumsum([c.size for c in chunks])])
        if int(starts[-1]) != n:
            raise ValueError("packed inputs do not fill the staging range")
        parts = self._pack_threads
        bounds = [n * w // parts for w in range(parts + 1)]

        def copy_range(lo, hi):
            c = int(np.searchsorted(starts, lo, side="right")) - 1
            while lo < hi:
                c0 = int(starts[c])
                take = min(hi, int(starts[c + 1])) - lo
                dst[lo:lo + take] = chunks[c][lo - c0:lo - c0 + take]
                lo += take
                c += 1

        return [self._pool.submit(copy_range, bounds[w], bounds[w + 1]) for w in range(parts)
                if bounds[w] < bounds[w + 1]]

    def _pack_wait(self, futs):
        for f in futs:
            f.result()

    _OWN_RESULTS = 1 << 17  # result values from which a launch set gets its own buffer
    # R^k count from which a new resident plan's contraction schedule is tuned
    # once, on its first request (a few milliseconds, then every later request)
    _TUNE_MIN_RK = 50_000

    def _host_need(self, n_in, end):
        """Pinned staging values a launch set needs: its inputs, and its
        results too when they come back through the staging."""
        return n_in if end - n_in >= self._OWN_RESULTS else end

    def _results_out(self, job, host, dev, n_in, end):
        """One device-to-host copy of a launch set's packed results on the
        current stream.  Large results go to a pinned buffer of the set's
        own (torch's caching host allocator) whose views become the results,
        so they are never copied again on the host; small ones go through
        the slot's staging and are copied out once when the set finishes."""
        torch = self._torch
        if end - n_in >= self._OWN_RESULTS:
            out = torch.empty(end - n_in, dtype=torch.float64, pin_memory=True)
            job["own"] = True
        else:
            out = host[n_in:end]
            job["own"] = False
        out.copy_(dev[n_in:end], non_blocking=True)
        job["host"] = out

    def _group(self, atoms):
        """Cached layout of a launch set whose atoms (those of one multi-atom
        request, or single requests) all have built resident states (the
        same group again, next SCF iteration), or None: (states, plans, plan
        handles, input offsets, output sizes, output offsets, Q present)."""
        try:
            gkey = tuple([(a.kind, a.state) for a in atoms])
        except TypeError:
            return None
        if any(k[1] is None for k in gkey):
            return None
        g = self._groups.get(gkey)
        if g is None:
            sts = [self.states.get(k) for k in gkey]
            if any(st is None or st.plan is None for st in sts):
                return None
            import ctypes
            plans = [st.plan for st in sts]
            sizes = []
            for a, pl in zip(atoms, plans):
                sizes.append(pl.n_orb * pl.n)
                if a.Q is not None:
                    sizes.append(pl.n_orb * pl.n)
            offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
            outs = np.array([pl.size for pl in plans], dtype=np.int64)
            ooff = offs[-1] + np.concatenate([[0], np.cumsum(outs)]).astype(np.int64)
            hasq = [a.Q is not None for a in atoms]
            handles = (ctypes.c_void_p * len(plans))(*[pl._h.value for pl in plans])
            g = (sts, plans, handles, offs, outs, ooff, hasq)
            self._groups[gkey] = g
            if len(self._groups) > 256:
                self._groups.popitem(last=False)
        sts, plans, _, _, _, _, hasq = g
        for k, st, pl in zip(gkey, sts, plans):
            if self.states.get(k) is not st or st.plan is not pl:  # a state was rebuilt
                del self._groups[gkey]
                return None
        # the same objects and shapes as when the layout was made
        f64 = np.float64
        for a, st, pl, q in zip(atoms, sts, plans, hasq):
            if (not self._same(st.match, a) or type(a.P) is not np.ndarray
                    or a.P.dtype != f64 or a.P.shape != (pl.n_orb, pl.n) or (a.Q is not None) != q
                    or (q and (type(a.Q) is not np.ndarray or a.Q.dtype != f64
                               or a.Q.shape != a.P.shape))):
                return None
        return g

    def _launch_group(self, atoms, g, job, stream):
        """A cached layout: one concatenation into the staging, one upload,
        one batch launch from prebuilt pointer tables, one download."""
        from .tensor import run_batch_arrays
        sts, plans, handles, offs, outs, ooff, hasq = g
        for st in sts:
            self._set_busy(st, True)
        chunks = []
        for a, q in zip(atoms, hasq):
            chunks.append(a.P.reshape(-1))
            if q:
                chunks.append(a.Q.reshape(-1))
        n_in = int(offs[-1])
        end = int(ooff[-1])
        host, dev = self._staging(job["sid"], self._host_need(n_in, end), end)
        hv = host.numpy()
        self._pack(chunks, hv[:n_in])
        base = np.uint64(dev.data_ptr())
        addr = base + np.uint64(8) * offs[:-1].astype(np.uint64)
        n = len(plans)
        pp = np.empty(n, dtype=np.uint64)
        qq = np.zeros(n, dtype=np.uint64)
        k = 0
        for i, q in enumerate(hasq):
            pp[i] = addr[k]
            k += 1
            if q:
                qq[i] = addr[k]
                k += 1
        oo = base + np.uint64(8) * ooff[:-1].astype(np.uint64)
        dev[:n_in].copy_(host[:n_in], non_blocking=True)
        run_batch_arrays(handles, n, pp, qq, oo, stream=stream)
        job["group"] = (n_in, end, ooff, outs)
        self._results_out(job, host, dev, n_in, end)
        if job["reqs"][0].kind == "slater_batch":
            job["states"].append(list(sts))
        else:
            job["states"].extend([st] for st in sts)

    def _launch_slater(self, reqs, job, stream):
        """One launch set: the inputs of every request are packed into the
        slot's pinned staging and moved with one copy, the plans run as one
        batch on raw device offsets of the slot's device staging, and all
        results come back with one copy."""
        from .tensor import SlaterTensor, run_batch_raw
        if len(reqs) == 1 and reqs[0].kind == "slater_batch":
            atoms = reqs[0].atoms
        elif all(r.kind == "slater" for r in reqs):
            atoms = reqs
        else:
            atoms = None
        if atoms:
            g = self._group(atoms)
            if g is not None:
                self._launch_group(atoms, g, job, stream)
                return
        entries = []  # [request index, atom index or -1, atom, state, P, Q]
        f64 = np.float64
        # the states of every request first (a request whose state cannot be
        # had fails alone), then the shape checks request by request against
        # the request itself (plans may not exist yet)
        claimed = []
        for req in reqs:
            atoms = req.atoms if req.kind == "slater_batch" else (req,)
            sts, err = [], None
            try:
                for a in atoms:
                    sts.append(self._state_for(a))
            except Exception as exc:
                err = exc
            claimed.append((atoms, sts, err))
        for i, (req, (atoms, sts, err)) in enumerate(zip(reqs, claimed)):
            batch = req.kind == "slater_batch"
            got = []
            try:
                if err is not None:
                    raise err
                for j, (a, st) in enumerate(zip(atoms, sts)):
                    P, Q = a.P, a.Q
                    if type(P) is not np.ndarray or P.dtype != f64:
                        P = np.asarray(P, dtype=f64)
                    if Q is not None and (type(Q) is not np.ndarray or Q.dtype != f64):
                        Q = np.asarray(Q, dtype=f64)
                    shape = (len(a.ls), int(np.size(a.gr)))
                    if P.shape != shape or (Q is not None and Q.shape != shape):
                        raise ValueError(f"orbital samples must be {shape}")
                    got.append([i, j if batch else -1, a, st, P, Q])
                entries.extend(got)
            except Exception as exc:  # the request fails alone
                self._fail_request(job, i, exc, atoms, sts)
            job["states"].append(sts)
        if not entries:
            return
        # staging layout from the requests (a plan's size is its R^k count),
        # the inputs packed on the worker threads while the new plans are made
        chunks, in_off = [], []
        off = 0
        for e in entries:
            in_off.append(off)
            chunks.append(e[4].reshape(-1))
            off += e[4].size
            if e[5] is not None:
                chunks.append(e[5].reshape(-1))
                off += e[5].size
        n_in = off
        out_off, out = [], n_in
        for e in entries:
            out_off.append(out)
            out += e[3].plan.size if e[3].plan is not None else self._info(e[2])[1]
        host, dev = self._staging(job["sid"], self._host_need(n_in, out), out)
        hv = host.numpy()
        packing = self._pack_start(chunks, hv[:n_in])
        tune, failed = self._make_plans(job, entries, stream)
        self._pack_wait(packing)
        base = dev.data_ptr()
        pp, qq, oo, plans = [], [], [], []
        results = {}
        for e, i0, o0 in zip(entries, in_off, out_off):
            i, j, a, st = e[0], e[1], e[2], e[3]
            if i in failed:
                continue
            plan = st.plan
            if o0 + plan.size > out:
                raise RuntimeError("R^k plan size differs from its request's count")
            pp.append(base + 8 * i0)
            qq.append(base + 8 * (i0 + e[4].size) if e[5] is not None else None)
            oo.append(base + 8 * o0)
            results[(i, j)] = (o0, plan.size)
            plans.append(plan)
        if not plans:
            return
        dev[:n_in].copy_(host[:n_in], non_blocking=True)
        for t, st in tune:  # times every contraction schedule on the staged inputs
            if t not in plans:
                continue
            k = plans.index(t)
            try:
                t.tune_raw(pp[k], qq[k], oo[k], stream=stream)
            except (RuntimeError, ValueError, MemoryError):  # tuning is optional: the
                pass  # plan keeps the schedule it had (ax_rk_tensor_tune restores it)
            # the other K splits' job tables (and a larger partial workspace)
            # now belong to the resident state
            st.nbytes = max(st.nbytes, t.footprint)
        run_batch_raw(plans, pp, qq, oo, stream=stream)
        job["results"] = results
        job["n_in"] = n_in
        self._results_out(job, host, dev, n_in, out)

    def _fail_request(self, job, i, exc, atoms, sts):
        """Request i of the launch set fails alone: every state the scheduler
        reserved for it is released with it, also those of atoms after the
        failing one."""
        job["errors"][i] = exc
        seen = {id(st) for st in sts}
        for a in atoms:
            k = self._key(a)
            st = self.states.get(k) if k is not None else None
            if st is not None and id(st) not in seen:
                sts.append(st)
                seen.add(id(st))

    def _make_plans(self, job, entries, stream):
        """Plans for the entries whose state has none: all of one precision in
        one library call (one staging upload, one shared set-up event, queued
        on the launch set's own stream); when that fails, one by one, so that
        a bad request fails alone.  Returns (new resident plans to tune,
        indices of the requests that failed)."""
        from .tensor import SlaterTensor
        tune, failed = [], set()
        need = {}
        for e in entries:
            if e[3].plan is None:
                need.setdefault(e[2].precision, []).append(e)
        for prec, es in need.items():
            try:
                made = SlaterTensor.create_many([(e[2].gr, e[2].grw, e[2].ls, e[2].kmax)
                                                 for e in es], prec, stream=stream)
            except (RuntimeError, ValueError, MemoryError):
                made = None
            for k, e in enumerate(es):
                i, st = e[0], e[3]
                if i in failed:
                    if made is not None:
                        made[k].close()
                    continue
                try:
                    plan = made[k] if made is not None else SlaterTensor(
                        e[2].gr, e[2].grw, e[2].ls, e[2].kmax, prec)
                except Exception as exc:
                    failed.add(i)
                    req = job["reqs"][i]
                    atoms = req.atoms if req.kind == "slater_batch" else (req,)
                    self._fail_request(job, i, exc, atoms, [])
                    continue
                st.plan = plan
                self.plans_created += 1
                # a resident state of a large tensor is tuned on its first request
                if st.key is not None and plan.size >= self._TUNE_MIN_RK:
                    tune.append((plan, st))
        return tune, failed

    def _launch_fock(self, req, job, stream):
        from .fock_yk import FockYk
        torch = self._torch
        st = None
        try:
            st = self._state_for(req)
            if st.fock is None:
                key = _grid_key(req.bsp_data["gr"], req.bsp_data["grw"]) + str(
                    req.bsp_data["bsp_p"].shape)
                fk = self._fock_cache.get(key)
                if fk is None:
                    fk = FockYk(req.bsp_data)
                    self._fock_cache[key] = fk
                    if len(self._fock_cache) > 8:
                        self._fock_cache.pop(next(iter(self._fock_cache)))
                st.fock = fk
            blocks = st.fock.build(req.P, req.Q, req.occ, req.exchange_coef, req.targets)
            host = {}
            for name, t in blocks.items():
                h = torch.empty(t.shape, dtype=torch.float64, pin_memory=True)
                h.copy_(t, non_blocking=True)
                host[name] = h
            job["keep"].append(blocks)
            job["host"] = host
            job["states"].append([st])
        except Exception as exc:
            job["errors"][0] = exc
            if st is None:  # the state the scheduler reserved is released too
                k = self._key(req)
                st = self.states.get(k) if k is not None else None
            job["states"].append([st] if st is not None else [])

    def done(self, job):
        return job["event"].query()

    def finish(self, job):
        job["event"].synchronize()
        results = []
        for i, (req, sts, err) in enumerate(zip(job["reqs"], job["states"], job["errors"])):
            if err is not None:
                results.append(err)
            elif "group" in job:
                n_in, end, ooff, outs = job["group"]
                if "values" not in job:  # views of the set's own buffer, or one copy
                    v = job["host"].numpy()
                    job["values"] = v if job["own"] else v.copy()
                vals = job["values"]
                first = 0 if req.kind == "slater_batch" else i  # atom index in the set
                got = []
                for j, st in enumerate(sts):
                    o = int(ooff[first + j]) - n_in
                    got.append(SlaterResult(vals[o:o + int(outs[first + j])], st.plan.offsets,
                                            st.plan.m, slices=st.plan.slices))
                results.append(got if req.kind == "slater_batch" else got[0])
            elif req.kind in ("slater", "slater_batch"):
                if "values" not in job:  # views of the set's own buffer, or one copy
                    v = job["host"].numpy()
                    job["values"] = v if job["own"] else v.copy()
                hv = job["values"]
                n_in = job["n_in"]
                got = []
                for j, st in enumerate(sts):
                    o, k = job["results"][(i, j if req.kind == "slater_batch" else -1)]
                    o -= n_in
                    got.append(SlaterResult(hv[o:o + k], st.plan.offsets, st.plan.m,
                                            slices=st.plan.slices))
                results.append(got if req.kind == "slater_batch" else got[0])
            else:  # the pinned result buffers are handed over (the arrays keep them alive)
                results.append(FockResult({k: v.numpy() for k, v in job["host"].items()}))
            for st in sts:
                self._release(st)
        self._slot_need[job["sid"]] = 0
        self._free_streams.append(job["sid"])
        return results

    def _release(self, st):
        self._set_busy(st, False)
        if st.key is None:
            st.close()

    def close(self):
        self._torch.cuda.synchronize()
        for st in self.states.values():
            st.close()
        self.states.clear()
        self._groups.clear()
        self._idle_bytes = 0
        self._fock_cache.clear()
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None


def _count_rk(ls, kmax):
    from .workloads import count_rk
    return int(count_rk(np.asarray(ls), int(kmax)))


class SlaterServer:
    """Scheduler + GPU executor.  ``budget_bytes`` defaults to 80 % of the
    device memory free at start-up."""

    def __init__(self, budget_bytes=None, max_inflight=4, max_batch=16):
        import torch
        self.executor = GpuExecutor(n_streams=max_inflight)
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            budget_bytes = int(0.8 * free)
        self.scheduler = Scheduler(self.executor, budget_bytes, max_inflight=max_inflight,
                                   max_batch=max_batch)

    def submit(self, req) -> Future:
        return self.scheduler.submit(req)

    def reset_stats(self):
        self.scheduler.reset_stats()

    def stats(self):
        s = self.scheduler.stats()
        s["plans_created"] = self.executor.plans_created
        s["resident_states"] = len(self.executor.states)
        return s

    def close(self):
        self.scheduler.close()
        self.executor.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
