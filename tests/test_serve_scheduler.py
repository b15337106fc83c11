"""CPU tests of the request scheduler (continuous batching, admission by
footprint, residency, error propagation) with a stand-in executor whose
"device work" is a timer."""

import threading
import time

import pytest

from atomx_amd.serve import Scheduler, percentile


class Req:
    def __init__(self, rid, nbytes, dur=0.002, kind="slater", state=None, key="split",
                 fail=False):
        self.rid, self.nbytes, self.dur = rid, nbytes, dur
        self.kind, self.state, self.key, self.fail = kind, state, key, fail


class FakeExecutor:
    def __init__(self):
        self.lock = threading.Lock()
        self.inflight_bytes = 0
        self.peak_inflight_bytes = 0
        self.batches = []
        self.states = {}          # state -> [bytes, busy]
        self.state_overlap = False
        self.launched_states = set()

    def ready(self, req):
        return req.state is None or req.state not in self.states or not self.states[req.state][1]

    def footprint(self, req):
        return req.nbytes

    def reserve(self, req):
        if req.state is not None:
            st = self.states.setdefault(req.state, [req.nbytes, False])
            if st[1]:
                self.state_overlap = True
            st[1] = True

    def resident_bytes(self):
        return sum(b for b, busy in self.states.values() if not busy)

    def evict(self, nbytes):
        freed = 0
        for s in list(self.states):
            if freed >= nbytes:
                break
            b, busy = self.states[s]
            if not busy:
                freed += b
                del self.states[s]
        return freed

    def batch_key(self, req):
        return (req.kind, req.key) if req.kind == "slater" else (req.kind, req.rid)

    def launch(self, reqs):
        with self.lock:
            nb = sum(r.nbytes for r in reqs)
            self.inflight_bytes += nb
            self.peak_inflight_bytes = max(self.peak_inflight_bytes, self.inflight_bytes)
            self.batches.append([r.rid for r in reqs])
        return {"reqs": reqs, "t_end": time.perf_counter() + max(r.dur for r in reqs), "nb": nb}

    def done(self, h):
        return time.perf_counter() >= h["t_end"]

    def finish(self, h):
        with self.lock:
            self.inflight_bytes -= h["nb"]
        out = []
        for r in h["reqs"]:
            if r.state is not None:
                self.states[r.state][1] = False
            out.append(RuntimeError(f"request {r.rid} failed") if r.fail else ("ok", r.rid))
        return out


def test_all_complete_fifo_batches():
    ex = FakeExecutor()
    sch = Scheduler(ex, budget_bytes=10**9, max_inflight=2, max_batch=4)
    futs = [sch.submit(Req(i, 100)) for i in range(37)]
    res = [f.result(timeout=10) for f in futs]
    sch.close()
    assert res == [("ok", i) for i in range(37)]
    order = [rid for b in ex.batches for rid in b]
    assert order == list(range(37)), "launch order must be FIFO"
    assert all(1 <= len(b) <= 4 for b in ex.batches)
    st = sch.stats()
    assert st["completed"] == 37 and st["failed"] == 0 and st["peak_inflight"] <= 2
    assert st["p50_ms"] <= st["p99_ms"]


def test_batches_split_on_kind_and_precision():
    ex = FakeExecutor()
    sch = Scheduler(ex, budget_bytes=10**9, max_inflight=1, max_batch=16)
    reqs = [Req(0, 1), Req(1, 1), Req(2, 1, key="fp64"), Req(3, 1, kind="fock"), Req(4, 1)]
    # submit everything before the scheduler can launch, so batching is decided on the queue
    futs = [sch.submit(r) for r in reqs]
    for f in futs:
        f.result(timeout=10)
    sch.close()
    for b in ex.batches:
        kinds = {(reqs[i].kind, reqs[i].key) for i in b}
        assert len(kinds) == 1
        assert not any(reqs[i].kind == "fock" for i in b) or len(b) == 1


def test_admission_respects_budget():
    ex = FakeExecutor()
    sch = Scheduler(ex, budget_bytes=1000, max_inflight=8, max_batch=16)
    futs = [sch.submit(Req(i, 300, dur=0.003)) for i in range(20)]
    for f in futs:
        f.result(timeout=10)
    sch.close()
    assert ex.peak_inflight_bytes <= 1000
    assert sch.stats()["peak_bytes"] <= 1000


def test_oversized_request_fails_alone():
    ex = FakeExecutor()
    sch = Scheduler(ex, budget_bytes=1000, max_inflight=2)
    f_big = sch.submit(Req(0, 5000))
    f_ok = sch.submit(Req(1, 10))
    with pytest.raises(MemoryError):
        f_big.result(timeout=10)
    assert f_ok.result(timeout=10) == ("ok", 1)
    sch.close()


def test_executor_error_reaches_its_future_only():
    ex = FakeExecutor()
    sch = Scheduler(ex, budget_bytes=10**6, max_inflight=2)
    f1 = sch.submit(Req(0, 10, fail=True))
    f2 = sch.submit(Req(1, 10))
    with pytest.raises(RuntimeError, match="request 0"):
        f1.result(timeout=10)
    assert f2.result(timeout=10) == ("ok", 1)
    sch.close()
    assert sch.stats()["failed"] == 1


def test_resident_state_never_in_flight_twice_and_is_evicted():
    ex = FakeExecutor()
    sch = Scheduler(ex, budget_bytes=1000, max_inflight=4, max_batch=8)
    futs = [sch.submit(Req(i, 400, dur=0.002, state="atomA" if i % 2 == 0 else None))
            for i in range(12)]
    for f in futs:
        f.result(timeout=10)
    assert not ex.state_overlap
    assert "atomA" in ex.states  # resident after completion
    # a different state that needs the memory evicts the idle resident one
    f = sch.submit(Req(99, 900, state="atomB"))
    assert f.result(timeout=10) == ("ok", 99)
    sch.close()
    assert "atomA" not in ex.states and "atomB" in ex.states


def test_percentile_nearest_rank():
    v = list(range(1, 101))
    assert percentile(v, 50) == 50 and percentile(v, 99) == 99 and percentile(v, 100) == 100
    assert percentile([3.0], 99) == 3.0


def test_threaded_packing_equals_concatenate():
    """The executor's input packing (equal ranges of the concatenation, one
    per thread, crossing chunk boundaries, empty chunks included) writes
    exactly np.concatenate of the chunks and nothing past them."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from atomx_amd.serve import GpuExecutor
    ex = object.__new__(GpuExecutor)  # packing state only: no device needed
    ex._pack_threads = 5
    ex._pool = ThreadPoolExecutor(5)
    ex._PACK_SERIAL = 0
    rng = np.random.default_rng(0)
    try:
        for _ in range(200):
            chunks = [rng.standard_normal(int(rng.integers(0, 40)))
                      for _ in range(int(rng.integers(1, 10)))]
            n = sum(c.size for c in chunks)
            dst = np.full(n + 3, np.nan)
            ex._pack(chunks, dst[:n])
            assert np.array_equal(dst[:n], np.concatenate(chunks))
            assert np.isnan(dst[n:]).all()
        with pytest.raises(ValueError):
            ex._pack([np.ones(4)], np.empty(5))
    finally:
        ex._pool.shutdown()


def test_scheduler_failure_rejects_later_submits():
    """An exception that escapes the executor ends the scheduler thread:
    every outstanding future fails with it and later submits raise instead
    of queueing work that would never run."""

    class Broken(FakeExecutor):
        def launch(self, reqs):
            raise KeyError("executor bug")

    sch = Scheduler(Broken(), budget_bytes=10**6, max_inflight=2)
    f = sch.submit(Req(0, 10))
    with pytest.raises(KeyError):
        f.result(timeout=10)
    sch._thread.join(timeout=10)
    with pytest.raises(RuntimeError, match="stopped after an error"):
        sch.submit(Req(1, 10))
