"""GPU: the request server.  Every result must be bitwise what the same
request gives when computed alone, whatever else is in flight; values are
checked against the CPU oracle (FP64 R^k at 1e-13 of sum |D||Y|, Fock pieces
at 1e-12 relative to the largest entry); admission keeps the accounted
device bytes under the budget; resident states reuse their plan."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _slater_req(nv, n_orb, kmax, seed, precision="split", state=None):
    from atomx_amd.serve import SlaterRequest
    from atomx_amd.workloads import synthetic_grid, synthetic_orbitals
    gr, grw = synthetic_grid(nv=nv, ng=8, r_min=1e-5, r_max=60.0)
    P, Q, ls = synthetic_orbitals(gr, grw, n_orb=n_orb, seed=seed)
    return SlaterRequest(gr, grw, P, Q, ls, kmax, precision=precision, state=state)


def _fock_req(seed, n_orb=4, state=None):
    from atomx_amd import bspline as B
    from atomx_amd.serve import FockRequest
    t, ns = B.make_physics_knots(ks=7, Z=6, r_max=40.0)
    bd = B.prepare_bspline_data(t, ns, 7)
    rng = np.random.default_rng(seed)
    decay = np.exp(-np.arange(ns) / 6.0)
    P = rng.normal(size=(n_orb, ns)) * decay
    Q = 0.01 * rng.normal(size=(n_orb, ns)) * decay
    occ = rng.integers(1, 3, size=n_orb).astype(float)
    coef = np.zeros((n_orb, n_orb, 3))
    for tt in range(n_orb):
        for b in range(n_orb):
            coef[tt, b, (tt + b) % 3] = 0.1 * (1 + tt + 2 * b)
    return FockRequest(bd, P, Q, occ, coef, state=state)


def _alone_slater(req):
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    plan = SlaterTensor(req.gr, req.grw, req.ls, req.kmax, precision=req.precision)
    out = dv.host(plan.run(req.P, req.Q))
    plan.close()
    return out


def _alone_at(req, schedule):
    """The request alone on a plan set to a given contraction schedule."""
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    plan = SlaterTensor(req.gr, req.grw, req.ls, req.kmax, precision=req.precision)
    plan.set_schedule(*schedule)
    out = dv.host(plan.run(req.P, req.Q))
    plan.close()
    return out


def _alone_fock(req):
    from atomx_amd import device as dv
    from atomx_amd.fock_yk import FockYk
    blocks = FockYk(req.bsp_data).build(req.P, req.Q, req.occ, req.exchange_coef)
    return {k: dv.host(v) for k, v in blocks.items()}


def test_mixed_stream_same_as_alone():
    from atomx_amd.serve import SlaterServer
    reqs = [_slater_req(60 + 10 * i, 3 + (i % 5), 2 + (i % 4), 10 + i,
                        precision="split" if i % 3 else "fp64") for i in range(14)]
    reqs.insert(3, _fock_req(1))
    reqs.insert(9, _fock_req(2, n_orb=3))
    with SlaterServer(max_inflight=3, max_batch=6) as srv:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=120) for f in futs]
        st = srv.stats()
    print(st)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["launches"] < len(reqs), "requests were not batched"
    for i, (req, r) in enumerate(zip(reqs, res)):
        if req.kind == "slater":
            assert np.array_equal(r.values, _alone_slater(req)), f"request {i}"
        else:
            alone = _alone_fock(req)
            for k in alone:
                assert np.array_equal(r.blocks[k], alone[k]), f"request {i} block {k}"


def test_slater_request_matches_oracle():
    from atomx_amd.serve import SlaterServer
    req = _slater_req(50, 4, 3, 5, precision="fp64")
    with SlaterServer() as srv:
        r = srv.submit(req).result(timeout=60)
    a, b = np.triu_indices(4)
    rho = req.P[a] * req.P[b] + req.Q[a] * req.Q[b]
    from atomx_amd.workloads import valid_pairs
    nv, ng = req.gr.shape
    w = req.grw.ravel()
    mo, mi = O.yk_partial_matrices(ng)
    for k in range(req.kmax + 1):
        V = valid_pairs(np.asarray(req.ls), k)
        if len(V) == 0:
            continue
        i, j = np.triu_indices(len(V))
        p, q = V[i], V[j]
        ref = O.rk_quartets(rho, req.gr, req.grw, nv, ng, p, q, np.full(len(p), k))
        got = r.values[r.offsets[k]:r.offsets[k] + len(p)]
        ys = {x: np.abs(O.yk_flat(rho[x], req.gr.ravel(), w, w.size, k, nv, ng, mo, mi))
              for x in set(q.tolist())}
        scale = np.array([np.sum(np.abs(w * rho[x]) * ys[y]) for x, y in zip(p, q)])
        assert np.max(np.abs(got - ref) / scale) < 1e-13, k


def test_fock_request_matches_oracle():
    from atomx_amd.serve import SlaterServer
    req = _fock_req(7)
    with SlaterServer() as srv:
        r = srv.submit(req).result(timeout=60)
    bd = req.bsp_data
    n_orb = req.P.shape[0]
    for t in range(n_orb):
        vpp, vqq = O.direct_potential_yk(req.P, req.Q, req.occ, bd, target=t)
        kpp = np.zeros_like(vpp)
        kqq = np.zeros_like(vqq)
        for b in range(n_orb):
            for k in range(req.exchange_coef.shape[2]):
                c = req.exchange_coef[t, b, k]
                if c != 0.0:
                    xp, xq = O.compute_exchange_matrix(req.P[b], req.Q[b], bd, k)
                    kpp += c * xp
                    kqq += c * xq
        for name, ref in (("VJ_PP", vpp), ("VJ_QQ", vqq), ("K_PP", kpp), ("K_QQ", kqq)):
            got = r.blocks[name][t]
            assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), (t, name)


def test_resident_state_and_budget():
    from atomx_amd.serve import GpuExecutor, SlaterServer
    base = _slater_req(80, 5, 3, 21, state="atom-1")
    iters = []
    for it in range(3):  # SCF iterations: same shape, new orbitals
        r = _slater_req(80, 5, 3, 21, state="atom-1")
        r.P = base.P * (1.0 + 0.01 * it)
        r.Q = base.Q * (1.0 - 0.01 * it)
        iters.append(r)
    others = [_slater_req(70, 4, 2, 30 + i) for i in range(6)]
    ordered = iters[:1] + others[:3] + iters[1:] + others[3:]
    alone = [_alone_slater(r) for r in ordered]
    # ample memory: the resident state keeps one plan across its iterations
    with SlaterServer(max_inflight=4) as srv:
        res = [f.result(timeout=120) for f in [srv.submit(r) for r in ordered]]
        st = srv.stats()
    print(st)
    for r, ref in zip(res, alone):
        assert np.array_equal(r.values, ref)
    assert st["plans_created"] == 1 + len(others)
    # tight memory: admission keeps the accounted bytes under the budget
    # (idle resident states are evicted when needed) and results do not change
    probe = GpuExecutor(n_streams=1)
    need = max(probe.footprint(r) for r in ordered)
    probe.close()
    budget = int(2.5 * need)
    with SlaterServer(budget_bytes=budget, max_inflight=4) as srv:
        res = [f.result(timeout=120) for f in [srv.submit(r) for r in ordered]]
        st = srv.stats()
    print(st)
    assert st["peak_bytes"] <= budget
    for r, ref in zip(res, alone):
        assert np.array_equal(r.values, ref)


@pytest.mark.parametrize("own_results", [False, True])
def test_batch_request_equals_its_atoms_alone(own_results):
    """A SlaterBatchRequest (resident states per atom, mixed with single
    requests) returns, per atom, bitwise what the atom gives alone; with
    results copied out of the staging, or (own_results) as views of each
    launch set's own pinned buffer."""
    from atomx_amd.serve import SlaterBatchRequest, SlaterServer
    atoms = [_slater_req(24 + 8 * i, 2 + i % 4, 2 + i % 3, 40 + i, precision="fp64",
                         state=("b", i)) for i in range(5)]
    single = _slater_req(30, 3, 2, 99, precision="fp64")
    with SlaterServer(max_inflight=2, max_batch=4) as srv:
        if own_results:
            srv.executor._OWN_RESULTS = 1  # every launch set gets its own buffer
        for it in range(2):
            fb = srv.submit(SlaterBatchRequest(atoms))
            fs = srv.submit(single)
            got = fb.result()
            if it == 0:
                first = got  # results are views of their launch set's own buffer
            assert len(got) == len(atoms)
            for a, res in zip(atoms, got):
                assert np.array_equal(res.values, _alone_slater(a))
            assert np.array_equal(fs.result().values, _alone_slater(single))
        assert srv.stats()["plans_created"] == len(atoms) + 2
        # later iterations: new request objects and orbitals on the same
        # states (the cached group layout), then one atom with wrong shapes
        from dataclasses import replace
        for it in range(2):
            fresh = [replace(a, P=a.P * (1.0 + 0.01 * (it + 1)),
                             Q=None if a.Q is None else a.Q * (1.0 - 0.01 * it)) for a in atoms]
            got = srv.submit(SlaterBatchRequest(fresh)).result()
            for a, res in zip(fresh, got):
                assert np.array_equal(res.values, _alone_slater(a))
        bad = list(atoms)
        bad[2] = replace(atoms[2], P=atoms[2].P[:, :-1])
        with pytest.raises(ValueError):
            srv.submit(SlaterBatchRequest(bad)).result()
        got = srv.submit(SlaterBatchRequest(atoms)).result()
        for a, res in zip(atoms, got):
            assert np.array_equal(res.values, _alone_slater(a))
        # the same states as single requests, twice (launch sets of four
        # recurring requests take the cached layout the second time)
        for it in range(2):
            fresh = [replace(a, P=a.P * (1.0 + 0.02 * it)) for a in atoms]
            futs = [srv.submit(a) for a in fresh]
            for a, f in zip(fresh, futs):
                assert np.array_equal(f.result().values, _alone_slater(a))
    # the first results are unchanged by every launch set served after them
    for a, res in zip(atoms, first):
        assert np.array_equal(res.values, _alone_slater(a))


def test_resident_state_is_tuned_once():
    """A new resident state of a large enough tensor is tuned on its first
    request; its results equal, bitwise, an alone run at the tuned schedule,
    and later requests keep that schedule."""
    from atomx_amd import device as dv
    from atomx_amd.serve import SlaterServer
    from atomx_amd.tensor import SlaterTensor
    req = _slater_req(120, 9, 4, 7, precision="fp64", state=("tuned", 0))
    with SlaterServer(max_inflight=2, max_batch=4) as srv:
        srv.executor._TUNE_MIN_RK = 1
        first = srv.submit(req).result()
        st = srv.executor.states[("slater", ("tuned", 0))]
        sched = st.plan.schedule()
        again = srv.submit(req).result()
        assert st.plan.schedule() == sched
    alone = SlaterTensor(req.gr, req.grw, req.ls, req.kmax, precision="fp64")
    alone.set_schedule(*sched)
    ref = dv.host(alone.run(req.P, req.Q))
    alone.close()
    print("tuned schedule", sched)
    assert np.array_equal(first.values, ref)
    assert np.array_equal(again.values, ref)


def test_bad_new_state_fails_alone_in_a_tuned_launch_set():
    """A launch set holding a good request and the first request of a new
    resident state whose orbitals have the wrong shape (with tuning of new
    states forced on): the bad request fails alone, the good one equals its
    alone result, and the executor keeps serving afterwards."""
    import time

    from atomx_amd.serve import GpuExecutor, SlaterServer
    good = _slater_req(60, 5, 3, 11, precision="fp64", state=("good", 0))
    bad = _slater_req(64, 6, 3, 12, precision="fp64", state=("bad", 0))
    bad.P = bad.P[:, :-8]  # wrong number of samples
    ex = GpuExecutor(n_streams=2)
    try:
        ex._TUNE_MIN_RK = 1
        for r in (good, bad):
            ex.reserve(r)
        h = ex.launch([good, bad])
        while not ex.done(h):
            time.sleep(1e-3)
        res = ex.finish(h)
        assert isinstance(res[1], ValueError)
        # the good state was tuned on this launch set: compare at its schedule
        sched = ex.states[("slater", ("good", 0))].plan.schedule()
        assert np.array_equal(res[0].values, _alone_at(good, sched))
        # the same states again: the bad state's plan was never tuned or listed
        ok = _slater_req(64, 6, 3, 12, precision="fp64", state=("bad", 0))
        for r in (good, ok):
            ex.reserve(r)
        h = ex.launch([good, ok])
        while not ex.done(h):
            time.sleep(1e-3)
        res = ex.finish(h)
        sched_ok = ex.states[("slater", ("bad", 0))].plan.schedule()
        assert np.array_equal(res[0].values, _alone_at(good, sched))
        assert np.array_equal(res[1].values, _alone_at(ok, sched_ok))
    finally:
        ex.close()
    with SlaterServer(max_inflight=2, max_batch=4) as srv:
        srv.executor._TUNE_MIN_RK = 1
        bad2 = _slater_req(64, 6, 3, 13, precision="fp64", state=("bad", 1))
        bad2.P = bad2.P[:, :-8]
        f_bad = srv.submit(bad2)
        f_good = srv.submit(good)
        with pytest.raises(ValueError):
            f_bad.result(timeout=60)
        got = f_good.result(timeout=60).values
        sched = srv.executor.states[("slater", ("good", 0))].plan.schedule()
        assert np.array_equal(got, _alone_at(good, sched))
        later = srv.submit(_slater_req(60, 5, 3, 14, precision="fp64"))
        assert later.result(timeout=60).values.size > 0
