"""GPU: the native serving loop (ax_server_*, NativeSlaterServer).

Every request's values equal bitwise those of its plan run alone (a batch
runs each plan's tiles, K splits and summation order unchanged), for mixed
grids, orbital counts, k ranges and both precision modes, with and without
the small component, through more launch sets than the server holds in
flight; admission rejects a request above the budget; a bad request fails
at submission; values held by the caller outlive the server; plans set up
after the server has closed work.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(n, seed):
    from atomx_amd.serve import SlaterRequest
    from atomx_amd.workloads import synthetic_grid, synthetic_orbitals
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nv = int(rng.integers(30, 180))
        n_orb = int(rng.integers(2, 9))
        kmax = int(rng.integers(1, 7))
        gr, grw = synthetic_grid(nv=nv, ng=8, r_min=1e-5, r_max=float(rng.uniform(30, 60)))
        P, Q, ls = synthetic_orbitals(gr, grw, n_orb=n_orb, seed=seed * 1000 + i)
        prec = "split" if i % 2 else "fp64"
        out.append(SlaterRequest(gr=gr, grw=grw, ls=ls, kmax=kmax, P=P,
                                 Q=Q if i % 3 else None, precision=prec))
    return out


def _alone(req):
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    pl = SlaterTensor(req.gr, req.grw, req.ls, req.kmax, req.precision)
    try:
        return dv.host(pl.run(req.P, req.Q)), pl.slices, pl.m, pl.offsets
    finally:
        pl.close()


def test_native_server_bitwise_equals_plans_alone():
    from atomx_amd.serve_native import NativeSlaterServer
    reqs = _requests(37, seed=11)
    with NativeSlaterServer(max_inflight=3, max_batch=5) as srv:
        tickets = [srv.submit(r) for r in reqs]
        res = [srv.result(t) for t in tickets]
        st = srv.stats()
    assert st["plans"] == len(reqs) and st["launches"] >= len(reqs) // 5
    for r, x in zip(reqs, res):
        ref, slices, m, offs = _alone(r)
        assert np.array_equal(x.values, ref)
        assert (x.slices, x.m, x.offsets) == (slices, m, offs)


def test_native_server_budget_errors_and_held_values():
    from atomx_amd.serve_native import NativeSlaterServer
    reqs = _requests(3, seed=12)
    srv = NativeSlaterServer(budget_bytes=1 << 20, max_inflight=2, max_batch=4)
    with pytest.raises(MemoryError):  # one request alone above the budget
        srv.submit(reqs[0])
    srv.close()
    with NativeSlaterServer(max_inflight=2, max_batch=4) as srv:
        bad = _requests(1, seed=13)[0]
        bad.ls = np.array(bad.ls).copy()
        bad.ls[0] = -1
        with pytest.raises(ValueError):
            srv.submit(bad)
        tickets = [srv.submit(r) for r in reqs]
        held = [srv.result(t) for t in tickets]
    # the server is gone; the values stay valid
    for r, x in zip(reqs, held):
        assert np.array_equal(x.values, _alone(r)[0])


def test_plans_after_server_close():
    """Plans set up after a server has closed (its streams destroyed) work:
    the server's one-plan launch sets used set-up staging of the same size
    class as the single plans made afterwards, and that staging went away
    with the server's streams instead of being handed out again."""
    from atomx_amd.serve_native import NativeSlaterServer
    from atomx_amd.tensor import SlaterTensor
    from atomx_amd import device as dv
    reqs = _requests(6, seed=14)
    with NativeSlaterServer(max_inflight=2, max_batch=1) as srv:
        res = [srv.result(t) for t in [srv.submit(r) for r in reqs]]
    for r, x in zip(reqs, res):
        assert np.array_equal(x.values, _alone(r)[0])
    many = SlaterTensor.create_many([(r.gr, r.grw, r.ls, r.kmax) for r in reqs[:2]], "fp64")
    for pl, r in zip(many, reqs[:2]):
        one = SlaterTensor(r.gr, r.grw, r.ls, r.kmax, "fp64")
        assert np.array_equal(dv.host(pl.run(r.P, r.Q)), dv.host(one.run(r.P, r.Q)))
        one.close()
        pl.close()


def _scf(atoms, iterations, seed):
    """SCF-style requests: every atom once per iteration under its own state,
    new orbitals each iteration."""
    from atomx_amd.serve import SlaterRequest
    rng = np.random.default_rng(seed)
    out = []
    for it in range(iterations):
        for a, r in enumerate(atoms):
            f = 1.0 + 1e-3 * (it + 1) * rng.standard_normal()
            out.append(SlaterRequest(gr=r.gr, grw=r.grw, ls=r.ls, kmax=r.kmax, P=r.P * f,
                                     Q=None if r.Q is None else r.Q * f,
                                     precision=r.precision, state=("atom", a)))
    return out


def test_native_resident_states():
    """Requests naming a resident state reuse the plan of its first request:
    one plan made per state over three SCF iterations, every value bitwise
    that of the plan run alone on the same orbitals."""
    from atomx_amd.serve_native import NativeSlaterServer
    reqs = _scf(_requests(6, seed=21), 3, seed=22)
    with NativeSlaterServer(max_inflight=3, max_batch=4) as srv:
        res = [srv.result(t) for t in [srv.submit(r) for r in reqs]]
        st = srv.stats()
        assert st["plans"] == len(reqs)
        assert st["plans_made"] == 6 and st["resident_states"] == 6
        srv.drop_state(("atom", 0))
        assert srv.stats()["resident_states"] == 5
    for r, x in zip(reqs, res):
        assert np.array_equal(x.values, _alone(r)[0])


def test_native_state_shape_and_eviction():
    """A request whose grid differs from its state's is refused at submit;
    with four states of one shape under a budget of two of their requests,
    idle states give up their plans (made again on their next request) and
    the values stay bitwise those of the plans alone."""
    import ctypes

    from atomx_amd._lib import PRECISIONS, load
    from atomx_amd.serve import SlaterRequest
    from atomx_amd.serve_native import NativeSlaterServer
    from atomx_amd.workloads import count_rk, synthetic_grid, synthetic_orbitals
    atoms = []
    for i in range(4):
        gr, grw = synthetic_grid(nv=120, ng=8, r_min=1e-5, r_max=40.0 + i)
        P, Q, ls = synthetic_orbitals(gr, grw, n_orb=6, seed=230 + i)
        atoms.append(SlaterRequest(gr=gr, grw=grw, ls=ls, kmax=4, P=P, Q=Q,
                                   precision="split" if i % 2 else "fp64"))
    lib = load()

    def need(r):
        ls = np.ascontiguousarray(r.ls, dtype=np.int32)
        nv, ng = np.shape(r.gr)
        plan = int(lib.ax_rk_tensor_estimate(nv, ng, ls.size,
                                             ls.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                             int(r.kmax), PRECISIONS[r.precision]))
        n_in = ls.size * nv * ng * (1 if r.Q is None else 2)
        return plan, plan + 8 * (n_in + count_rk(r.ls, r.kmax))

    sizes = [need(r) for r in atoms]
    assert all(plan > total - plan for plan, total in sizes)  # plans outweigh staging
    budget = 2 * max(total for _, total in sizes)
    reqs = _scf(atoms, 3, seed=24)
    with NativeSlaterServer(budget_bytes=budget, max_inflight=2, max_batch=2) as srv:
        srv.result(srv.submit(reqs[0]))
        a = atoms[0]
        other = SlaterRequest(gr=a.gr * 1.0001, grw=a.grw, ls=a.ls, kmax=a.kmax, P=a.P, Q=a.Q,
                              precision=a.precision, state=("atom", 0))
        with pytest.raises(ValueError, match="another shape"):
            srv.submit(other)
        res = [srv.result(t) for t in [srv.submit(r) for r in reqs]]
        st = srv.stats()
    assert st["plans_made"] > 4  # evicted states made their plans again
    assert st["peak_bytes"] <= budget
    for r, x in zip(reqs, res):
        assert np.array_equal(x.values, _alone(r)[0])
