"""GPU: batched launches over several R^k plans and HIP-graph replay.

The batch runs every plan's tiles, K splits and summation order unchanged,
so each plan's output must be bitwise identical alone and in any batch
(different grids, n_gauss values, orbital counts and k ranges; more plans than
one launch holds; a grid too large for the LDS-resident scan).  Values are
checked against the CPU oracle at 1e-12 (FP64) relative to sum_x |D||Y|.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# (nv, ng, n_orb, kmax, seed); nv = 2100 exceeds the LDS scan's capacity
SHAPES = [(60, 8, 7, 4, 1), (90, 7, 5, 6, 2), (40, 8, 12, 2, 3), (75, 9, 4, 3, 4),
          (2100, 8, 3, 2, 5)]


def _problem(nv, ng, n_orb, seed):
    from atomx_amd.workloads import synthetic_grid, synthetic_orbitals
    gr, grw = synthetic_grid(nv=nv, ng=ng, r_min=1e-5, r_max=60.0)
    P, Q, ls = synthetic_orbitals(gr, grw, n_orb=n_orb, seed=seed)
    return gr, grw, P, Q, ls


@pytest.fixture(scope="module")
def problems():
    return [_problem(nv, ng, n_orb, seed) + (kmax,) for nv, ng, n_orb, kmax, seed in SHAPES]


def _oracle_scaled_err(plan, out, gr, grw, P, Q, kmax):
    """max |GPU - oracle| / sum_x |w rho_p| |Y^k[rho_q]| over all values of the plan."""
    a, b = plan.pairs(P.shape[0])
    rho = P[a] * P[b] + Q[a] * Q[b]
    nv, ng = gr.shape
    w = grw.ravel()
    mo, mi = O.yk_partial_matrices(ng)
    worst = 0.0
    for k in range(kmax + 1):
        if plan.m[k] == 0:
            continue
        p, q = plan.quartets(k)
        ref = O.rk_quartets(rho, gr, grw, nv, ng, p, q, np.full(len(p), k))
        got = out[plan.offsets[k]:plan.offsets[k] + len(p)]
        ys = {j: np.abs(O.yk_flat(rho[j], gr.ravel(), w, w.size, k, nv, ng, mo, mi))
              for j in set(q.tolist())}
        scale = np.array([np.sum(np.abs(w * rho[i]) * ys[j]) for i, j in zip(p, q)])
        worst = max(worst, float(np.max(np.abs(got - ref) / scale)))
    return worst


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_batch_bitwise_equals_alone(problems, precision):
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor, run_batch
    plans, ins, alone = [], [], []
    for gr, grw, P, Q, ls, kmax in problems:
        pl = SlaterTensor(gr, grw, ls, kmax, precision=precision)
        Pd, Qd = dv.f64(P), dv.f64(Q)
        plans.append(pl)
        ins.append((Pd, Qd))
        alone.append(dv.host(pl.run(Pd, Qd)))
    outs = run_batch(plans, [a for a, _ in ins], [b for _, b in ins])
    for i, (o, ref) in enumerate(zip(outs, alone)):
        got = dv.host(o)
        print(f"{precision} plan {i} {SHAPES[i]}: {got.size} values")
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), f"plan {i} differs in the batch"
    # the mix in another order and with a repeated plan
    order = [3, 0, 2, 1, 0]
    outs2 = run_batch([plans[i] for i in order], [ins[i][0] for i in order],
                      [ins[i][1] for i in order])
    for j, i in enumerate(order):
        assert np.array_equal(dv.host(outs2[j]), alone[i])
    # one plan listed twice with different orbitals: each listing gets its own
    P0h, Q0h = problems[0][2], problems[0][3]
    P0b, Q0b = dv.f64(0.5 * P0h), dv.f64(0.75 * Q0h)
    alone_b = dv.host(plans[0].run(P0b, Q0b))
    outs3 = run_batch([plans[0], plans[1], plans[0]], [ins[0][0], ins[1][0], P0b],
                      [ins[0][1], ins[1][1], Q0b])
    assert np.array_equal(dv.host(outs3[0]), alone[0])
    assert np.array_equal(dv.host(outs3[1]), alone[1])
    assert np.array_equal(dv.host(outs3[2]), alone_b)
    assert not np.array_equal(alone_b, alone[0])
    if precision == "fp64":
        gr, grw, P, Q, ls, kmax = problems[1]
        assert _oracle_scaled_err(plans[1], alone[1], gr, grw, P, Q, kmax) < 1e-13
    for pl in plans:
        pl.close()


def test_contraction_splits_match_oracle(problems):
    """The FP64 contraction (each tile's K chunk split across the four waves
    of its work-group) at several K splits and block orders: oracle parity
    (grids of 630 and 675 nodes end in a partial 8-node stage; 3-12 orbitals
    leave partial 16x16 blocks at the segment edge), agreement between
    schedules, and bitwise alone-vs-batch with different K splits in one
    batch."""
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor, run_batch
    plans, ins, alone = [], [], []
    for i, (gr, grw, P, Q, ls, kmax) in enumerate(problems):
        pl = SlaterTensor(gr, grw, ls, kmax, precision="fp64")
        Pd, Qd = dv.f64(P), dv.f64(Q)
        pl.set_schedule("tile-major", 1)
        ref = dv.host(pl.run(Pd, Qd))
        scale = np.max(np.abs(ref))
        cap = max(s for s in range(1, 17) if _accepts_split(pl, s))
        for split in sorted({2, 3, cap}):
            if split > cap:
                continue
            for order in SlaterTensor.ORDERS:
                pl.set_schedule(order, split)
                got = dv.host(pl.run(Pd, Qd))
                err = np.max(np.abs(got - ref)) / scale
                print(f"plan {i} {SHAPES[i]} {order} split {split}: {err:.2e} vs split 1")
                assert err <= 1e-13
            if i < 4 and split == cap:
                assert _oracle_scaled_err(pl, got, gr, grw, P, Q, kmax) < 1e-12
        if i < 4:
            assert _oracle_scaled_err(pl, ref, gr, grw, P, Q, kmax) < 1e-12
        pl.set_schedule(("tile-major", "xcd-contiguous", "balanced-k-ranges")[i % 3],
                        min(1 + i, cap))
        plans.append(pl)
        ins.append((Pd, Qd))
        alone.append(dv.host(pl.run(Pd, Qd)))
    outs = run_batch(plans, [a for a, _ in ins], [b for _, b in ins])
    for i, (o, ref) in enumerate(zip(outs, alone)):
        assert np.array_equal(dv.host(o), ref), f"plan {i} differs in the batch"
    for pl in plans:
        pl.close()


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_single_chunk_tiles_write_every_value(problems, precision):
    """At K split 1 the static orders' contraction writes the packed values
    itself (no reduction launch; the balanced order still cuts tiles into
    pieces): every value of an output pre-filled with NaN is written, under
    every block order, and agrees with the default schedule's result."""
    import torch

    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    for i in (0, 1, 3):
        gr, grw, P, Q, ls, kmax = problems[i]
        pl = SlaterTensor(gr, grw, ls, kmax, precision=precision)
        Pd, Qd = dv.f64(P), dv.f64(Q)
        ref = dv.host(pl.run(Pd, Qd))
        scale = np.max(np.abs(ref))
        for order in SlaterTensor.ORDERS:
            pl.set_schedule(order, 1)
            out = torch.full((pl.size,), float("nan"), dtype=torch.float64, device=dv.device())
            got = dv.host(pl.run(Pd, Qd, out))
            assert not np.isnan(got).any(), f"plan {i} {order}: values left unwritten"
            err = np.max(np.abs(got - ref)) / scale
            print(f"{precision} plan {i} {SHAPES[i]} {order} split 1: {err:.2e}")
            assert err <= 1e-13
        pl.close()


def test_dynamic_order_repeats_and_replays(problems):
    """The dynamic contraction order resets its own ticket: ten launches in a
    row and three replays of a captured HIP graph (eager launches between
    them) all give the same bits."""
    import torch

    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    gr, grw, P, Q, ls, kmax = problems[0]
    plan = SlaterTensor(gr, grw, ls, kmax, precision="fp64")
    cap = max(s for s in range(1, 17) if _accepts_split(plan, s))
    plan.set_schedule("dynamic-heaviest-first", cap)
    Pd, Qd = dv.f64(P), dv.f64(Q)
    ref = dv.host(plan.run(Pd, Qd))
    for _ in range(10):
        assert np.array_equal(dv.host(plan.run(Pd, Qd)), ref)
    out = dv.empty((plan.size,))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run(Pd, Qd, out)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run(Pd, Qd, out)
    for rep in range(3):
        out.fill_(0.0)
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(dv.host(out), ref), f"replay {rep}"
        assert np.array_equal(dv.host(plan.run(Pd, Qd)), ref)
        torch.cuda.synchronize()
    del g
    torch.cuda.synchronize()
    plan.close()


def _accepts_split(plan, split):
    try:
        plan.set_schedule("tile-major", split)
    except ValueError:
        return False
    return True


def test_batch_more_plans_than_one_launch():
    """35 plans (two launch chunks of at most 32)."""
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor, run_batch
    plans, Ps, Qs, alone = [], [], [], []
    for i in range(35):
        gr, grw, P, Q, ls = _problem(12 + i, 8, 2 + i % 4, 100 + i)
        pl = SlaterTensor(gr, grw, ls, 1 + i % 3, precision="fp64")
        plans.append(pl)
        Ps.append(dv.f64(P))
        Qs.append(dv.f64(Q))
        alone.append(dv.host(pl.run(Ps[-1], Qs[-1])))
    outs = run_batch(plans, Ps, Qs)
    for o, ref in zip(outs, alone):
        assert np.array_equal(dv.host(o), ref)
    for pl in plans:
        pl.close()


def test_graph_replay_matches_eager(problems):
    """A batch captured into a HIP graph and replayed three times, with eager
    work between the replays, reproduces the eager result each time."""
    import torch

    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor, run_batch
    sel = problems[:3]
    plans = [SlaterTensor(gr, grw, ls, kmax, precision="split")
             for gr, grw, P, Q, ls, kmax in sel]
    Ps = [dv.f64(p[2]) for p in sel]
    Qs = [dv.f64(p[3]) for p in sel]
    eager = [dv.host(o) for o in run_batch(plans, Ps, Qs)]
    outs = [dv.empty((pl.size,)) for pl in plans]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run_batch(plans, Ps, Qs, outs)  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run_batch(plans, Ps, Qs, outs)
    for rep in range(3):
        for o in outs:
            o.fill_(0.0)
        g.replay()
        torch.cuda.synchronize()
        for i, (o, ref) in enumerate(zip(outs, eager)):
            assert np.array_equal(dv.host(o), ref), f"replay {rep} plan {i}"
        # eager work between replays
        run_batch(plans[:1], Ps[:1], Qs[:1])
        torch.cuda.synchronize()
    del g
    torch.cuda.synchronize()
    for pl in plans:
        pl.close()


def test_schedule_tuner(problems):
    """The tuner keeps a valid schedule, leaves that schedule's result in out
    (a later run reproduces it bitwise), and every schedule agrees with the
    default one to FP64 rounding."""
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    gr, grw, P, Q, ls, kmax = problems[0]
    for precision in ("fp64", "split"):
        plan = SlaterTensor(gr, grw, ls, kmax, precision=precision)
        Pd, Qd = dv.f64(P), dv.f64(Q)
        ref = dv.host(plan.run(Pd, Qd))
        out = dv.empty((plan.size,))
        order, split, ms = plan.tune(Pd, Qd, out, reps=2)
        print(f"{precision}: tuned order={order} split={split} {ms * 1e3:.1f} us")
        assert order in SlaterTensor.ORDERS and split >= 1 and ms > 0
        tuned = dv.host(out)
        assert np.array_equal(dv.host(plan.run(Pd, Qd)), tuned)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(tuned - ref)) <= 1e-13 * scale
        for o in SlaterTensor.ORDERS:
            plan.set_schedule(o, 1)
            got = dv.host(plan.run(Pd, Qd))
            assert np.max(np.abs(got - ref)) <= 1e-13 * scale
        with pytest.raises(ValueError):
            plan.set_schedule("tile-major", 64)
        plan.close()


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_pair_product_plan_matches_oracle(precision):
    """Explicit band pairs of B-spline products, every pair at every rank
    (the CI / CCSD product workload at a small size)."""
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor, run_batch
    from atomx_amd.workloads import bspline_product_workload
    gr, grw, B, pa, pb = bspline_product_workload(ns=30, ks=6, r_max=40.0)
    kmax = 3
    plan = SlaterTensor(gr, grw, None, kmax, precision=precision, pairs=(pa, pb),
                        n_fn=B.shape[0])
    out = dv.host(plan.run(B, None))
    nv, ng = gr.shape
    rho = B[pa] * B[pb]
    w = grw.ravel()
    mo, mi = O.yk_partial_matrices(ng)
    # B_0 .. B_{ks-1} are nonzero in the first cell [0, r_min], where r^k spans
    # many decades across the cell's Gauss nodes and the partial-interval rule
    # of the Y^k scan cancels heavily: any FP64 summation order (the
    # reference's included) is uncertain there at ~1e-12 of sum|D||Y| for
    # k >= 2.  Quartets away from the origin are held to 1e-13, quartets
    # touching the first cell to 1e-9.  On this 25-cell grid one K-tile spans
    # ~29 bits of r^-(k+1), beyond the split error bound, so a split plan
    # contracts in FP64 (slices == 0).
    tol = 1e-13
    ks = 6
    assert plan.slices == 0
    for k in range(kmax + 1):
        assert plan.m[k] == len(pa)
        p, q = plan.quartets(k)
        ref = O.rk_quartets(rho, gr, grw, nv, ng, p, q, np.full(len(p), k))
        got = out[plan.offsets[k]:plan.offsets[k] + len(p)]
        ys = {j: np.abs(O.yk_flat(rho[j], gr.ravel(), w, w.size, k, nv, ng, mo, mi))
              for j in set(q.tolist())}
        scale = np.array([np.sum(np.abs(w * rho[i]) * ys[j]) for i, j in zip(p, q)])
        rel = np.abs(got - ref) / scale
        origin = (pa[p] < ks) | (pa[q] < ks)
        err, err0 = float(np.max(rel[~origin])), float(np.max(rel[origin]))
        print(f"{precision} k={k}: max err / sum|D||Y| = {err:.2e} (first-cell quartets {err0:.2e})")
        assert err < tol, k
        assert err0 < 1e-9, k
    # in a batch with an all-pairs plan: bitwise the same
    gr2, grw2, P2, Q2, ls2 = _problem(50, 8, 5, 9)
    other = SlaterTensor(gr2, grw2, ls2, 2, precision=precision)
    o1, o2 = run_batch([plan, other], [dv.f64(B), dv.f64(P2)], [None, dv.f64(Q2)])
    assert np.array_equal(dv.host(o1), out)
    assert np.array_equal(dv.host(o2), dv.host(other.run(P2, Q2)))
    plan.close()
    other.close()


def test_slice_count_follows_the_grid():
    """7 slices where a 64-node K-tile spans a small radius ratio (the
    400-cell synthetic grid), 8 on coarser grids (200 cells to 60 a.u. at
    k <= 8, the 113-cell B-spline product grid at k <= 6), and an FP64
    contraction where even 8 slices would not meet the bound (100 cells at
    k <= 8)."""
    from atomx_amd.tensor import SlaterTensor
    from atomx_amd.workloads import bspline_product_workload, synthetic_grid
    ls = np.array([0, 1, 2], dtype=np.int32)
    for nv, expect in ((400, 7), (200, 8), (100, 0)):
        gr, grw = synthetic_grid(nv=nv, ng=8)
        pl = SlaterTensor(gr, grw, ls, 8, precision="split")
        assert pl.slices == expect, nv
        pl.close()
        pl = SlaterTensor(gr, grw, ls, 8, precision="fp64")
        assert pl.slices == 0
        pl.close()
    gr, grw, B, pa, pb = bspline_product_workload()
    pl = SlaterTensor(gr, grw, None, 6, precision="split", pairs=(pa, pb), n_fn=B.shape[0])
    assert pl.slices == 8
    pl.close()


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_create_many_equals_single_plans(problems, precision):
    """Plans set up by one ax_rk_tensor_create_many call (one staging upload,
    batched r-power tables, one shared set-up event) describe and compute
    bitwise what plans created one by one do, run alone or batched, and are
    closed one by one in any order (the shared event outlives the first)."""
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor, run_batch
    specs = [(gr, grw, ls, kmax) for gr, grw, P, Q, ls, kmax in problems]
    many = SlaterTensor.create_many(specs, precision)
    single = [SlaterTensor(gr, grw, ls, kmax, precision) for gr, grw, ls, kmax in specs]
    for a, b in zip(many, single):
        assert (a.size, a.m, a.offsets, a.slices, a.n_orb, a.n) == \
            (b.size, b.m, b.offsets, b.slices, b.n_orb, b.n)
    ins = [(P, Q) for _, _, P, Q, _, _ in problems]
    ref = [dv.host(p.run(P, Q)) for p, (P, Q) in zip(single, ins)]
    got = [dv.host(p.run(P, Q)) for p, (P, Q) in zip(many, ins)]
    for r, g in zip(ref, got):
        assert np.array_equal(r, g)
    outs = run_batch(many, [P for P, _ in ins], [Q for _, Q in ins])
    for r, o in zip(ref, outs):
        assert np.array_equal(r, dv.host(o))
    for k in (3, 0, 4, 1, 2):  # the shared set-up event is released by the last
        many[k].close()
    for p in single:
        p.close()


def test_create_many_on_stream(problems):
    """Set-up queued on a caller's stream for more plans than one set-up
    launch takes (16 per launch; 20 here): every plan computes bitwise what
    a plan created alone does, run on that stream or on another one, which
    waits for the set-up on the device."""
    import torch

    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    specs = [(gr, grw, ls, kmax) for gr, grw, P, Q, ls, kmax in problems] * 4
    ins = [(dv.to_dev(P), dv.to_dev(Q)) for _, _, P, Q, _, _ in problems]
    single = [SlaterTensor(gr, grw, ls, kmax, "split") for gr, grw, ls, kmax in specs[:5]]
    ref = [dv.host(p.run(P, Q)) for p, (P, Q) in zip(single, ins)]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    many = SlaterTensor.create_many(specs, "split", stream=s1)
    outs = []
    for i, p in enumerate(many):
        st = s1 if i % 2 == 0 else s2
        with torch.cuda.stream(st):
            outs.append(p.run(*ins[i % 5], stream=st))
    torch.cuda.synchronize()
    for i, o in enumerate(outs):
        assert np.array_equal(dv.host(o), ref[i % 5]), i
    for p in many + single:
        p.close()


def test_create_many_failure(problems):
    """A bad specification fails the whole call with the library's error (the
    library frees the plans it built before it); a later call on good
    specifications works."""
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    gr, grw, P, Q, ls, kmax = problems[0]
    bad = np.array(ls, dtype=np.int32).copy()
    bad[0] = -1
    with pytest.raises((RuntimeError, ValueError), match="negative l"):
        SlaterTensor.create_many([(gr, grw, ls, kmax), (gr, grw, bad, kmax)], "fp64")
    ok = SlaterTensor.create_many([(gr, grw, ls, kmax)], "fp64")
    single = SlaterTensor(gr, grw, ls, kmax, "fp64")
    assert np.array_equal(dv.host(ok[0].run(P, Q)), dv.host(single.run(P, Q)))
    ok[0].close()
    single.close()
