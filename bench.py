"""Benchmark: R^k integrals/s of the batched all-quartet Slater-integral tensor.

Workload (one GPU): exponential grid of 400 intervals x 8 Gauss points = 3200
nodes, r from 1e-5 to 60 a.u.; 40 synthetic radial orbitals (seed 0,
P = r^(l+1) exp(-Z r/n)(1 + sum c_j r^j), Q = P'/2c); every R^k, k = 0..8,
over quartets valid by the triangle and parity rules, each unordered
pair-pair once.  One step = one full pass (pair densities, FP64 Y^k scans,
contraction) with the orbitals already resident in HBM.

Prints ONE JSON line (rank 0).  The split mode is compared with the FP64 mode
on every value before timing, and only qualifies for the headline value at
<= 1e-11; the CPU baseline leg recomputes the whole tensor with the FP64 CPU
oracle and reports both modes' errors against it.  Further legs: the B-spline
product workload (pair x pair R^k, k = 0..6) and the served request stream
(p50/p99 latency).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "R^k integrals/s at \u22641e-11 rel err vs FP64 oracle; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK = 78.6      # TFLOP/s dense
INT8_MFMA_PEAK = 5000.0    # TOP/s dense


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=("split", "fp64", "auto"), default="auto")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tensor", action="store_true",
                    help="skip the benchmark tensor (and the CPU leg): run the other legs alone")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all available (<= 16)")
    ap.add_argument("--nv", type=int, default=400)
    ap.add_argument("--ng", type=int, default=8)
    ap.add_argument("--n-orb", type=int, default=40)
    ap.add_argument("--kmax", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--bank-steps", type=int, default=10,
                    help="timed steps of the B-spline product workload (0 = skip that leg)")
    ap.add_argument("--ion-steps", type=int, default=20,
                    help="timed SCF-iteration Fock builds of the 26 iron ions (0 = skip)")
    ap.add_argument("--no-ion-scf", action="store_true",
                    help="skip the batched SCF of the 26 iron ions")
    ap.add_argument("--no-neon", action="store_true", help="skip the neon Dirac-Fock SCF leg")
    ap.add_argument("--no-correlation", action="store_true",
                    help="skip the MBPT(2) and spinor-export leg (it runs on the neon leg's state)")
    ap.add_argument("--serve-requests", type=int, default=300,
                    help="requests in the served Poisson stream (0 = skip the serving leg)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB device-to-device once "
                         "(a kernel of known traffic for calibrating HBM byte counters)")
    return ap.parse_args()


def neon_oracle_scf(bd):
    """The neon SCF of the neon leg with the oracle's FP64 two-electron Fock
    blocks and energy (CPU baseline leg: the checker of the GPU SCF)."""
    import oracle as O
    from atomx_amd import dirac as D

    shells = D.neon_subshells()
    occ = [s.occupation for s in shells]
    kap = [s.kappa for s in shells]
    banks = O.Banks(bd)
    tables = O.angular_tables(occ, kap)

    class Backend:
        def fock(self, i, P, Q):
            return O.build_hf_matrix_dbsr(i, P, Q, occ, kap, bd, banks, skip_pq_exchange=True,
                                          tables=tables)

        def energy(self, P, Q):
            return O.two_electron_energy(P, Q, occ, kap, bd, banks)

    scf = D.DiracFockSCF(10, shells, bd, backend=Backend()).bare_nucleus_start()
    return scf.run(), scf.iterations


def agreement(got, ref, offsets, m):
    """Error of ``got`` against ``ref`` (packed per rank): the largest
    |got - ref| / max|ref of that rank|, and the largest elementwise relative
    error over values with |ref| >= 1e-3 max|ref of that rank|."""
    worst_norm = worst_el = 0.0
    for k, mk in enumerate(m):
        if mk == 0:
            continue
        a, b = offsets[k], offsets[k] + mk * (mk + 1) // 2
        g, r = got[a:b], ref[a:b]
        big = float(np.max(np.abs(r)))
        if big == 0.0:
            continue
        worst_norm = max(worst_norm, float(np.max(np.abs(g - r))) / big)
        mask = np.abs(r) >= 1e-3 * big
        worst_el = max(worst_el, float(np.max(np.abs(g[mask] - r[mask]) / np.abs(r[mask]))))
    return worst_norm, worst_el


_T0 = time.perf_counter()


def progress(leg):
    """One line on stderr per finished leg (the JSON line stays alone on stdout)."""
    print(f"bench: {leg} done at {time.perf_counter() - _T0:.1f} s", file=sys.stderr, flush=True)


def time_steps(fn, steps, stream):
    import torch
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps  # ms


def cpu_baseline(P, Q, gr, grw, ls, kmax, threads, min_seconds=6.0, native=True):
    """The oracle's C FP64 restatement (Y^k per (pair, k) once, then one
    sum_x w rho_p Y_q per R^k), OpenMP over pairs/quartets, full workload.
    native: the -O2 -march=native build for this host (the timed baseline);
    otherwise the strict checker build.  Returns (R^k/s, seconds, repetitions,
    values in the tensor's packed order)."""
    import oracle as O
    from atomx_amd.workloads import pair_list, valid_pairs
    nv, ng = gr.shape
    a, b = pair_list(P.shape[0])
    rho = P[a] * P[b] + Q[a] * Q[b]
    qp, qq, qk = [], [], []
    for k in range(kmax + 1):
        V = valid_pairs(ls, k)
        i, j = np.triu_indices(len(V))
        qp.append(V[i]); qq.append(V[j]); qk.append(np.full(i.size, k))
    qp, qq, qk = (np.concatenate(x).astype(np.int32) for x in (qp, qq, qk))
    reps, t0 = 0, time.perf_counter()
    while True:
        vals = O.rk_quartets(rho, gr, grw, nv, ng, qp, qq, qk, nthreads=threads,
                             native=native)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    return reps * qp.size / dt, dt, reps, vals


def bank_leg(args):
    """CI / CCSD B-spline products: 120 B-splines of order 8 on an exponential
    grid to 50 a.u. (r_min 1e-5), 8 Gauss nodes, band pairs B_i B_{i+s}
    (s < 8, 932 pairs), the pair x pair R^k matrix for k = 0..6 in both
    precision modes; per-k error of split against FP64 (whose parity with
    the oracle the tests establish)."""
    import numpy as np
    import torch

    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    from atomx_amd.workloads import bspline_product_workload

    kmax = 6
    gr, grw, B, pa, pb = bspline_product_workload(ns=120, ks=8, r_min=1e-5, r_max=50.0)
    nv, ng = gr.shape
    Bd = dv.f64(B)
    stream = torch.cuda.current_stream()
    res, outs, plans = {}, {}, {}
    for prec in ("fp64", "split"):
        pl = SlaterTensor(gr, grw, None, kmax, precision=prec, pairs=(pa, pb), n_fn=B.shape[0])
        o = dv.empty((pl.size,))
        order, split, _ = pl.tune(Bd, None, o)
        for _ in range(args.warmup):
            pl.run(Bd, None, out=o)
        ms = time_steps(lambda: pl.run(Bd, None, out=o), args.bank_steps, stream)
        res[prec] = {"ms_per_step": ms, "rk_per_s": pl.size / (ms * 1e-3),
                     "schedule": {"order": order, "k_split": split}}
        if prec == "split":
            res[prec]["slices"] = pl.slices
        outs[prec] = dv.host(o)
        plans[prec] = pl
    pl = plans["fp64"]
    per_k = []
    for k in range(kmax + 1):
        a, b = pl.offsets[k], pl.offsets[k] + pl.m[k] * (pl.m[k] + 1) // 2
        f, sp = outs["fp64"][a:b], outs["split"][a:b]
        big = float(np.max(np.abs(f)))
        mask = np.abs(f) >= 1e-3 * big
        per_k.append({"k": k, "split_max_err_over_max": float(np.max(np.abs(sp - f)) / big),
                      "split_max_rel_err_where_abs_ge_1e-3_max":
                          float(np.max(np.abs(sp[mask] - f[mask]) / np.abs(f[mask])))})
    for p_ in plans.values():
        p_.close()
    return {"workload": "120 B-splines (order 8, exponential grid r 1e-5..50, 113 x 8 nodes), "
                        "932 band pairs B_i B_{i+s} (s<8), pair x pair R^k for k=0..6",
            "n_rk": int(pl.size), "modes": res, "split_error_vs_fp64_per_k": per_k}


def ions_leg(args):
    """Batched uneven SCF iteration: the direct + exchange Fock build (DBSR
    cell algorithm, FP64) of every target of all 26 iron charge states
    (ns ~ U{40..160}, order 7-8, 1-10 orbitals, k <= 6), the Fock engine's
    five launches replayed from a captured HIP graph; eager, and the older
    per-(ion, block, k) launch path, for comparison."""
    import torch

    from atomx_amd.batch_scf import IonBatch
    from atomx_amd.workloads import fe_ion_workload

    ions = fe_ion_workload(seed=args.seed)
    batch = IonBatch(ions).capture()
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        batch.replay()
    ms_graph = time_steps(batch.replay, args.ion_steps, stream)
    ms_eager = time_steps(batch.run_eager, max(3, args.ion_steps // 2), stream)
    launches = IonBatch(ions, path="launches").capture()
    launches.replay()
    ms_launches = time_steps(launches.replay, max(3, args.ion_steps // 4), stream)
    n_eq = batch.rk_equivalents
    return {"workload": "26 Fe ions (charge 0..25), ns~U{40..160}, ks~U{7,8} (kinetic balance), "
                        "ground-configuration subshells (1-10 orbitals), k<=6, all targets",
            "ns": [int(i["bsp_data"]["ns"]) for i in ions],
            "orbitals": [len(i["kappas"]) for i in ions],
            "rk_equivalents_per_iteration": n_eq,
            "ms_per_iteration_graph": ms_graph, "ms_per_iteration_eager": ms_eager,
            "ms_per_iteration_per_block_launch_path_graph": ms_launches,
            "engine_plan_device_bytes": batch.engine.nbytes,
            "rk_equivalents_per_s": n_eq / (ms_graph * 1e-3), "dtype": "f64"}


def ion_scf_leg(args):
    """The 26 iron charge states solved: independent Gauss-Seidel Dirac-Fock
    SCFs from the bare nucleus to an orbital change below 1e-9, sweep s of
    every ion together (stage i = the Fock blocks of orbital i of every ion
    in one engine run on the GPU, then the ions' FP64 eigensolves in
    parallel host threads)."""
    import json as _json

    from atomx_amd.batch_solve import BatchDiracFock, fe_ion_subshells
    from atomx_amd.workloads import fe_ion_workload

    ions = fe_ion_workload(seed=args.seed)
    batch = BatchDiracFock([(26, fe_ion_subshells(c), ion["bsp_data"])
                            for c, ion in enumerate(ions)])
    try:
        t0 = time.perf_counter()
        e = batch.run(tol=1e-9)
        dt = time.perf_counter() - t0
        sweeps = [s.sweeps for s in batch.systems]
        res = {"workload": "26 Fe ions (charge 0..25) of the Fock-build leg, Z=26, point nucleus, "
                           "bare-nucleus start, tol 1e-9 on the orbitals",
               "E_total_Ha": [float(x) for x in e], "sweeps_per_ion": sweeps,
               "sweeps": batch.sweeps, "seconds": dt,
               "ms_per_sweep": float(np.mean(batch.sweep_ms)),
               "gpu_fock_ms_per_sweep": float(np.mean(batch.gpu_ms)),
               "host_copy_ms_per_sweep": float(np.mean(batch.copy_ms)),
               "host_eigensolve_ms_per_sweep": float(np.mean(batch.solve_ms)),
               "eigensolve_threads": batch.threads,
               "fock_rk_equivalents_per_sweep": batch.rk_equivalents_per_sweep}
    finally:
        batch.close()
    gold = os.path.join(ROOT, "tests", "golden", "fe_ion_scf_oracle.json")
    if args.seed == 0 and os.path.exists(gold):  # the oracle SCF energies of this workload
        ref = _json.load(open(gold))["ions"]
        res["max_abs_dE_vs_oracle_scf_Ha"] = max(abs(x - ref[str(c)]["E_total_Ha"])
                                                  for c, x in enumerate(res["E_total_Ha"]))
    return res


def neon_leg(args):
    """Ne (Z = 10) Dirac-Fock ground state 1s2 2s2 2p6 on 60 B-splines of
    order 7 (exponential knots, r 1e-5..40, 7 Gauss points per interval,
    point nucleus), from the bare-nucleus orbitals to convergence: the
    two-electron Fock blocks and energies on the GPU, the one-body matrices
    and the FP64 generalised eigensolves on the host."""
    import torch

    from atomx_amd import bspline as B
    from atomx_amd import dirac as D

    t = B.make_knots(60, 7, 1e-5, 40.0)
    bd = B.prepare_bspline_data(t, 60, 7)
    scf = D.DiracFockSCF(10, D.neon_subshells(), bd).bare_nucleus_start()
    for i in range(len(scf.shells)):  # set-up outside the timed solve: plans, energy tables
        scf.backend.plan(i).run()
    scf.energy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = scf.run()
    dt = time.perf_counter() - t0
    n_eq = scf.backend.rk_equivalents()
    return {"workload": "Ne 1s2 2s2 2p6, 60 B-splines order 7, r 1e-5..40, point nucleus",
            "E_total_Ha": e, "sweeps": scf.iterations, "seconds": dt,
            "ms_per_sweep": dt / scf.iterations * 1e3,
            "orbital_uploads": scf.backend.uploads,
            "fock_rk_equivalents_per_sweep": n_eq, "_bd": bd, "_scf": scf}


def correlation_leg(neon):
    """Consumers of the Slater integrals in correlation methods, on the
    converged neon ground state of the neon leg (60 B-splines): MBPT(2) over
    every physical V^N virtual with l <= 4 (kappa_max = 4, no caps), helium's
    MBPT(2) on 80 B-splines against the reference's published value, and the
    CC/FSCC signed spinor-integral export (1s..2p occupied; 3s 4s 3p 4p 3d
    4d 4f virtuals from the bare-nucleus complement: 60 spinors, packed).
    Each is timed on its second call (the first builds grid banks and
    library handles)."""
    import torch

    from atomx_amd import bspline as B
    from atomx_amd import dirac as D
    from atomx_amd.mbpt import mbpt2_energy
    from atomx_amd.spinor_integrals import build_spinor_integrals

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    scf, bd = neon["_scf"], neon["_bd"]
    ne, dt_ne = timed(lambda: mbpt2_energy(scf, kappa_max=4))
    t = B.make_knots(80, 7, 1e-5, 40.0)
    bd_he = B.prepare_bspline_data(t, 80, 7)
    he_scf = D.DiracFockSCF(2, [D.Subshell("1s", 1, -1, 2.0)], bd_he).bare_nucleus_start()
    e_he = he_scf.run(tol=1e-9)
    he, dt_he = timed(lambda: mbpt2_energy(he_scf, kappa_max=4))
    orbitals = {(sh.n, sh.kappa): (scf.P[i], scf.Q[i]) for i, sh in enumerate(scf.shells)}
    vir = []
    for kappa, shells in ((-1, (3, 4)), (1, (3, 4)), (-2, (3, 4)), (2, (3, 4)), (-3, (3, 4)),
                          (3, (4,)), (-4, (4,))):
        for (P, Q, _), n in zip(scf.virtual_orbitals(kappa, len(shells)), shells):
            orbitals[n, kappa] = (P, Q)
            vir.append((n, kappa))
    occ = [(1, -1), (2, -1), (2, 1), (2, -2)]
    si, dt_si = timed(lambda: build_spinor_integrals(orbitals, bd, 10, bd["ns"],
                                                     occupied_keys=occ, virtual_keys=vir))
    return {"mbpt2_neon": {"workload": "Ne 1s2 2s2 2p6 (60 B-splines), kappa_max 4, every "
                                       "physical V^N virtual",
                           "E2_Ha": ne.correlation_energy, "virtuals": sum(ne.n_virtuals.values()),
                           "occupied_pairs": len(ne.pair_contributions), "seconds": dt_ne},
            "mbpt2_helium": {"workload": "He 1s2 (80 B-splines order 7, r 1e-5..40), kappa_max 4, "
                                         "every physical V^N virtual",
                             "E_DF_Ha": e_he, "E2_Ha": he.correlation_energy,
                             "E2_reference_published_Ha": -0.036968,
                             "virtuals": sum(he.n_virtuals.values()), "seconds": dt_he},
            "spinor_export": {"workload": "Ne, 1s..2p occupied, 3s 4s 3p 4p 3d 4d 4f virtuals "
                                          "(bare-nucleus complement), Dirac-Coulomb, packed",
                              "spinors": len(si.labels), "packed_values": int(si.packed_eri_Ha.size),
                              "E_reference_Ha": si.reference_energy_Ha,
                              "abs_diff_vs_E_DF_Ha": abs(si.reference_energy_Ha - neon["E_total_Ha"]),
                              "seconds": dt_si}}


TRAFFIC_TABLE = "profiles/c2_hbm_traffic.txt"


def hbm_traffic(args, kernel, order, k_split):
    """HBM bytes per launch of the contraction kernel at this schedule, from
    the calibrated FETCH_SIZE + WRITE_SIZE table measured on the default
    workload for every schedule the tuner can choose (None for other
    workloads, or a kernel the table does not hold)."""
    from atomx_amd.tensor import SlaterTensor
    default = (args.nv, args.ng, args.n_orb, args.kmax, args.seed) == (400, 8, 40, 8, 0)
    path = os.path.join(ROOT, TRAFFIC_TABLE)
    if not default or not os.path.exists(path):
        return None
    key = f"{SlaterTensor.ORDERS.index(order)}:{int(k_split)}"
    for line in open(path):
        f = line.split()
        if len(f) >= 6 and f[0] == "traffic:" and f[1] == kernel and f[2] == key:
            return int(f[5])
    return None


def serving_leg(args):
    """Poisson stream of Slater-integral and Fock-build requests through the
    server (split precision, continuous batching): p50/p99 latency and R^k/s;
    then the same Slater requests submitted at once (saturated batched
    throughput) and one at a time in plain FP64 (create, run, copy back,
    destroy per request), whose results also give each request's error."""
    import time

    import numpy as np

    from atomx_amd import device as dv
    from atomx_amd.serve import SlaterServer, percentile
    from atomx_amd.tensor import SlaterTensor
    from atomx_amd.workloads import count_rk, serving_stream

    stream = serving_stream(args.serve_requests, rate=args.serve_rate, seed=args.seed)
    warm = serving_stream(24, rate=1e9, seed=args.seed + 1)
    slater = [r for _, r in stream if r.kind == "slater"]
    n_rk = {id(r): count_rk(r.ls, r.kmax) for r in slater}
    # objects alive before the served passes (inputs, earlier legs) leave the
    # collector's generations, as a serving process does with its long-lived
    # state: the requests' own objects then do not trigger rescans of them
    # (served and one-at-a-time passes alike)
    import gc
    gc.collect()
    gc.freeze()
    with SlaterServer(max_inflight=4, max_batch=16) as srv:
        for f in [srv.submit(r) for _, r in warm]:
            f.result()
        srv.reset_stats()
        t0 = time.perf_counter()
        futs = []
        for t_arr, req in stream:
            delay = t0 + t_arr - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futs.append(srv.submit(req))
        res = [f.result() for f in futs]
        t_open = time.perf_counter() - t0
        st_open = srv.stats()
        # saturated: every Slater request submitted at once, six passes; the
        # first grows the server's pinned and device staging to these launch
        # sets (reported as the cold rate), the median of the others is timed
        t_sat_all = []
        for _ in range(6):
            srv.reset_stats()
            t0 = time.perf_counter()
            for f in [srv.submit(r) for r in slater]:
                f.result()
            t_sat_all.append(time.perf_counter() - t0)
        st_sat = srv.stats()
        t_sat_cold, t_sat = t_sat_all[0], float(np.median(t_sat_all[1:]))
    # the same saturated passes through the native serving loop (ax_server_*)
    from atomx_amd.serve_native import NativeSlaterServer
    t_nat = []
    with NativeSlaterServer(max_inflight=4, max_batch=16) as nsrv:
        for i in range(6):  # the first pass grows the server's staging (cold)
            nres = None
            if i == 1:
                st_cold = nsrv.stats()
            t0 = time.perf_counter()
            tickets = [nsrv.submit(r) for r in slater]
            nres = [nsrv.result(t) for t in tickets]
            t_nat.append(time.perf_counter() - t0)
        st_nat = nsrv.stats()
    host_ms = ("setup_ms", "pack_wait_ms", "alloc_ms", "queue_ms")
    by_req = {id(req): r for (_, req), r in zip(stream, res)}
    nat_diff = max(float(np.max(np.abs(x.values - by_req[id(q)].values)))
                   for q, x in zip(slater, nres))
    del nres
    lat = {"slater": [], "fock": []}
    for (_, req), r in zip(stream, res):
        lat[req.kind].append(r.latency_s * 1e3)
    # plain per-request FP64 on the GPU, one request at a time
    for r in slater[:8]:  # warm-up of the same path
        pl = SlaterTensor(r.gr, r.grw, r.ls, r.kmax, precision="fp64")
        dv.host(pl.run(r.P, r.Q))
        pl.close()
    fp64 = {}
    t0 = time.perf_counter()
    for r in slater:
        pl = SlaterTensor(r.gr, r.grw, r.ls, r.kmax, precision="fp64")
        fp64[id(r)] = dv.host(pl.run(r.P, r.Q))
        pl.close()
    t_seq = time.perf_counter() - t0
    worst_norm = worst_elem = 0.0
    for (_, req), r in zip(stream, res):
        if req.kind != "slater":
            continue
        ref = fp64[id(req)]
        big = np.max(np.abs(ref))
        worst_norm = max(worst_norm, float(np.max(np.abs(r.values - ref)) / big))
        mask = np.abs(ref) >= 1e-3 * big
        worst_elem = max(worst_elem, float(np.max(np.abs(r.values[mask] - ref[mask])
                                                  / np.abs(ref[mask]))))
    total_rk = sum(n_rk.values())
    all_lat = lat["slater"] + lat["fock"]
    gc.unfreeze()
    scf = scf_serving(args, slater)
    return {
        "workload": (f"{len(stream)} requests, Poisson arrivals at {args.serve_rate:g}/s (seed "
                     f"{args.seed}); n_orb~logU[5,60], N~logU[1000,6000] (8 Gauss nodes per "
                     f"cell), k_max~logU[2,10]; {len(slater)} all-R^k requests (split "
                     f"precision), {len(stream) - len(slater)} Y^k Fock-build requests"),
        "p50_ms": percentile(all_lat, 50), "p99_ms": percentile(all_lat, 99),
        "p50_ms_slater": percentile(lat["slater"], 50),
        "p99_ms_slater": percentile(lat["slater"], 99),
        "p50_ms_fock": percentile(lat["fock"], 50) if lat["fock"] else None,
        "p99_ms_fock": percentile(lat["fock"], 99) if lat["fock"] else None,
        "open_loop_s": t_open, "rk_total": total_rk,
        "rk_per_s_open_loop": total_rk / t_open,
        "launches_open_loop": st_open["launches"], "peak_inflight": st_open["peak_inflight"],
        "peak_device_bytes": st_open["peak_bytes"],
        "saturated_rk_per_s": total_rk / t_sat, "saturated_launches": st_sat["launches"],
        "saturated_cold_rk_per_s": total_rk / t_sat_cold,
        "native_saturated_rk_per_s": total_rk / float(np.median(t_nat[1:])),
        "native_saturated_cold_rk_per_s": total_rk / t_nat[0],
        "native_saturated_vs_sequential_fp64": t_seq / float(np.median(t_nat[1:])),
        "native_launches": st_nat["launches"] // len(t_nat),
        "native_pass_ms": [round(1e3 * t, 2) for t in t_nat],
        "native_host_ms_cold_pass": {k: round(st_cold[k], 2) for k in host_ms},
        "native_host_ms_per_warm_pass": {k: round((st_nat[k] - st_cold[k]) / (len(t_nat) - 1), 2)
                                         for k in host_ms},
        "native_peak_device_bytes": st_nat["peak_bytes"],
        "saturated_pass_ms": [round(1e3 * t, 2) for t in t_sat_all],
        "native_vs_python_server_max_abs_diff": nat_diff,
        "saturated_cold_vs_sequential_fp64": t_seq / t_sat_cold,
        "split_requests_sliced": sum(1 for (_, q), r in zip(stream, res)
                                     if q.kind == "slater" and r.slices > 0),
        "split_requests_fp64_fallback": sum(1 for (_, q), r in zip(stream, res)
                                            if q.kind == "slater" and r.slices == 0),
        "sequential_fp64_rk_per_s": total_rk / t_seq,
        "saturated_vs_sequential_fp64": t_seq / t_sat,
        "max_err_vs_fp64_over_max": worst_norm,
        "max_rel_err_vs_fp64_where_abs_ge_1e-3_max": worst_elem,
        "scf_serving": scf,
    }


def _plain_per_atom(atoms, reqs):
    """Plain per-atom FP64: each atom's plan created beforehand, then per
    request: pinned staging, upload, the plan's launches, copy back, wait.
    Returns the seconds for ``reqs`` (request r belongs to atom r % len)."""
    import time

    import torch

    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor

    plans = [SlaterTensor(r.gr, r.grw, r.ls, r.kmax, precision="fp64") for r in atoms]
    dev = [(dv.empty(r.P.shape), dv.empty(r.Q.shape), dv.empty((p.size,)))
           for r, p in zip(atoms, plans)]
    width = max(2 * r.P.size for r in atoms) + max(p.size for p in plans)
    stage = torch.empty(width, dtype=torch.float64, pin_memory=True)
    s_np = stage.numpy()

    def plain(a, req):
        P, Q, out = dev[a]
        k = req.P.size
        s_np[:k] = req.P.ravel()
        s_np[k:2 * k] = req.Q.ravel()
        P.copy_(stage[:k].view(P.shape), non_blocking=True)
        Q.copy_(stage[k:2 * k].view(Q.shape), non_blocking=True)
        plans[a].run(P, Q, out)
        res = stage[2 * k:2 * k + out.numel()]
        res.copy_(out, non_blocking=True)
        torch.cuda.synchronize()
        return res.numpy().copy()

    for a, r in enumerate(atoms):  # warm-up
        plain(a, r)
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        plain(i % len(atoms), r)
    dt = time.perf_counter() - t0
    for p in plans:
        p.close()
    return dt


def _scf_requests(atoms, iterations, precision, seed):
    import numpy as np

    from atomx_amd.serve import SlaterRequest
    rng = np.random.default_rng(seed)
    out = []
    for it in range(iterations):
        f = 1.0 + 1e-3 * it * rng.standard_normal()
        for a, r in enumerate(atoms):
            out.append(SlaterRequest(r.gr, r.grw, r.P * f, r.Q * f, r.ls, r.kmax, precision,
                                     state=("atom", a)))
    return out


def _served(reqs, warm):
    """Seconds to serve ``reqs`` submitted at once (after ``warm`` untimed).
    The objects alive before the timed part (inputs, earlier legs) are moved
    out of the collector's generations first, as a serving process does with
    its long-lived state, so that collections triggered by the requests' own
    objects do not rescan them."""
    import gc
    import time

    from atomx_amd.serve import SlaterServer
    with SlaterServer(max_inflight=4, max_batch=16) as srv:
        for f in [srv.submit(r) for r in warm]:
            f.result()
        gc.collect()
        gc.freeze()
        srv.reset_stats()
        t0 = time.perf_counter()
        for f in [srv.submit(r) for r in reqs]:
            f.result()
        dt = time.perf_counter() - t0
        st = srv.stats()
        st["plans_created"] = srv.stats()["plans_created"]
    gc.unfreeze()
    return dt, st


def _served_native(reqs, warm):
    """Seconds to serve ``reqs`` submitted at once through the native
    serving loop with resident states (after ``warm`` untimed, which makes
    the states' plans); the same collector treatment as ``_served``."""
    import gc
    import time

    from atomx_amd.serve_native import NativeSlaterServer
    with NativeSlaterServer(max_inflight=4, max_batch=16) as srv:
        for t in [srv.submit(r) for r in warm]:
            srv.result(t)
        gc.collect()
        gc.freeze()
        t0 = time.perf_counter()
        res = [srv.result(t) for t in [srv.submit(r) for r in reqs]]
        dt = time.perf_counter() - t0
        st = srv.stats()
        del res
    gc.unfreeze()
    return dt, st


def scf_serving(args, slater, iterations=6):
    """SCF-style serving: every atom sends one all-R^k request per SCF
    iteration (new orbitals, same grid and shells) under a resident state;
    iteration 0 creates the states untimed, iterations 1.. are submitted at
    once.  Compared with the plain per-atom FP64 form (``_plain_per_atom``).
    Two atom sets: 48 atoms of the request mix (large: 1000-6000 nodes, up
    to 60 orbitals, GPU-bound per atom) and 256 small atoms and ions (40-160
    cells, 1-10 orbitals: launch-bound per atom), the latter served both as
    one request per atom and as one request per group of 32 atoms per
    iteration (``SlaterBatchRequest``); per-atom requests also through the
    native serving loop with its resident states."""
    from atomx_amd.serve import SlaterBatchRequest
    from atomx_amd.workloads import count_rk, small_atom_requests

    out = {}
    for name, atoms in (("mix48", slater[:48]),
                        ("small256", small_atom_requests(256, seed=args.seed + 11))):
        reqs = _scf_requests(atoms, iterations, "split", args.seed + 7)
        n = len(atoms)
        n_rk = sum(count_rk(r.ls, r.kmax) for r in reqs[n:])
        t_req, st = _served(reqs[n:], reqs[:n])
        t_nat, st_nat = _served_native(reqs[n:], reqs[:n])
        t_plain = _plain_per_atom(atoms, reqs[n:])
        row = {
            "atoms": n, "timed_iterations": iterations - 1, "rk_total": n_rk,
            "plain_per_atom_fp64_rk_per_s": n_rk / t_plain,
            "served_per_atom_requests_rk_per_s": n_rk / t_req,
            "served_per_atom_requests_vs_plain": t_plain / t_req,
            "served_per_atom_p50_ms": st["p50_ms"], "served_per_atom_p99_ms": st["p99_ms"],
            "native_served_per_atom_rk_per_s": n_rk / t_nat,
            "native_served_per_atom_vs_plain": t_plain / t_nat,
            "native_plans_made": st_nat["plans_made"],
        }
        if name == "small256":
            group = 32
            breqs = [SlaterBatchRequest(reqs[i:i + group]) for i in range(0, len(reqs), group)]
            n_warm = n // group
            t_grp, st_g = _served(breqs[n_warm:], breqs[:n_warm])
            row.update({
                "served_group32_requests_rk_per_s": n_rk / t_grp,
                "served_group32_requests_vs_plain": t_plain / t_grp,
                "served_group32_p50_ms": st_g["p50_ms"], "served_group32_p99_ms": st_g["p99_ms"],
            })
        out[name] = row
    return out


def headline(args):
    """The benchmark tensor in both modes; returns (line, context of the CPU leg)."""
    import torch
    from atomx_amd import device as dv
    from atomx_amd.tensor import SlaterTensor
    from atomx_amd.workloads import count_rk, synthetic_grid, synthetic_orbitals

    gr, grw = synthetic_grid(nv=args.nv, ng=args.ng, r_min=1e-5, r_max=60.0)
    P, Q, ls = synthetic_orbitals(gr, grw, n_orb=args.n_orb, seed=args.seed)
    n_rk = count_rk(ls, args.kmax)
    n = gr.size
    Pd, Qd = dv.f64(P), dv.f64(Q)
    stream = torch.cuda.current_stream()

    modes = {}
    plans = {}
    host_out = {}
    for prec in ("fp64", "split"):
        plan = SlaterTensor(gr, grw, ls, args.kmax, precision=prec)
        out = dv.empty((plan.size,))
        order, split, tuned_ms = plan.tune(Pd, Qd, out)  # schedule autotuner (untimed set-up)
        torch.cuda.synchronize()
        host_out[prec] = dv.host(out)
        plans[prec] = (plan, out)
        modes[prec] = {"schedule": {"order": order, "k_split": split,
                                    "contract_reduce_ms": tuned_ms}}
        if prec == "split":
            modes[prec]["slices"] = plan.slices
    p0 = plans["fp64"][0]
    norm, el = agreement(host_out["split"], host_out["fp64"], p0.offsets, p0.m)
    modes["split"]["vs_fp64_max_err_over_max"] = norm
    modes["split"]["vs_fp64_max_rel_err"] = el

    # whole-step timing, and both modes for the record
    for name, (pl, o) in plans.items():
        for _ in range(args.warmup):
            pl.run(Pd, Qd, out=o)
        ms = time_steps(lambda: pl.run(Pd, Qd, out=o), args.steps, stream)
        modes[name]["ms_per_step"] = ms
        modes[name]["rk_per_s"] = n_rk / (ms * 1e-3)

    # the headline is the faster mode among those meeting the accuracy bar
    ok = ["fp64"] + (["split"] if modes["split"]["vs_fp64_max_rel_err"] <= 1e-11 else [])
    if args.precision == "auto":
        prec = max(ok, key=lambda m: modes[m]["rk_per_s"])
    else:
        prec = args.precision
    plan, out = plans[prec]

    # per-stage timing (HIP events on the launch stream)
    stages = {}
    for st in SlaterTensor.STAGES:
        stages[st] = time_steps(lambda: plan.run_stage(st, Pd, Qd, out), args.steps, stream)
    contract_ms = stages["contract"]
    flops = 2.0 * n * n_rk
    if prec == "fp64":
        peak = FP64_MFMA_PEAK
    else:
        peak = INT8_MFMA_PEAK / (plan.slices * (plan.slices + 1) // 2)
    achieved = flops / (contract_ms * 1e-3) / 1e12
    kernel = "rk_gemm_f64_kernel" if prec == "fp64" else f"rk_gemm_split_kernel<{plan.slices}>"
    k_split = modes[prec]["schedule"]["k_split"]
    order = modes[prec]["schedule"]["order"]
    roofline = {"bound": "mfma", "kernel": kernel, "achieved": achieved,
                "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": hbm_traffic(args, kernel, order, k_split),
                "traffic_source": TRAFFIC_TABLE, "order": order, "k_split": k_split,
                "algorithmic_flops_per_launch": flops, "avg_launch_ms": contract_ms}

    ms = modes[prec]["ms_per_step"]
    value = n_rk / (ms * 1e-3)
    line = {
        "metric": METRIC,
        "value": value, "unit": "R^k/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == "fp64" else "int8-slices/f64",
        "data": "synthetic",
        "config": {"workload": f"all-quartet R^k tensor: {args.n_orb} synthetic orbitals, "
                               f"{args.nv}x{args.ng} = {n} Gauss nodes (r 1e-5..60), k=0..{args.kmax}",
                   "n_rk": n_rk, "precision": prec, "parallelism": "single-gpu"},
        "accuracy": modes[prec].get("vs_fp64_max_rel_err", 0.0),
        "modes": modes, "stage_ms": stages, "roofline": roofline,
    }
    ctx = {"P": P, "Q": Q, "gr": gr, "grw": grw, "ls": ls, "n_rk": n_rk, "host_out": host_out,
           "p0": p0, "modes": modes, "prec": prec, "plans": plans}
    return line, ctx


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    if args.no_tensor:  # the other legs alone (profiling one leg)
        line = {"metric": METRIC, "value": None, "unit": "R^k/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                "data": "synthetic", "config": {"workload": "selected legs only"}}
        ctx = None
    else:
        line, ctx = headline(args)
        progress("benchmark tensor")
    if args.bank_steps > 0:
        line["bspline_products"] = bank_leg(args)
        progress("B-spline products")
    if args.ion_steps > 0:
        line["fe_ions"] = ions_leg(args)
        progress("iron-ion Fock builds")
    if not args.no_ion_scf:
        line["fe_ions_scf"] = ion_scf_leg(args)
        progress("iron-ion SCF")
    neon = None
    if not args.no_neon:
        neon = neon_leg(args)
        line["neon_dirac_fock"] = {k: v for k, v in neon.items() if not k.startswith("_")}
        progress("neon SCF")
        if not args.no_correlation:
            line["correlation"] = correlation_leg(neon)
            progress("MBPT(2) and spinor export")
    if args.serve_requests > 0:
        serving = serving_leg(args)
        progress("serving")
        line["serving"] = serving
        line["latency_p50_ms"] = serving["p50_ms"]
        line["latency_p99_ms"] = serving["p99_ms"]
    if not args.no_cpu and ctx is not None:
        import oracle as O
        P, Q, gr, grw, ls = (ctx[k] for k in ("P", "Q", "gr", "grw", "ls"))
        n_rk, host_out, p0, modes, prec = (ctx[k] for k in ("n_rk", "host_out", "p0", "modes",
                                                            "prec"))
        host = O.host_cpu()
        threads = args.cpu_threads or max(1, min(16, host["affinity_cpus"]))
        # the checker: the strict oracle build, once; then the timed native build
        _, _, _, ref = cpu_baseline(P, Q, gr, grw, ls, args.kmax, threads, min_seconds=0.0,
                                    native=False)
        v, dt, reps, nat = cpu_baseline(P, Q, gr, grw, ls, args.kmax, threads)
        v1, dt1, reps1, _ = cpu_baseline(P, Q, gr, grw, ls, args.kmax, 1, min_seconds=4.0)
        for m in ("fp64", "split"):  # the oracle's values check every GPU output
            norm, el = agreement(host_out[m], ref, p0.offsets, p0.m)
            modes[m]["vs_oracle_max_err_over_max"] = norm
            modes[m]["vs_oracle_max_rel_err"] = el
        line["accuracy"] = modes[prec]["vs_oracle_max_rel_err"]
        line["accuracy_reference"] = "CPU FP64 oracle, all values"
        if neon is not None:  # the same SCF with the oracle's two-electron restatement
            e_cpu, sweeps = neon_oracle_scf(neon["_bd"])
            line["neon_dirac_fock"]["E_total_oracle_scf_Ha"] = e_cpu
            line["neon_dirac_fock"]["abs_diff_vs_oracle_scf_Ha"] = abs(neon["E_total_Ha"] - e_cpu)
        line["cpu_baseline"] = {
            "value": v, "unit": "R^k/s", "cores": threads, "kind": "port",
            "sample": f"full workload ({n_rk} R^k) repeated {reps}x in {dt:.1f} s with {threads} "
                      f"OpenMP threads",
            "build": "oracle/slater_oracle.c, gcc -O2 -march=native -fno-fast-math -fopenmp, "
                     "built on this host",
            "native_vs_checker_max_err_over_max": agreement(nat, ref, p0.offsets, p0.m)[0],
            "single_thread": {"value": v1, "sample": f"{reps1}x in {dt1:.1f} s"},
            "host": host}
        progress("CPU baseline")
    print(json.dumps(line), flush=True)
    if ctx is not None:
        for pl, _ in ctx["plans"].values():
            pl.close()
    if args.copy_mib > 0:
        import torch
        src = torch.ones(args.copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy calibration: {args.copy_mib} MiB read + {args.copy_mib} MiB written", flush=True)


if __name__ == "__main__":
    main()
