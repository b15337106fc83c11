"""GPU parity of the banded B-spline assembly (integrals.py:18-56) and of the
Y^k-form Fock pieces against the CPU oracle (1e-13 / 1e-12 relative to the
largest entry; the arithmetic is the same, only FMA contraction differs)."""

import numpy as np
import pytest

import oracle as O
from helpers import physics_workspace

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kb", [False, True])
def test_assemble_integral_matches_oracle(kb):
    from atomx_amd import device as dv
    from atomx_amd._lib import call
    bd, t = physics_workspace(Z=3, ks=8, r_max=30.0, kinetic_balance=kb)
    nv, ns = bd["nv"], bd["ns"]
    gr, grw = bd["gr"], bd["grw"]
    bp, bq = bd["bsp_p"], bd["bsp_q"]
    ksp, ksq = bd["ksp"], bd["ksq"]
    wts = np.stack([grw * np.cos(gr) / (1.0 + gr), grw * gr ** 2 * np.exp(-gr)])
    out = dv.empty((2, ns, ns))
    bpd, bqd, wd = dv.f64(bp), dv.f64(bq), dv.f64(wts)  # alive until the launch is queued
    call("ax_assemble_integral", nv, gr.shape[1], ksp, ksq, ns, dv.ptr(bpd), dv.ptr(bqd),
         dv.ptr(wd), 2, dv.ptr(out), dv.stream_handle())
    got = dv.host(out)
    for w in range(2):
        ref = O.assemble_integral(nv, ksp, ksq, ns, bp, bq, wts[w])
        assert np.max(np.abs(got[w] - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_fock_yk_direct_and_exchange_match_oracle():
    from atomx_amd import device as dv
    from atomx_amd.fock_yk import FockYk
    bd, t = physics_workspace(Z=4, ks=7, r_max=35.0, kinetic_balance=True)
    ns = bd["ns"]
    rng = np.random.default_rng(3)
    decay = np.exp(-np.arange(ns) / 5.0)
    P = rng.normal(size=(3, ns)) * decay
    Q = 0.02 * rng.normal(size=(3, ns)) * decay
    occ = [2.0, 1.0, 3.0]
    fk = FockYk(bd)
    fP, fQ = fk.samples(P, Q)
    vpp, vqq = (dv.host(x) for x in fk.direct(fP, fQ, occ, [0, 1, 2]))
    for tt in range(3):
        rpp, rqq = O.direct_potential_yk(P, Q, occ, bd, target=tt)
        assert np.max(np.abs(vpp[tt] - rpp)) <= 1e-12 * np.max(np.abs(rpp))
        assert np.max(np.abs(vqq[tt] - rqq)) <= 1e-12 * np.max(np.abs(rqq))
        assert np.array_equal(vpp[tt], vpp[tt].T)
    for b, k in ((0, 0), (1, 2), (2, 1)):
        kp, kq = (dv.host(x) for x in fk.exchange_matrix(fP, fQ, b, k))
        rp, rq = O.compute_exchange_matrix(P[b], Q[b], bd, k)
        assert np.max(np.abs(kp - rp)) <= 1e-12 * np.max(np.abs(rp))
        assert np.max(np.abs(kq - rq)) <= 1e-12 * np.max(np.abs(rq))
