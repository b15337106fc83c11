"""GPU: the orbital-mode Fock blocks (direct potential, per-source cached
exchange, intra-shell potential) against the oracle's restatement, the
cache invalidation of an updated orbital, and the orbital-mode neon SCF
against the DBSR-mode SCF (same closed-shell fixed point)."""

import numpy as np
import pytest

import oracle as O
from oracle import orbital_mode as OM
from helpers import dbsr_state

pytestmark = pytest.mark.gpu


def _close(a, b, tol=1e-12):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) <= tol * scale


def test_neon_fixture_blocks_match_oracle():
    from atomx_amd.orbital_mode import OrbitalModeFock
    bd, P, Q, occ, kap, ref = dbsr_state()
    banks = O.Banks(bd)
    om = OrbitalModeFock(bd, kap, occ)
    om.load(P, Q)
    for t in range(len(kap)):
        print(f"target {t}", flush=True)
        got = om.blocks(t)
        want = OM.orbital_fock_blocks(t, P, Q, occ, kap, bd, banks)
        for g, w in zip(got, want):
            assert _close(g, w), t


def test_open_shell_ion_blocks_and_update():
    from atomx_amd.orbital_mode import OrbitalModeFock
    from atomx_amd.workloads import fe_ion_workload
    ions = fe_ion_workload(seed=3, ns_range=(40, 60))
    ion = next(i for i in ions if any(abs(k) >= 3 for k in i["kappas"])
               and len(i["kappas"]) >= 6)
    bd, kap, occ = ion["bsp_data"], ion["kappas"], ion["occupations"]
    print(f"Fe{ion['charge']}+ ns={bd['ns']} kappas={kap} occ={occ}", flush=True)
    banks = O.Banks(bd)
    P, Q = ion["P"], ion["Q"]
    om = OrbitalModeFock(bd, kap, occ)
    om.load(P, Q)
    targets = range(len(kap))
    for t in targets:
        got = om.blocks(t)
        want = OM.orbital_fock_blocks(t, P, Q, occ, kap, bd, banks)
        for g, w in zip(got, want):
            assert _close(g, w), t
    # update one orbital: cached contractions of the others are kept, the
    # result equals a fresh build bitwise
    P2, Q2 = P.copy(), Q.copy()
    P2[1] *= 1.01
    Q2[1] *= 0.99
    om.update(1, P2[1], Q2[1])
    fresh = OrbitalModeFock(bd, kap, occ, om.cache)
    fresh.load(P2, Q2)
    for t in targets:
        for g, w in zip(om.blocks(t), fresh.blocks(t)):
            assert np.array_equal(g, w), t


def test_orbital_mode_neon_scf_matches_dbsr_mode():
    from atomx_amd import bspline as B
    from atomx_amd import dirac as D
    bd = B.prepare_bspline_data(B.make_knots(60, 7, 1e-5, 40.0), 60, 7)
    shells = D.neon_subshells()
    e_orb = D.DiracFockSCF(10, shells, bd, mode="orbital").bare_nucleus_start().run()
    print(f"E_total orbital mode {e_orb:.13f}", flush=True)
    assert abs(e_orb - (-128.69196949331734)) < 1e-10
