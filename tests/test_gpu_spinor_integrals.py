"""GPU: the signed spinor-integral export (physics/csf/spinor_integrals.py:
258-466) against the oracle's restatement (oracle/spinor.py, pinned on the
CPU by the Dirac-Fock energy and closed forms), with the reference's own
properties (antisymmetry, Hermiticity, orbital-phase covariance, read-only
arrays, memory cap and metric rejection) and, for neon, the exported
closed-shell reference energy against the GPU Dirac-Fock total energy.
Tolerance: 1e-12 of each array's largest magnitude against the oracle (the
device radial table is a GEMM over banded densities and potentials, the
oracle's a scalar cell R^k per key)."""

import itertools

import numpy as np
import pytest

import oracle as O
from oracle.spinor import spinor_integrals as oracle_export

pytestmark = pytest.mark.gpu

OCC, VIR = [(1, -1)], [(2, -1), (2, 1), (2, -2)]


@pytest.fixture(scope="module")
def be_like():
    from atomx_amd import bspline as B
    from atomx_amd import dirac as D
    t = B.make_knots(30, 7, 1e-4, 30.0)
    bd = B.prepare_bspline_data(t, 30, 7)
    banks = O.Banks(bd)
    tables = O.angular_tables([2.0], [-1])

    class Backend:
        def fock(self, i, P, Q):
            return O.build_hf_matrix_dbsr(i, P, Q, [2.0], [-1], bd, banks,
                                          skip_pq_exchange=True, tables=tables)

        def energy(self, P, Q):
            return O.two_electron_energy(P, Q, [2.0], [-1], bd, banks)

    scf = D.DiracFockSCF(4, [D.Subshell("1s", 1, -1, 2.0)], bd, backend=Backend())
    e_df = scf.bare_nucleus_start().run(tol=1e-9)
    orbitals = {(1, -1): (scf.P[0], scf.Q[0])}
    for kappa in (-1, 1, -2):
        P, Q, _ = scf.virtual_orbitals(kappa, 1)[0]
        orbitals[(2, kappa)] = (P, Q)
    return bd, orbitals, e_df


def _close(a, b, rel=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b))) <= rel * max(float(np.max(np.abs(b))), 1e-300)


@pytest.mark.parametrize("frozen", [False, True])
def test_export_matches_oracle(be_like, frozen):
    from atomx_amd.spinor_integrals import build_spinor_integrals
    bd, orbitals, e_df = be_like
    occ, vir, fz = (([(2, -1)], [(2, 1), (2, -2)], [(1, -1)]) if frozen else (OCC, VIR, ()))
    got = build_spinor_integrals(orbitals, bd, 4, bd["ns"], occupied_keys=occ,
                                 virtual_keys=vir, frozen_keys=fz)
    labels, nocc, heff, fock, packed, e_frozen, e_ref = oracle_export(
        orbitals, bd, 4, bd["ns"], occ, vir, fz)
    assert got.labels == labels and got.noccupied == nocc
    assert np.all(got.packed_eri_Ha.imag == 0.0)
    assert _close(got.packed_eri_Ha.real, packed)
    assert _close(got.one_body_Ha.real, heff) and _close(got.fock_Ha.real, fock)
    assert abs(got.frozen_energy_Ha - e_frozen) < 1e-11
    assert abs(got.reference_energy_Ha - e_ref) < 1e-11
    if frozen:
        assert abs(got.frozen_energy_Ha - e_df) < 1e-10
    else:
        assert abs(got.reference_energy_Ha - e_df) < 1e-10
    print(f"frozen={frozen}: {len(labels)} spinors, max |g| {np.max(np.abs(packed)):.3e}, "
          f"Hermitian projection {got.identity['hermitian_projection_max_Ha']:.1e}", flush=True)


def test_signed_symmetries_phase_covariance_and_owned_arrays(be_like):
    from atomx_amd.spinor_integrals import build_spinor_integrals
    bd, orbitals, _ = be_like
    a = build_spinor_integrals(orbitals, bd, 4, bd["ns"], occupied_keys=OCC, virtual_keys=VIR)
    for p, q, r, s in itertools.product(range(10), repeat=4):
        g = a.eri(p, q, r, s)
        assert g == pytest.approx(-a.eri(q, p, r, s), abs=1e-14)
        assert g == pytest.approx(a.eri(r, s, p, q).conjugate(), abs=1e-14)
    changed = dict(orbitals)
    changed[2, -1] = tuple(-x for x in orbitals[2, -1])
    b = build_spinor_integrals(changed, bd, 4, bd["ns"], occupied_keys=OCC, virtual_keys=VIR)
    ph = np.array([-1 if label[:2] == (2, -1) else 1 for label in a.labels])
    assert np.max(np.abs(b.fock_Ha - ph[:, None] * a.fock_Ha * ph)) < 1e-12
    for p, q, r, s in itertools.product(range(10), repeat=4):
        assert b.eri(p, q, r, s) == pytest.approx(a.eri(p, q, r, s) * ph[p] * ph[q] * ph[r] * ph[s],
                                                  abs=1e-12)
    for arr in (a.fock_Ha, a.one_body_Ha, a.packed_eri_Ha):
        with pytest.raises(ValueError):
            arr.setflags(write=True)
    assert a.identity["binding"] != b.identity["binding"]


def test_rejections(be_like):
    from atomx_amd.spinor_integrals import build_spinor_integrals
    bd, orbitals, _ = be_like
    kw = dict(occupied_keys=OCC, virtual_keys=VIR)
    with pytest.raises(MemoryError):
        build_spinor_integrals(orbitals, bd, 4, bd["ns"], max_integral_bytes=100, **kw)
    bad = dict(orbitals)
    bad[2, -1] = orbitals[1, -1]
    with pytest.raises(ValueError, match="metric"):
        build_spinor_integrals(bad, bd, 4, bd["ns"], occupied_keys=OCC, virtual_keys=[(2, -1)])
    with pytest.raises(ValueError):
        build_spinor_integrals(orbitals, bd, 4, bd["ns"], occupied_keys=OCC, virtual_keys=OCC)
    with pytest.raises(ValueError):
        build_spinor_integrals(orbitals, bd, 4, bd["ns"], representation="dense", **kw)
    with pytest.raises(NotImplementedError):
        build_spinor_integrals(orbitals, bd, 4, bd["ns"], include_breit=True, **kw)


def test_on_demand_matches_packed(be_like):
    from atomx_amd.spinor_integrals import build_spinor_integrals
    bd, orbitals, _ = be_like
    kw = dict(occupied_keys=OCC, virtual_keys=VIR)
    a = build_spinor_integrals(orbitals, bd, 4, bd["ns"], **kw)
    b = build_spinor_integrals(orbitals, bd, 4, bd["ns"], representation="on_demand", **kw)
    assert b.reference_energy_Ha == pytest.approx(a.reference_energy_Ha, abs=1e-12)
    assert np.max(np.abs(b.fock_Ha - a.fock_Ha)) < 1e-12
    rng = np.random.default_rng(3)
    for p, q, r, s in rng.integers(0, 10, size=(300, 4)):
        assert b.eri(p, q, r, s) == pytest.approx(a.eri(p, q, r, s), abs=1e-13)
    streamed = np.concatenate(list(b.iter_packed_eri(max_chunk_bytes=16 * 37)))
    assert streamed.size == a.packed_eri_Ha.size
    assert np.max(np.abs(streamed - a.packed_eri_Ha)) < 1e-14
    assert b.diagnostics()["resident_eri_bytes"] == 0


def test_neon_reference_energy_is_the_gpu_dirac_fock_energy():
    """Ne 1s2 2s2 2p6 (60 splines, order 7) solved on the GPU; the export
    with 3s 3p 3d virtuals (28 spinors) reproduces its total energy, also
    with the 1s frozen (then as the frozen-core constant plus the rest)."""
    import time

    from atomx_amd import bspline as B
    from atomx_amd import dirac as D
    from atomx_amd.spinor_integrals import build_spinor_integrals
    t = B.make_knots(60, 7, 1e-5, 40.0)
    bd = B.prepare_bspline_data(t, 60, 7)
    scf = D.DiracFockSCF(10, D.neon_subshells(), bd).bare_nucleus_start()
    e_df = scf.run(tol=1e-9)
    orbitals = {(sh.n, sh.kappa): (scf.P[i], scf.Q[i]) for i, sh in enumerate(scf.shells)}
    for kappa, n in ((-1, 3), (1, 3), (-2, 3), (2, 3), (-3, 3)):
        P, Q, _ = scf.virtual_orbitals(kappa, 1)[0]
        orbitals[(n, kappa)] = (P, Q)
    occ = [(1, -1), (2, -1), (2, 1), (2, -2)]
    vir = [(3, -1), (3, 1), (3, -2), (3, 2), (3, -3)]
    t0 = time.perf_counter()
    si = build_spinor_integrals(orbitals, bd, 10, bd["ns"], occupied_keys=occ, virtual_keys=vir)
    dt = time.perf_counter() - t0
    print(f"Ne export: {len(si.labels)} spinors, {si.packed_eri_Ha.size} packed values in "
          f"{dt * 1e3:.1f} ms; E_ref {si.reference_energy_Ha:.10f}, E_DF {e_df:.10f}", flush=True)
    assert len(si.labels) == 28
    assert abs(si.reference_energy_Ha - e_df) < 1e-9
    fz = build_spinor_integrals(orbitals, bd, 10, bd["ns"], occupied_keys=occ[1:],
                                virtual_keys=vir, frozen_keys=occ[:1])
    assert abs(fz.reference_energy_Ha - e_df) < 1e-9
    # the frozen 1s2 alone: its two one-body elements and the one pair
    # integral of the unfrozen export (spinors 0, 1 are 1s m = -1/2, +1/2)
    e_1s2 = 2.0 * si.one_body_Ha[0, 0].real + si.eri(0, 1, 0, 1).real
    assert abs(fz.frozen_energy_Ha - e_1s2) < 1e-10
    assert np.max(np.abs(fz.fock_Ha - si.fock_Ha[2:, 2:])) < 1e-10
