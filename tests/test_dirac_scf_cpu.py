"""CPU pins of the host side of the Dirac-Fock SCF (one-body Dirac matrices
with kinetic balance, the Bloch boundary term and the Fermi nucleus; the
reduced eigensolve) on the DBSR_HF neon spinors of the reference fixture.
On these fixed spinors the reference test (tests/test_dbsr_common_spinors.py
:38-61) holds E_1e, E_2e, E_total and every orbital's Fock expectation to
2e-10 Ha; the two-electron parts come from the oracle here."""

import numpy as np
import pytest

import oracle as O
from atomx_amd import dirac as D
from helpers import dbsr_state


@pytest.fixture(scope="module")
def ne():
    bd, P, Q, occ, kap, ref = dbsr_state()
    banks = O.Banks(bd)
    return bd, P, Q, occ, kap, ref, banks


class _OracleBackend:
    def __init__(self, bd, occ, kap, banks):
        self.bd, self.occ, self.kap, self.banks = bd, occ, kap, banks
        self.tables = O.angular_tables(occ, kap)

    def fock(self, i, P, Q):
        return O.build_hf_matrix_dbsr(i, P, Q, self.occ, self.kap, self.bd, self.banks,
                                      skip_pq_exchange=True, tables=self.tables)

    def energy(self, P, Q):
        return O.two_electron_energy(P, Q, self.occ, self.kap, self.bd, self.banks)


def _scf(ne, backend=None):
    bd, P, Q, occ, kap, ref, banks = ne
    shells = D.neon_subshells()
    scf = D.DiracFockSCF(10, shells, bd, nuclear="fermi", A=20,
                         backend=backend or _OracleBackend(bd, occ, kap, banks))
    return scf.set_orbitals(P, Q)


def test_fermi_parameters_of_neon():
    """c and a of the Fermi model reproduce the tabulated RMS radius
    (3.0055 fm) through the closed-form moment (Parpia & Mohanty)."""
    a, c = D.fermi_params(10, 20)
    rms_fm = D._rms_of(a * D.BOHR_RADIUS_FM, c * D.BOHR_RADIUS_FM)
    assert abs(rms_fm - 3.0055) < 1e-12
    assert abs(a * D.BOHR_RADIUS_FM - 2.30 / (4 * np.log(3.0))) < 1e-15
    # far outside the nucleus the potential is -Z/r
    r = np.array([1e-3, 0.1, 1.0])
    assert np.allclose(D.nuclear_potential(10, r, "fermi", 20), -10 / r, rtol=1e-13)
    # at the origin it is finite and below the point potential's value at c
    assert D.nuclear_potential(10, np.array([0.0]), "fermi", 20)[0] < -10 / c


def test_one_and_two_electron_energies_on_fixture_spinors(ne):
    bd, P, Q, occ, kap, ref, banks = ne
    e_tot, e1, e2 = _scf(ne).energy()
    assert e1 == pytest.approx(ref["components"]["E_1e"], abs=2e-10, rel=0)
    assert e2 == pytest.approx(ref["components"]["E_2e"], abs=2e-10, rel=0)
    assert e_tot == pytest.approx(ref["reference"]["total_energy_Ha"], abs=2e-10, rel=0)


def test_fock_expectations_on_fixture_spinors(ne):
    bd, P, Q, occ, kap, ref, banks = ne
    scf = _scf(ne)
    ns = scf.ns
    for i, sh in enumerate(scf.shells):
        H, S = scf.fock_matrix(i)
        h = np.tril(H) + np.tril(H, -1).T
        v = np.concatenate((scf.P[i], scf.Q[i]))
        assert v @ S @ v == pytest.approx(1.0, abs=5e-13, rel=0)
        assert v @ h @ v == pytest.approx(
            ref["reference"]["orbital_fock_expectations_Ha"][sh.label], abs=2e-10, rel=0)


def test_one_sweep_keeps_the_converged_spinors(ne):
    """The fixture spinors are a fixed point of this SCF: two Gauss-Seidel
    sweeps change the large components by < 1e-6 and then < 1e-8 of their
    maxima, every eigenvalue stays within 5e-8 Ha of the fixture's Fock
    expectation (each sees the earlier orbitals' small update; the printed
    orbital energies of the fixture are a rounded figure 5e-7 Ha away), and
    the total energy stays on the reference's -128.69192584304 Ha within
    2e-10 Ha."""
    bd, P, Q, occ, kap, ref, banks = ne
    scf = _scf(ne)
    assert scf.sweep() < 1e-6
    assert scf.sweep() < 1e-8
    for i, sh in enumerate(scf.shells):
        assert scf.eps[i] == pytest.approx(
            ref["reference"]["orbital_fock_expectations_Ha"][sh.label], abs=5e-8, rel=0)
    assert scf.energy()[0] == pytest.approx(ref["reference"]["total_energy_Ha"], abs=2e-10, rel=0)
