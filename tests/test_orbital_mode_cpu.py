"""CPU pins of the oracle's restatement of the DF solver's orbital mode
(methods/dirac_fock.py:840-1068, :1469-1551) on the DBSR_HF neon spinors of
the reference fixture: for these closed shells the orbital-mode Fock
expectation of every orbital equals the fixture's Fock expectation
(tests/test_dbsr_common_spinors.py:38-61 tolerance, 2e-10 Ha), and the
orbital-mode SCF keeps the fixture spinors converged."""

import numpy as np
import pytest

import oracle as O
from oracle import orbital_mode as OM
from atomx_amd import dirac as D
from helpers import dbsr_state


@pytest.fixture(scope="module")
def ne():
    bd, P, Q, occ, kap, ref = dbsr_state()
    return bd, P, Q, occ, kap, ref, O.Banks(bd)


class _OrbitalOracle:
    def __init__(self, bd, occ, kap, banks):
        self.bd, self.occ, self.kap, self.banks = bd, occ, kap, banks

    def fock(self, i, P, Q):
        return OM.orbital_fock_blocks(i, P, Q, self.occ, self.kap, self.bd, self.banks)

    def energy(self, P, Q):
        return O.two_electron_energy(P, Q, self.occ, self.kap, self.bd, self.banks)


def test_orbital_mode_fock_expectations_match_fixture(ne):
    bd, P, Q, occ, kap, ref, banks = ne
    scf = D.DiracFockSCF(10, D.neon_subshells(), bd, nuclear="fermi", A=20,
                         backend=_OrbitalOracle(bd, occ, kap, banks)).set_orbitals(P, Q)
    for i, sh in enumerate(scf.shells):
        H, S = scf.fock_matrix(i)
        v = np.concatenate((scf.P[i], scf.Q[i]))
        e = v @ H @ v / (v @ S @ v)
        assert e == pytest.approx(ref["reference"]["orbital_fock_expectations_Ha"][sh.label],
                                  abs=2e-10, rel=0)


def test_intra_shell_term_vanishes_for_j_half_and_closes_the_2p32_shell(ne):
    bd, P, Q, occ, kap, ref, banks = ne
    for t in range(3):   # 1s1/2, 2s1/2, 2p1/2: j = 1/2, no k > 0 self term
        vpp, vqq = OM.intra_shell_potential(t, P, Q, occ, kap, bd, banks)
        assert not vpp.any() and not vqq.any()
    vpp, _ = OM.intra_shell_potential(3, P, Q, occ, kap, bd, banks)
    # 2p3/2 (q = 4): -(q-1) Lambda(2) / (2j(2j+1)) F^2 contraction, negative definite
    ns = bd["ns"]
    p = np.zeros(ns)
    p[:P.shape[1]] = P[3]
    assert p @ vpp @ p < 0


def test_orbital_mode_scf_keeps_fixture_spinors(ne):
    bd, P, Q, occ, kap, ref, banks = ne
    scf = D.DiracFockSCF(10, D.neon_subshells(), bd, nuclear="fermi", A=20,
                         backend=_OrbitalOracle(bd, occ, kap, banks)).set_orbitals(P, Q)
    assert scf.sweep() < 1e-6
    assert scf.energy()[0] == pytest.approx(ref["reference"]["total_energy_Ha"], abs=2e-10, rel=0)
