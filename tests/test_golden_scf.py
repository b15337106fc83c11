"""CPU: the iron-ion oracle SCF energies in tests/golden/fe_ion_scf_oracle.json
(the checker of the GPU batched SCF) are what the oracle computes now, for
the ions that converge in seconds; and the fixture covers every charge
state with the workload's bases."""

import json
import os

import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fe_ion_scf_oracle.json")


def test_fixture_covers_the_workload():
    from atomx_amd.workloads import fe_ion_workload
    doc = json.load(open(GOLDEN))
    ions = fe_ion_workload(seed=doc["seed"])
    assert sorted(int(c) for c in doc["ions"]) == list(range(26))
    for c, ion in enumerate(ions):
        row = doc["ions"][str(c)]
        assert row["ns"] == ion["bsp_data"]["ns"] and row["ks"] == ion["bsp_data"]["ks"]
        assert row["orbitals"] == len(ion["kappas"])
    # Fe25+ has one electron: its SCF energy is the 1s Dirac eigenvalue of a
    # point nucleus, c^2 (sqrt(1 - (Z/c)^2) - 1), up to the basis error
    import math

    from atomx_amd.dirac import C_SQ
    c = math.sqrt(C_SQ)
    exact = C_SQ * (math.sqrt(1.0 - (26.0 / c) ** 2) - 1.0)
    e25 = doc["ions"]["25"]["E_total_Ha"]
    assert e25 >= exact and e25 - exact < 1e-6


@pytest.mark.parametrize("charge", [8, 15, 25])
def test_fixture_matches_oracle_scf(charge):
    from oracle.scf import fe_ion_energy
    doc = json.load(open(GOLDEN))
    e, sweeps = fe_ion_energy(charge, seed=doc["seed"], tol=doc["tol"])
    row = doc["ions"][str(charge)]
    assert abs(e - row["E_total_Ha"]) <= 1e-10
    assert sweeps == row["sweeps"]
