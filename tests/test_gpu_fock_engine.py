"""GPU: the batched Fock engine (one plan of five launches for every target
of every ion) against the per-(block, k) launch path and the oracle's
restatement of methods/dirac_fock.py:1128-1319 (_build_hf_matrix_dbsr).

Ions on both sides of the ns >= 96 semiseparable direct switch, orders 7
and 8 with kinetic balance, 1-10 orbitals.  Tolerance: 1e-12 of each
block's largest entry against the oracle; 1e-13 against the launch path
(the same contractions summed in another order)."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def ions():
    from atomx_amd.workloads import fe_ion_workload
    all_ions = fe_ion_workload(seed=5, ns_range=(40, 150))
    return [all_ions[i] for i in (0, 3, 9, 14, 20, 25)]


@pytest.mark.parametrize("skip", [False, True])
def test_engine_matches_launch_path_and_oracle(ions, skip):
    from atomx_amd.batch_scf import IonBatch
    eng = IonBatch(ions, skip_pq_exchange=skip)
    ref = IonBatch(ions, skip_pq_exchange=skip, path="launches")
    assert eng.rk_equivalents == ref.rk_equivalents
    got = eng.host(eng.run_eager())
    want = ref.host(ref.run_eager())
    for ion, g_ion, w_ion in zip(ions, got, want):
        bd = ion["bsp_data"]
        print(f"Fe{ion['charge']}+ ns={bd['ns']} ks={bd['ks']} orbitals={len(ion['kappas'])}",
              flush=True)
        banks = O.Banks(bd)
        tables = O.angular_tables(ion["occupations"], ion["kappas"])
        for t in w_ion:
            oracle = O.build_hf_matrix_dbsr(t, ion["P"], ion["Q"], ion["occupations"],
                                            ion["kappas"], bd, banks, tables=tables,
                                            skip_pq_exchange=skip)
            for name, a, b, c in zip(("PP", "QQ", "PQ", "QP"), g_ion[t], w_ion[t], oracle):
                assert _rel(a, c) <= 1e-12, (ion["charge"], t, name, "oracle")
                assert _rel(a, b) <= 1e-13, (ion["charge"], t, name, "launch path")
                if name in ("PP", "QQ"):
                    assert np.array_equal(a, a.T)


def test_engine_graph_replays_follow_new_orbitals(ions):
    import torch

    from atomx_amd.batch_scf import IonBatch
    batch = IonBatch(ions).capture()
    rng = np.random.default_rng(9)
    for it in range(3):
        Ps = [ion["P"] * (1.0 + 0.1 * it) + 0.01 * rng.normal(size=ion["P"].shape)
              for ion in ions]
        Qs = [ion["Q"] * (1.0 - 0.05 * it) for ion in ions]
        batch.load(Ps, Qs)
        eager = batch.host(batch.run_eager())
        torch.cuda.synchronize()
        batch.replay()
        torch.cuda.synchronize()
        for g_ion, e_ion in zip(batch.host(), eager):
            for t in e_ion:
                for a, b in zip(g_ion[t], e_ion[t]):
                    assert np.array_equal(a, b), it
