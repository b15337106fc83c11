"""GPU: the batched Dirac-Fock SCF of the 26 iron charge states (uneven
bases ns 40..160, orders 7-8 with kinetic balance, 1-10 orbitals, k <= 6),
every ion converged from the bare nucleus, against the same Gauss-Seidel
SCF of each ion run alone on the oracle backend (oracle/scf.py).
Tolerance: 1e-10 Ha on every total energy; both sweep until the orbitals
change by less than 1e-9."""

import multiprocessing as mp
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_energy(charge):
    from oracle.scf import fe_ion_energy
    return fe_ion_energy(charge)


def test_all_26_iron_ions_scf_matches_oracle_scf():
    from concurrent.futures import as_completed

    from atomx_amd.batch_solve import BatchDiracFock, fe_ion_subshells
    from atomx_amd.workloads import fe_ion_workload
    ions = fe_ion_workload(seed=0)
    batch = BatchDiracFock([(26, fe_ion_subshells(c), ion["bsp_data"])
                            for c, ion in enumerate(ions)])
    try:
        e_gpu = batch.run(tol=1e-9)
        sweeps = [s.sweeps for s in batch.systems]
        print(f"batched SCF: {batch.sweeps} sweeps, {np.mean(batch.sweep_ms):.1f} ms per sweep "
              f"(Fock builds {np.mean(batch.gpu_ms):.2f} ms of GPU, copies "
              f"{np.mean(batch.copy_ms):.1f} ms, eigensolves {np.mean(batch.solve_ms):.1f} ms)",
              flush=True)
    finally:
        batch.close()
    # the oracle SCFs afterwards, in CPU worker processes (no GPU), one
    # thread each: a spawned child takes the environment at spawn time
    ctx = mp.get_context("spawn")
    ref = {}
    names = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
    saved = {n: os.environ.get(n) for n in names}
    os.environ.update({n: "1" for n in names})
    try:
        with ProcessPoolExecutor(max_workers=16, mp_context=ctx) as pool:
            futs = {pool.submit(_oracle_energy, c): c for c in range(26)}
            for f in as_completed(futs, timeout=600):
                c = futs[f]
                ref[c] = f.result()
                print(f"oracle SCF of Fe{c}+ done ({len(ref)} of 26)", flush=True)
    finally:
        for n, v in saved.items():
            if v is None:
                os.environ.pop(n, None)
            else:
                os.environ[n] = v
    worst = 0.0
    for c, e in enumerate(e_gpu):
        e_ref, n_ref = ref[c]
        d = abs(e - e_ref)
        worst = max(worst, d)
        print(f"Fe{c}+: E = {e:.10f} Ha, oracle SCF {e_ref:.10f} Ha, |dE| = {d:.1e} "
              f"(sweeps {sweeps[c]} / {n_ref})", flush=True)
        assert d <= 1e-10, c
    print(f"max |dE| = {worst:.2e} Ha")
