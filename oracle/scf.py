"""The Dirac-Fock SCF of one system with the oracle's two-electron blocks
and energy — TEST INFRASTRUCTURE ONLY (the checker of the GPU SCF drivers;
imported by tests/ and bench.py's CPU leg, never by the product).

The SCF driver itself (one-body Dirac matrices, eigensolve, Gauss-Seidel
order) is the product's host code (atomx_amd.dirac.DiracFockSCF), run here
with oracle.build_hf_matrix_dbsr and oracle.two_electron_energy in place of
the GPU backend, so a GPU SCF and this one differ only in how the
two-electron pieces are computed."""

from __future__ import annotations


def oracle_scf_energy(Z, shells, bsp_data, nuclear="point", tol=1e-9, max_sweeps=100):
    """(E_total, sweeps) of the Gauss-Seidel SCF from the bare nucleus,
    swept until the orbitals change by less than ``tol``."""
    import oracle as O
    from atomx_amd import dirac as D

    occ = [s.occupation for s in shells]
    kap = [s.kappa for s in shells]
    banks = O.Banks(bsp_data)
    tables = O.angular_tables(occ, kap)

    class Backend:
        def fock(self, i, P, Q):
            return O.build_hf_matrix_dbsr(i, P, Q, occ, kap, bsp_data, banks,
                                          skip_pq_exchange=True, tables=tables)

        def energy(self, P, Q):
            return O.two_electron_energy(P, Q, occ, kap, bsp_data, banks)

    scf = D.DiracFockSCF(Z, shells, bsp_data, nuclear=nuclear, backend=Backend())
    scf.bare_nucleus_start()
    for sweep in range(1, max_sweeps + 1):
        if scf.sweep() < tol:
            return scf.energy()[0], sweep
    raise RuntimeError(f"oracle SCF did not converge in {max_sweeps} sweeps")


def fe_ion_energy(charge, seed=0, tol=1e-9):
    """E_total and sweeps of one iron ion of ``fe_ion_workload(seed)``."""
    from atomx_amd.batch_solve import fe_ion_subshells
    from atomx_amd.workloads import fe_ion_workload
    ion = fe_ion_workload(seed=seed)[charge]
    return oracle_scf_energy(26, fe_ion_subshells(charge), ion["bsp_data"], tol=tol)
