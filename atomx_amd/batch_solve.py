"""Dirac-Fock SCF of many atoms or ions at once: independent Gauss-Seidel
solves whose per-orbital Fock builds share GPU launches.

Each system follows its own solver exactly as ``DiracFockSCF`` (and the
reference, methods/dirac_fock.py:2057-2204) does: within a sweep orbital i
is rebuilt from the orbitals as updated so far (the Fock blocks of target i
from the current P, Q, then the reduced generalised eigensolve under
orthogonality to the other orbitals of its kappa).  Systems are independent
(the reference runs them as separate solves, methods/concurrent_solve.py:
72-114), so sweep s of every system can advance together:

  stage i of a sweep: one ``FockEngine`` run builds the Fock blocks of
  target i of every system that has an orbital i (five launches for all
  of them), one device-to-host copy brings them back, and the systems'
  eigensolves run in parallel host threads (LAPACK releases the
  interpreter lock) while the updated orbitals are written back to the
  resident device stacks for stage i + 1.

The eigensolve stays FP64 on the host, unchanged from the single-system
driver.  A system whose orbitals change by less than ``tol`` in a sweep is
converged and leaves the later sweeps; its energy is evaluated once, at the
end.
"""

from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import device as dv
from .dirac import DiracFockSCF, Subshell, one_blas_thread, solve_dirac

__all__ = ["BatchDiracFock", "fe_ion_subshells"]


def fe_ion_subshells(charge):
    """Subshells of the iron ion of the given charge (ground configuration,
    average level; ``workloads.fe_configuration``)."""
    from .workloads import fe_configuration
    return [Subshell(label, int(label[0]), kappa, occ)
            for label, kappa, occ in fe_configuration(charge)]


class _System:
    """Host state of one system: a ``DiracFockSCF`` whose two-electron
    blocks come from the shared engine runs."""

    def __init__(self, Z, subshells, bsp_data, nuclear, builder):
        self.scf = DiracFockSCF(Z, subshells, bsp_data, nuclear=nuclear, backend=_NoBackend())
        self.fb = builder
        self.change = 0.0
        self.converged = False
        self.sweeps = 0


class _NoBackend:
    def fock(self, i, P, Q):  # pragma: no cover - the batch driver supplies the blocks
        raise RuntimeError("Fock blocks come from the batch engine")

    def energy(self, P, Q):  # pragma: no cover
        raise RuntimeError("energies come from the batch driver")


class BatchDiracFock:
    """systems: sequence of (Z, subshells, bsp_data); ``nuclear`` as in
    ``DiracFockSCF``.  threads: host threads for the eigensolves (default:
    the CPUs this process may use, at most 16).  diis: Pulay extrapolation
    per system, as ``DiracFockSCF.sweep(diis=True)`` applies it."""

    def __init__(self, systems, nuclear="point", threads=None, target_group=4, diis=False):
        from .dirac_fock import FockBuilder, FockEngine
        dv.require_gpu()
        self.systems = []
        builders = []
        for Z, shells, bd in systems:
            fb = FockBuilder(bd, [s.kappa for s in shells], [s.occupation for s in shells])
            builders.append(fb)
            self.systems.append(_System(Z, shells, bd, nuclear, fb))
        # resident coefficient stacks shared by the stage plans
        self.P = [dv.zeros((fb.n_orb, fb.ns)) for fb in builders]
        self.Q = [dv.zeros((fb.n_orb, fb.ns)) for fb in builders]
        n_stages = max(fb.n_orb for fb in builders)
        self.stages = []
        for i in range(n_stages):
            members = [a for a, fb in enumerate(builders) if fb.n_orb > i]
            eng = FockEngine([builders[a] for a in members], skip_pq_exchange=True,
                             target_group=target_group, targets=[[i]] * len(members),
                             P=[self.P[a] for a in members], Q=[self.Q[a] for a in members])
            self.stages.append((members, eng))
        self.rk_equivalents_per_sweep = sum(eng.rk_equivalents for _, eng in self.stages)
        if threads is None:
            try:
                threads = len(os.sched_getaffinity(0))
            except AttributeError:
                threads = os.cpu_count() or 1
            threads = max(1, min(16, threads))
        self.threads = threads
        self.diis = diis
        self.extrapolations = 0
        self._pool = ThreadPoolExecutor(threads) if threads > 1 else None
        self.gpu_ms = []       # GPU time of the Fock builds of each sweep (all stages)
        self.copy_ms = []      # host time of the block downloads and orbital uploads
        self.solve_ms = []     # host time of the eigensolves
        self.sweep_ms = []     # wall time of each sweep
        self.sweeps = 0

    # -- helpers -------------------------------------------------------------------
    def _map(self, fn, items):
        # one BLAS thread per eigensolve: the parallelism is across systems
        with one_blas_thread():
            if self._pool is None:
                return [fn(x) for x in items]
            return list(self._pool.map(fn, items))

    def _upload(self, a, rows=None):
        """Writes system a's coefficient rows (all, or the given ones) to the device."""
        scf = self.systems[a].scf
        P = dv.f64(scf.P if rows is None else scf.P[rows])
        Q = dv.f64(scf.Q if rows is None else scf.Q[rows])
        if rows is None:
            self.P[a].copy_(P)
            self.Q[a].copy_(Q)
        else:
            self.P[a][rows] = P
            self.Q[a][rows] = Q

    def _solve(self, a, i, blocks):
        """The eigensolve of orbital i of system a with its Fock blocks."""
        sysm = self.systems[a]
        scf = sysm.scf
        sh = scf.shells[i]
        H, S = scf.h0[sh.kappa]
        H = H.copy()
        pp, qq, pq, qp = blocks
        ns = scf.ns
        H[:ns, :ns] += pp
        H[ns:, ns:] += qq
        H[:ns, ns:] += pq
        H[ns:, :ns] += qp
        others = [np.concatenate((scf.P[j], scf.Q[j])) for j, o in enumerate(scf.shells)
                  if j != i and o.kappa == sh.kappa]
        cons = np.column_stack(others) if others else None
        w, v = solve_dirac(H, S, ns, scf.nsp, scf.nsq, orthogonal_to=cons)
        P, Q, e = scf._root(i, w, v, cons is not None)
        pmax = np.max(np.abs(scf.P[i]))
        if pmax > 1e-15:
            if np.dot(P, scf.P[i]) < 0:
                P, Q = -P, -Q
            sysm.change = max(sysm.change, float(np.max(np.abs(P - scf.P[i])) / pmax))
        scf.P[i], scf.Q[i], scf.eps[i] = P, Q, e

    # -- driver --------------------------------------------------------------------
    def start(self):
        """Bare-nucleus orbitals for every system (host), uploaded."""
        self._map(lambda s: s.scf.bare_nucleus_start(), self.systems)
        for a in range(len(self.systems)):
            self._upload(a)
        return self

    def sweep(self):
        """One Gauss-Seidel sweep of every unconverged system."""
        import torch
        t0 = time.perf_counter()
        gpu = copy = solve = 0.0
        for sysm in self.systems:
            sysm.change = 0.0
        # DIIS as DiracFockSCF.sweep applies it: five or more orbitals, fourth sweep on
        x_in = {a: s.scf._flat() for a, s in enumerate(self.systems)
                if self.diis and not s.converged and len(s.scf.shells) >= 5 and s.sweeps >= 3}
        for i, (members, eng) in enumerate(self.stages):
            live = [r for r, a in enumerate(members) if not self.systems[members[r]].converged]
            if not live:
                continue
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            eng.run()
            ev1.record()
            ta = time.perf_counter()
            host = [dv.host(eng.H[r][0]) for r in live]
            tb = time.perf_counter()
            gpu += ev0.elapsed_time(ev1)
            self._map(lambda rb: self._solve(members[rb[0]], i, rb[1]),
                      list(zip(live, host)))
            tc = time.perf_counter()
            for r in live:
                self._upload(members[r], rows=[i])
            copy += (tb - ta) + (time.perf_counter() - tc)
            solve += tc - tb
        for a, x in x_in.items():
            sysm = self.systems[a]
            if sysm.scf.diis_step(x, sysm.change):
                self.extrapolations += 1
                self._upload(a)
        for sysm in self.systems:
            if not sysm.converged:
                sysm.sweeps += 1
        self.sweeps += 1
        self.gpu_ms.append(gpu)
        self.copy_ms.append(1e3 * copy)
        self.solve_ms.append(1e3 * solve)
        self.sweep_ms.append(1e3 * (time.perf_counter() - t0))
        return max((s.change for s in self.systems if not s.converged), default=0.0)

    def run(self, max_sweeps=100, tol=1e-9):
        """Sweeps until every system's orbitals change by less than ``tol``
        (relative to each large component's maximum); returns the total
        energies (E_total per system, Ha)."""
        self.start()
        for _ in range(max_sweeps):
            self.sweep()
            for sysm in self.systems:
                if not sysm.converged and sysm.change < tol:
                    sysm.converged = True
            if all(s.converged for s in self.systems):
                return self.energies()
        raise RuntimeError(f"{sum(not s.converged for s in self.systems)} systems did not "
                           f"converge in {max_sweeps} sweeps")

    def energies(self):
        """E_total of every system at its current orbitals (the GPU two-
        electron energy, methods/dirac_fock.py:1553-1661)."""
        out = []
        for sysm in self.systems:
            scf = sysm.scf
            e1 = 0.0
            for i, sh in enumerate(scf.shells):
                H, S = scf.h0[sh.kappa]
                v = np.concatenate((scf.P[i], scf.Q[i]))
                nrm = v @ S @ v
                if nrm > 1e-20:
                    e1 += sh.occupation * (v @ H @ v) / nrm
            out.append(e1 + sysm.fb.energy(scf.P, scf.Q)[0])
        return np.asarray(out)

    def close(self):
        if self._pool is not None:
            self._pool.shutdown()
            self._pool = None
        for _, eng in self.stages:
            eng.close()
