"""Per-iteration Fock build of many atoms or ions (uneven bases, orbital
counts and k ranges).

Each ion keeps its banks, factors, coefficient tables and coefficient-stack
buffers resident on the device.  The default path is one ``FockEngine``
plan for all ions: five launches per SCF iteration (direct densities,
semiseparable chains, direct contractions, exchange and assembly per output
tile, symmetrisation).  ``path="launches"`` keeps the earlier form, every
ion's prepared build (``FockBuilder.prepare``, launches per (block, k)),
for comparison.  Either is captured once into a HIP graph; an SCF iteration
writes the new coefficient stacks into the resident buffers and replays the
graph, which leaves every target's four Fock blocks (PP, QQ, PQ, QP) in
resident output tensors.

All targets of an ion are built from the orbitals of the same iteration (the
Jacobi order of one batched build).  The reference's solver updates each
orbital before the next target's build (methods/dirac_fock.py:2165-2168); a
caller that needs that order interleaves independent ions
(``atomx_amd.batch_solve``).
"""

from __future__ import annotations

import numpy as np

from . import device as dv
from .dirac_fock import FockBuilder, FockEngine

__all__ = ["IonBatch"]


class IonBatch:
    """ions: sequence of dicts with keys bsp_data, kappas, occupations (and
    optionally P, Q initial coefficient stacks (n_orb, ns)).
    path: "engine" (one plan for all ions) or "launches" (per-ion builds)."""

    def __init__(self, ions, skip_pq_exchange=False, path="engine", target_group=4):
        import torch
        dv.require_gpu()
        if path not in ("engine", "launches"):
            raise ValueError(f"unknown path {path!r}")
        self._torch = torch
        self.path = path
        self.builders = [FockBuilder(ion["bsp_data"], ion["kappas"], ion["occupations"])
                         for ion in ions]
        self.graph = None
        self.out = None
        if path == "engine":
            self.engine = FockEngine(self.builders, skip_pq_exchange=skip_pq_exchange,
                                     target_group=target_group)
            self.P, self.Q = self.engine.P, self.engine.Q
            self.rk_equivalents = self.engine.rk_equivalents
        else:
            self.engine = None
            self.steps = [fb.prepare(skip_pq_exchange=skip_pq_exchange) for fb in self.builders]
            self.P = [dv.zeros((fb.n_orb, fb.ns)) for fb in self.builders]
            self.Q = [dv.zeros((fb.n_orb, fb.ns)) for fb in self.builders]
            self.rk_equivalents = sum(s.rk_equivalents() for s in self.steps)
        if all("P" in ion for ion in ions):
            self.load([ion["P"] for ion in ions], [ion["Q"] for ion in ions])

    def load(self, P_list, Q_list):
        """Write new coefficient stacks into the resident buffers (current stream)."""
        for i, (P, Q) in enumerate(zip(P_list, Q_list)):
            ns = self.builders[i].ns
            for dst, src in ((self.P[i], P), (self.Q[i], Q)):
                src = np.asarray(src, dtype=np.float64)
                buf = np.zeros(dst.shape)
                m = min(src.shape[1], ns)
                buf[:, :m] = src[:, :m]
                dst.copy_(dv.f64(buf))

    def run_eager(self):
        """One build of every ion without the graph: list of {t: [PP, QQ, PQ, QP]}."""
        if self.engine is not None:
            self.engine.run()
            return self.engine.blocks()
        return [step.run(P, Q) for step, P, Q in zip(self.steps, self.P, self.Q)]

    def capture(self):
        """Capture every ion's build into one graph (after a warm-up run)."""
        torch = self._torch
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run_eager()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = self.run_eager()
        torch.cuda.synchronize()
        return self

    def replay(self):
        """One SCF iteration's Fock build of all ions (graph launch on the
        current stream); returns the resident outputs."""
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.out

    def host(self, out=None):
        out = self.out if out is None else out
        return [{t: tuple(dv.host(x) for x in blocks) for t, blocks in H.items()} for H in out]
