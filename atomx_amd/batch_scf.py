"""Per-iteration Fock build of many atoms or ions (uneven bases, orbital
counts and k ranges) replayed as one HIP graph.

Each ion keeps its banks, factors, coefficient tables and coefficient-stack
buffers resident on the device (``FockBuilder.prepare``).  The launches of
every ion's build are captured once into a single graph; an SCF iteration
writes the new coefficient stacks into the resident buffers (one
asynchronous copy per ion) and replays the graph, which leaves every
target's four Fock blocks (PP, QQ, PQ, QP) in resident output tensors.

All targets of an ion are built from the orbitals of the same iteration (the
Jacobi order of one batched build).  The reference's solver updates each
orbital before the next target's build (methods/dirac_fock.py:2165-2168); a
caller that needs that order builds one target per replay of a per-target
graph, or interleaves independent ions.
"""

from __future__ import annotations

import numpy as np

from . import device as dv
from .dirac_fock import FockBuilder

__all__ = ["IonBatch"]


class IonBatch:
    """ions: sequence of dicts with keys bsp_data, kappas, occupations (and
    optionally P, Q initial coefficient stacks (n_orb, ns))."""

    def __init__(self, ions, skip_pq_exchange=False):
        import torch
        dv.require_gpu()
        self._torch = torch
        self.builders, self.steps, self.P, self.Q = [], [], [], []
        for ion in ions:
            fb = FockBuilder(ion["bsp_data"], ion["kappas"], ion["occupations"])
            step = fb.prepare(skip_pq_exchange=skip_pq_exchange)
            self.builders.append(fb)
            self.steps.append(step)
            n_orb, ns = fb.n_orb, fb.ns
            self.P.append(dv.zeros((n_orb, ns)))
            self.Q.append(dv.zeros((n_orb, ns)))
        self.graph = None
        self.out = None
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
