"""Y^k-form Fock operator pieces on the GPU, batched over the targets of one
atom (the Fock-build requests of the request server).

* direct: methods/dirac_fock.py:929-965 (_build_direct_potential_yk) —
  rho_t = sum_j q_j (P_j^2 + Q_j^2) - (P_t^2 + Q_t^2), Y^0[rho_t], and the
  banded assembly <B_i | grw Y^0 | B_j> (numerics/bspline/integrals.py:18-56)
  in the P and Q bases, symmetrised.  All targets share one density launch,
  one row combination, one Y^0 scan and one assembly launch per basis.
* exchange: numerics/bspline/slater_yk.py:125-152 (compute_exchange_matrix)
  for every (source orbital, rank) with a nonzero coefficient, combined per
  target as sum_{b,k} c[t, b, k] K^k_b: (b, k) order inside chunks of 64
  terms, chunk sums added in order.
"""

from __future__ import annotations

import numpy as np

from . import device as dv
from . import grid as _grid
from ._lib import call
from .slater_yk import _pad, eval_orbitals_device
from .yk import yk_rows

__all__ = ["FockYk"]


class FockYk:
    """Device-resident basis and grid of one atom; ``build`` returns the Fock
    pieces of every target as device tensors."""

    def __init__(self, bsp_data):
        bd = bsp_data
        dv.require_gpu()
        self.nv, self.ns = int(bd["nv"]), int(bd["ns"])
        gr, grw = np.asarray(bd["gr"], float), np.asarray(bd["grw"], float)
        self.ng = gr.shape[1]
        self.n = self.nv * self.ng
        self.ksp, self.nsp = int(bd.get("ksp", bd["ks"])), int(bd.get("nsp", self.ns))
        self.ksq, self.nsq = int(bd.get("ksq", bd["ks"])), int(bd.get("nsq", self.ns))
        self.bsp_p_host = np.ascontiguousarray(bd["bsp_p"], dtype=np.float64)
        self.bsp_q_host = np.ascontiguousarray(bd.get("bsp_q", bd["bsp_p"]), dtype=np.float64)
        self.bsp_p = dv.f64(self.bsp_p_host)
        self.bsp_q = dv.f64(self.bsp_q_host)
        self.tables = _grid.get(gr.ravel(), grw.ravel(), self.nv, self.ng, 0)
        self.r = self.tables.r
        self.w = self.tables.w

    # -- orbital samples ------------------------------------------------------
    def samples(self, P_coef, Q_coef):
        P = np.atleast_2d(np.asarray(P_coef, float))
        Q = np.atleast_2d(np.asarray(Q_coef, float))
        fP = eval_orbitals_device(np.stack([_pad(p, self.nsp) for p in P]), self.bsp_p_host,
                                  self.nv, self.ksp, self.nsp)
        fQ = eval_orbitals_device(np.stack([_pad(q, self.nsq) for q in Q]), self.bsp_q_host,
                                  self.nv, self.ksq, self.nsq)
        return fP, fQ

    # -- direct potentials -------------------------------------------------------
    def direct(self, fP, fQ, occ, targets):
        """(V_PP, V_QQ), each (n_targets, ns, ns): Y^0 potential of the
        occupation-weighted density minus one electron of each target."""
        n_orb = fP.shape[0]
        targets = [int(t) for t in targets]
        nt = len(targets)
        n, ns = self.n, self.ns
        s = dv.stream_handle()
        rho = dv.empty((n_orb, n))
        idx = dv.i32(np.arange(n_orb, dtype=np.int32))
        call("ax_pair_density", dv.ptr(fP), dv.ptr(fQ), n, n, n_orb, dv.ptr(idx), dv.ptr(idx),
             None, dv.ptr(rho), n, s)
        coef = np.tile(np.asarray(occ, float), (nt, 1))
        for r, t in enumerate(targets):
            coef[r, t] -= 1.0
        cd = dv.f64(coef)
        rho_t = dv.empty((nt, n))
        call("ax_combine_rows", dv.ptr(rho), n, n_orb, dv.ptr(cd), nt, dv.ptr(rho_t), s)
        y0 = yk_rows(rho_t, self.tables, 0)
        wts = dv.empty((nt, n))
        call("ax_mul_rows", dv.ptr(wts), dv.ptr(y0), dv.ptr(self.w), nt, n, s)
        out = []
        for bsp, ks in ((self.bsp_p, self.ksp), (self.bsp_q, self.ksq)):
            V = dv.empty((nt, ns, ns))
            call("ax_assemble_integral", self.nv, self.ng, ks, ks, ns, dv.ptr(bsp), dv.ptr(bsp),
                 dv.ptr(wts), nt, dv.ptr(V), s)
            call("ax_symmetrize", dv.ptr(V), ns, nt, s)
            out.append(V)
        return tuple(out)

    # -- exchange ----------------------------------------------------------------
    def exchange_matrix(self, fP, fQ, b, k):
        """K^k_b for the P and Q bases (slater_yk.py:125-152) as device tensors."""
        s = dv.stream_handle()
        out = []
        for f, bsp, ks, nsx in ((fP, self.bsp_p, self.ksp, self.nsp),
                                (fQ, self.bsp_q, self.ksq, self.nsq)):
            K = dv.empty((nsx, nsx))
            call("ax_exchange_matrix_yk", dv.ptr(f[b]), dv.ptr(bsp), dv.ptr(self.r),
                 dv.ptr(self.w), self.nv, self.ng, ks, nsx, int(k), dv.ptr(K), s)
            out.append(K)
        return tuple(out)

    def exchange(self, fP, fQ, coef, targets, chunk=64):
        """(K_PP, K_QQ), each (n_targets, ns, ns): sum_{b,k} coef[t, b, k] K^k_b.

        The K^k_b of up to ``chunk`` (b, k) terms are stacked and combined for
        all targets in one row-combination launch per basis; chunks are
        added in (b, k) order."""
        coef = np.asarray(coef, float)
        targets = [int(t) for t in targets]
        nt = len(targets)
        s = dv.stream_handle()
        terms = [(b, k) for b in range(coef.shape[1]) for k in range(coef.shape[2])
                 if np.any(coef[targets, b, k] != 0.0)]
        KP = dv.zeros((nt, self.nsp, self.nsp))
        KQ = dv.zeros((nt, self.nsq, self.nsq))
        if terms:
            self.tables.ensure_k(max(k for _, k in terms))
        tb = self.tables
        for c0 in range(0, len(terms), chunk):
            part = terms[c0:c0 + chunk]
            cm = dv.f64(np.stack([coef[targets, b, k] for b, k in part], axis=1))  # (nt, n_part)
            ib = dv.i32([b for b, _ in part])
            ik = dv.i32([k for _, k in part])
            for f, bsp, ks, nsx, acc in ((fP, self.bsp_p, self.ksp, self.nsp, KP),
                                         (fQ, self.bsp_q, self.ksq, self.nsq, KQ)):
                stack = dv.empty((len(part), nsx, nsx))
                call("ax_exchange_matrix_yk_batch", dv.ptr(f), dv.ld(f), dv.ptr(bsp),
                     dv.ptr(self.w), dv.ptr(tb.rk), dv.ptr(tb.rkm), tb.kmax, self.nv, self.ng,
                     ks, nsx, len(part), dv.ptr(ib), dv.ptr(ik), dv.ptr(stack), s)
                if c0 == 0:
                    call("ax_combine_rows", dv.ptr(stack), nsx * nsx, len(part), dv.ptr(cm), nt,
                         dv.ptr(acc), s)
                else:
                    tmp = dv.empty((nt, nsx, nsx))
                    call("ax_combine_rows", dv.ptr(stack), nsx * nsx, len(part), dv.ptr(cm), nt,
                         dv.ptr(tmp), s)
                    call("ax_axpy", dv.ptr(acc), dv.ptr(tmp), nt * nsx * nsx, 1.0, s)
        return KP, KQ

    def build(self, P_coef, Q_coef, occ, exchange_coef, targets=None):
        """Direct potentials and exchange sums of ``targets`` (default: all):
        a dict of device tensors VJ_PP, VJ_QQ, K_PP, K_QQ, each (n_targets, ns, ns)."""
        fP, fQ = self.samples(P_coef, Q_coef)
        n_orb = fP.shape[0]
        targets = list(range(n_orb)) if targets is None else [int(t) for t in targets]
        vpp, vqq = self.direct(fP, fQ, occ, targets)
        kpp, kqq = self.exchange(fP, fQ, exchange_coef, targets)
        return {"VJ_PP": vpp, "VJ_QQ": vqq, "K_PP": kpp, "K_QQ": kqq}

    @staticmethod
    def footprint(n_orb, n_targets, ns, n, nv, ng, ks, chunk=64):
        """Device bytes a build holds at its peak (admission estimate)."""
        return 8 * (3 * n_orb * n + 3 * n_targets * n + 5 * n_targets * ns * ns
                    + (chunk + 2) * ns * ns + 2 * nv * ng * ks + 2 * (ns + 1) * n)
