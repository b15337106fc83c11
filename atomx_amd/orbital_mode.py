"""The DF solver's "orbital" SCF mode on the GPU: per-target direct
potential, per-source cached exchange and the intra-shell potential
(methods/dirac_fock.py:840-921, :969-1068, :1469-1551, assembled as in
:2205-2227), for occupation-product input.

For target a the two-electron blocks are
  PP = V_J,PP - K_PP + V_i,PP,   QQ = V_J,QQ - K_QQ + V_i,QQ,
  PQ = -K_PQ,                    QP = -K_QP,
with V_J the k = 0 case-1 contraction of sum_b w_b rho_b (w_b = q_b, and
q_a - 1 for the target), K the symmetrised sum over sources b != a and
exchange ranks k of q_b Lambda_ab^k / ((2j_a+1)(2j_b+1)) times the case-4
contraction of b's own full densities (which depends on b and k only, so it
is computed once and kept until ``invalidate(b)``), and V_i the even-k
case-1 contractions of the target's own density weighted by
-(q_a - 1) Lambda_aa^k / (2j_a (2j_a+1)).  Contractions use the full banks,
as the reference's orbital mode does.  Every step is a device kernel; the
blocks come back to the host for the FP64 eigensolve.
"""

from __future__ import annotations

import numpy as np

from . import device as dv
from ._lib import call
from .angular import exchange_k_range, exchange_lambda
from .convol import ConvolCache, axpy

__all__ = ["OrbitalModeFock"]


def _s():
    return dv.stream_handle()


class OrbitalModeFock:
    def __init__(self, bsp_data, kappas, occupations, cache=None):
        bd = bsp_data
        self.bd = bd
        self.kappas = [int(k) for k in kappas]
        self.occ = [float(q) for q in occupations]
        self.n_orb = len(self.kappas)
        ks = bd["ks"]
        self.ksp, self.ksq = int(bd.get("ksp", ks)), int(bd.get("ksq", ks))
        self.nsp, self.nsq = int(bd.get("nsp", bd["ns"])), int(bd.get("nsq", bd["ns"]))
        self.ns = int(bd["ns"])
        self.cache = cache if cache is not None else ConvolCache(bd)
        self._K = {}          # (b, k) -> device (4, ns, ns): PP, QQ, PQ, QP contractions
        self._P = self._Q = None
        # exchange coefficients per target: list of (b, k, c)
        self.exchange_terms = []
        for a, ka in enumerate(self.kappas):
            tja = 2 * abs(ka) - 1
            terms = []
            for b, kb in enumerate(self.kappas):
                if b == a:
                    continue
                tjb = 2 * abs(kb) - 1
                for k in exchange_k_range(ka, kb):
                    lam = exchange_lambda(ka, kb, k)
                    if lam < 1e-15:
                        continue
                    terms.append((b, k, self.occ[b] * lam / ((tja + 1.0) * (tjb + 1.0))))
            self.exchange_terms.append(terms)
        # intra-shell (k, coefficient) per target
        self.intra_terms = []
        for a, ka in enumerate(self.kappas):
            tja = 2 * abs(ka) - 1
            qa = self.occ[a]
            terms = []
            if tja > 1 and qa * (qa - 1.0) >= 1e-15:
                la = ka if ka > 0 else -ka - 1
                for k in range(2, tja + 1):
                    if (k + 2 * la) % 2:
                        continue
                    lam = exchange_lambda(ka, ka, k)
                    if lam < 1e-15:
                        continue
                    terms.append((k, -(qa - 1.0) * lam / (tja * (tja + 1.0))))
            self.intra_terms.append(terms)

    # -- orbitals -----------------------------------------------------------------
    def load(self, P, Q):
        """New coefficient stacks (n_orb, >= ns); drops every cached
        exchange contraction (use ``update`` for a single orbital)."""
        self._P = dv.f64(self._pad(P))
        self._Q = dv.f64(self._pad(Q))
        self._K.clear()

    def update(self, b, Pb, Qb):
        """Replace orbital b and drop its cached exchange contractions
        (dirac_fock.py:2296-2300)."""
        self._P[b].copy_(dv.f64(self._pad(np.asarray(Pb)[None]))[0])
        self._Q[b].copy_(dv.f64(self._pad(np.asarray(Qb)[None]))[0])
        self.invalidate(b)

    def invalidate(self, b=None):
        if b is None:
            self._K.clear()
        else:
            for key in [key for key in self._K if key[0] == b]:
                del self._K[key]

    def _pad(self, A):
        A = np.atleast_2d(np.asarray(A, dtype=np.float64))
        out = np.zeros((A.shape[0], self.ns))
        m = min(A.shape[1], self.ns)
        out[:, :m] = A[:, :m]
        return out

    # -- pieces -----------------------------------------------------------------------
    def _banded(self, C, ks, rows):
        out = dv.empty((len(rows), self.ns, ks))
        idx = dv.i32(np.asarray(rows, dtype=np.int32))
        call("ax_density_sym", dv.ptr(C), self.ns, dv.ptr(idx), dv.ptr(idx), len(rows), self.ns,
             self.ns, self.ns, ks, dv.ptr(out), _s())
        return out

    def _to_full(self, V, vb, ks, coeff):
        call("ax_banded_to_full_add", dv.ptr(V), self.ns * self.ns, dv.ptr(vb), self.ns * ks,
             vb.shape[0], self.ns, ks, float(coeff), _s())

    def direct(self, target):
        """(V_J,PP, V_J,QQ) device (ns, ns) of the k = 0 direct potential."""
        ns, n = self.ns, self.n_orb
        coef = np.array([self.occ[j] - 1.0 if j == target else self.occ[j] for j in range(n)])
        coef[np.abs(coef) < 1e-10] = 0.0
        rows = list(range(n))
        dens = {}
        for c, C, ks in (("p", self._P, self.ksp), ("q", self._Q, self.ksq)):
            per = self._banded(C, ks, rows)
            acc = dv.empty((1, ns, ks))
            cd = dv.f64(coef[None])
            call("ax_combine_rows", dv.ptr(per), ns * ks, n, dv.ptr(cd), 1, dv.ptr(acc), _s())
            dens[c] = acc
        vpp, vqq = dv.zeros((1, ns, ns)), dv.zeros((1, ns, ns))
        ch = self.cache
        self._to_full(vpp, ch.case1(0, "pppp", dens["p"], force="bank"), self.ksp, 1.0)
        self._to_full(vpp, ch.case1(0, "pqpq", dens["q"], force="bank"), self.ksp, 1.0)
        self._to_full(vqq, ch.case1(0, "qpqp", dens["p"], force="bank"), self.ksq, 1.0)
        self._to_full(vqq, ch.case1(0, "qqqq", dens["q"], force="bank"), self.ksq, 1.0)
        return vpp[0], vqq[0]

    def _source(self, b, k):
        """Case-4 contractions of source b's full densities at rank k
        (computed once per (b, k))."""
        got = self._K.get((b, k))
        if got is None:
            ns = self.ns
            one = dv.f64(np.ones((1, 1)))
            x = {}
            for name, A, B in (("pp", self._P, self._P), ("qq", self._Q, self._Q),
                               ("pq", self._P, self._Q), ("qp", self._Q, self._P)):
                xx = dv.empty((1, ns, ns))
                call("ax_exchange_density", dv.ptr(A[b:b + 1]), dv.ptr(B[b:b + 1]), ns, 1, ns,
                     dv.ptr(one), 1, dv.ptr(xx), _s())
                x[name] = xx
            got = dv.empty((4, ns, ns))
            ch = self.cache
            for i, (name, itype) in enumerate((("pp", "pppp"), ("qq", "qqqq"), ("pq", "pqpq"),
                                               ("qp", "qpqp"))):
                got[i].copy_(ch.case4(k, itype, x[name], False, force="bank")[0])
            self._K[(b, k)] = got
        return got

    def exchange(self, target):
        """(K_PP, K_QQ, K_PQ, K_QP) device (ns, ns), symmetrised."""
        ns = self.ns
        K = dv.zeros((4, ns, ns))
        for b, k, c in self.exchange_terms[target]:
            axpy(K, self._source(b, k), c)
        call("ax_fock_finalize", dv.ptr(K[0]), dv.ptr(K[1]), dv.ptr(K[2]), dv.ptr(K[3]), ns, 1.0,
             0, _s())
        return K[0], K[1], K[2], K[3]

    def intra_shell(self, target):
        """(V_i,PP, V_i,QQ) device (ns, ns); zero for j = 1/2 shells."""
        ns = self.ns
        vpp, vqq = dv.zeros((1, ns, ns)), dv.zeros((1, ns, ns))
        if not self.intra_terms[target]:
            return vpp[0], vqq[0]
        dp = self._banded(self._P, self.ksp, [target])
        dq = self._banded(self._Q, self.ksq, [target])
        ch = self.cache
        for k, c in self.intra_terms[target]:
            self._to_full(vpp, ch.case1(k, "pppp", dp, force="bank"), self.ksp, c)
            self._to_full(vqq, ch.case1(k, "qpqp", dp, force="bank"), self.ksq, c)
            self._to_full(vqq, ch.case1(k, "qqqq", dq, force="bank"), self.ksq, c)
            self._to_full(vpp, ch.case1(k, "pqpq", dq, force="bank"), self.ksp, c)
        return vpp[0], vqq[0]

    def blocks_device(self, target):
        vpp, vqq = self.direct(target)
        kpp, kqq, kpq, kqp = self.exchange(target)
        ipp, iqq = self.intra_shell(target)
        axpy(vpp, kpp, -1.0)
        axpy(vpp, ipp, 1.0)
        axpy(vqq, kqq, -1.0)
        axpy(vqq, iqq, 1.0)
        pq, qp = dv.zeros(kpq.shape), dv.zeros(kqp.shape)
        axpy(pq, kpq, -1.0)
        axpy(qp, kqp, -1.0)
        return vpp, vqq, pq, qp

    def blocks(self, target, P=None, Q=None):
        """Host (PP, QQ, PQ, QP) two-electron blocks of target's Fock matrix;
        P, Q given: load them first."""
        if P is not None:
            self.load(P, Q)
        return tuple(dv.host(x) for x in self.blocks_device(target))
