"""Batched all-quartet R^k tensor on one grid (device-resident plan).

Wraps ``ax_rk_tensor_*``: for orbital samples P, Q (n_orb, n) on the grid it
returns every R^k(ab;cd), k = 0..kmax, over pair-pairs valid by the triangle
and parity rules — R^k(ab;cd) = sum_x grw rho_ab Y^k[rho_cd] for pair(ab) <=
pair(cd) (numerics/bspline/slater.py:118-130 per integral).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import device as dv
from ._lib import PRECISIONS, call, check, load
from .workloads import pair_list, valid_pairs


class SlaterTensor:
    def __init__(self, gr, grw, ls, kmax, precision="fp64", pairs=None, n_fn=None):
        """All-pairs plan of orbitals with angular momenta ``ls`` (angular
        rule per rank), or, with ``pairs=(pa, pb)`` and ``n_fn`` functions, a
        pair-product plan of that explicit list valid at every rank."""
        dv.require_gpu()
        lib = load()
        gr = np.ascontiguousarray(gr, dtype=np.float64)
        grw = np.ascontiguousarray(grw, dtype=np.float64)
        if gr.ndim != 2 or gr.shape != grw.shape:
            raise ValueError("grid and weights must both be (nv, n_gauss)")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {tuple(PRECISIONS)}")
        self.nv, self.ng = gr.shape
        self.n = self.nv * self.ng
        self.kmax = int(kmax)
        self.precision = precision
        h = ctypes.c_void_p()
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        if pairs is None:
            self.ls = np.ascontiguousarray(ls, dtype=np.int32)
            self.n_orb = int(self.ls.size)
            self._pairs = None
            call("ax_rk_tensor_create", gr.ravel().ctypes.data_as(dp),
                 grw.ravel().ctypes.data_as(dp), self.nv, self.ng, self.n_orb,
                 self.ls.ctypes.data_as(ip), self.kmax, PRECISIONS[precision], ctypes.byref(h))
        else:
            pa = np.ascontiguousarray(pairs[0], dtype=np.int32)
            pb = np.ascontiguousarray(pairs[1], dtype=np.int32)
            if pa.shape != pb.shape or pa.ndim != 1:
                raise ValueError("pairs must be two equal-length index vectors")
            self.ls = None
            self.n_orb = int(n_fn)
            self._pairs = (pa, pb)
            call("ax_rk_tensor_create_pairs", gr.ravel().ctypes.data_as(dp),
                 grw.ravel().ctypes.data_as(dp), self.nv, self.ng, self.n_orb, int(pa.size),
                 pa.ctypes.data_as(ip), pb.ctypes.data_as(ip), self.kmax, PRECISIONS[precision],
                 ctypes.byref(h))
        self._attach(lib, h)

    def _attach(self, lib, h):
        self._h = h
        self._lib = lib
        k1 = self.kmax + 1
        size, slices = ctypes.c_int64(), ctypes.c_int32()
        pairs, offs = (ctypes.c_int32 * k1)(), (ctypes.c_int64 * k1)()
        rc = lib.ax_rk_tensor_describe(h, ctypes.byref(size), ctypes.byref(slices), pairs, offs)
        if rc:
            check(rc, "ax_rk_tensor_describe")
        self.size = size.value
        self.m = list(pairs)
        self.offsets = list(offs)
        self.slices = slices.value

    @classmethod
    def create_many(cls, specs, precision="fp64", stream=None):
        """All-pairs plans for ``specs`` = [(gr, grw, ls, kmax), ...] in one
        library call (one staging upload, batched scatter and r-power tables,
        one shared set-up event); each plan is then used and closed on its
        own.  ``stream``: the torch stream to queue the set-up on (the one
        the plans will run on, say); None = the library's own set-up stream
        of the calling thread."""
        dv.require_gpu()
        lib = load()
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {tuple(PRECISIONS)}")
        n = len(specs)
        if n == 0:
            return []
        objs, keep = [], []
        ptrs = [], [], []  # grid, weights, l addresses
        nvs, ngs, nos, kms = [], [], [], []
        f64, i32 = np.float64, np.int32
        for gr, grw, ls, kmax in specs:
            gr = np.ascontiguousarray(gr, dtype=f64)
            grw = np.ascontiguousarray(grw, dtype=f64)
            if gr.ndim != 2 or gr.shape != grw.shape:
                raise ValueError("grid and weights must both be (nv, n_gauss)")
            o = cls.__new__(cls)
            o.nv, o.ng = gr.shape
            o.n = o.nv * o.ng
            o.kmax = int(kmax)
            o.precision = precision
            o.ls = np.ascontiguousarray(ls, dtype=i32)
            o.n_orb = int(o.ls.size)
            o._pairs = None
            keep += [gr, grw]
            for lst, a in zip(ptrs, (gr, grw, o.ls)):
                lst.append(a.__array_interface__["data"][0])
            nvs.append(o.nv)
            ngs.append(o.ng)
            nos.append(o.n_orb)
            kms.append(o.kmax)
            objs.append(o)
        vp, ia = ctypes.c_void_p * n, ctypes.c_int32 * n
        hs = (ctypes.c_void_p * n)()
        call("ax_rk_tensor_create_many", n, vp(*ptrs[0]), vp(*ptrs[1]), ia(*nvs), ia(*ngs),
             ia(*nos), vp(*ptrs[2]), ia(*kms), ia(*([PRECISIONS[precision]] * n)),
             None if stream is None else dv.stream_handle(stream), hs)
        for o, h in zip(objs, hs):
            o._attach(lib, ctypes.c_void_p(h))
        return objs

    @property
    def footprint(self):
        """Device bytes the plan holds now (tuning or a schedule of another
        K split adds that split's job tables, and partial workspace when it
        needs more)."""
        return int(self._lib.ax_rk_tensor_footprint(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ax_rk_tensor_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, P, Q, out=None, stream=None):
        """P, Q device tensors (n_orb, n) FP64 (or host arrays); returns the
        packed output (device tensor of ``size`` values)."""
        P, Q = self.check_inputs(P, Q)
        if out is None:
            out = dv.empty((self.size,))
        call("ax_rk_tensor_run", self._h, dv.ptr(P), dv.ptr(Q), dv.ptr(out),
             dv.stream_handle(stream))
        return out

    STAGES = ("density", "scan", "slice", "contract", "reduce")
    # block orders of the contraction launch: three static job orders, one
    # round of resident work-groups taking jobs heaviest first from a ticket
    # (FP64), and one K slice per XCD with its tiles packed heaviest first
    # onto the least-loaded work-groups
    ORDERS = ("tile-major", "xcd-contiguous", "xcd-contiguous-split-fastest",
              "dynamic-heaviest-first", "balanced-k-ranges")

    def schedule(self):
        """(block order name, K split) of the contraction."""
        o, s = ctypes.c_int(), ctypes.c_int()
        call("ax_rk_tensor_get_schedule", self._h, ctypes.byref(o), ctypes.byref(s))
        return self.ORDERS[o.value], s.value

    def set_schedule(self, order, n_split):
        call("ax_rk_tensor_set_schedule", self._h, self.ORDERS.index(order), int(n_split))

    def tune(self, P, Q, out=None, reps=3, stream=None):
        """Times every contraction schedule on these inputs, keeps the fastest
        and returns (order, split, ms per contraction + reduction); ``out``
        then holds the tuned result."""
        P, Q = self.check_inputs(P, Q)
        if out is None:
            out = dv.empty((self.size,))
        ms = ctypes.c_float()
        call("ax_rk_tensor_tune", self._h, dv.ptr(P), dv.ptr(Q), dv.ptr(out), int(reps),
             dv.stream_handle(stream), ctypes.byref(ms))
        return self.schedule() + (float(ms.value),)

    def tune_raw(self, P_ptr, Q_ptr, out_ptr, reps=3, stream=None):
        """``tune`` on raw device addresses (Q_ptr None or 0 for no small
        component) whose shapes the caller has checked; synchronises the
        stream."""
        ms = ctypes.c_float()
        call("ax_rk_tensor_tune", self._h, P_ptr, Q_ptr or None, out_ptr, int(reps),
             dv.stream_handle(stream), ctypes.byref(ms))
        return self.schedule() + (float(ms.value),)

    def check_inputs(self, P, Q):
        P = dv.f64(P)
        Q = dv.f64(Q) if Q is not None else None
        if P.shape != (self.n_orb, self.n) or (Q is not None and Q.shape != P.shape):
            raise ValueError(f"orbitals must be ({self.n_orb}, {self.n})")
        return P, Q

    def run_stage(self, stage, P=None, Q=None, out=None, stream=None):
        """Launches one stage of ``run`` (per-kernel timing); P/Q/out must be
        the same device tensors as in the run being repeated."""
        call("ax_rk_tensor_run_stage", self._h, self.STAGES.index(stage), dv.ptr(P), dv.ptr(Q),
             dv.ptr(out), dv.stream_handle(stream))

    # -- host-side views of the packed output -----------------------------
    def rank_pairs(self, k):
        if self._pairs is not None:
            return np.arange(self._pairs[0].size, dtype=np.int64)
        return valid_pairs(self.ls, k)

    def quartets(self, k):
        """(p, q) pair indices of rank k's packed values (p <= q in V_k order)."""
        V = self.rank_pairs(k)
        i, j = np.triu_indices(len(V))
        return V[i], V[j]

    def unpack(self, out_host, k):
        """Rank k's values as a symmetric (m_k, m_k) matrix over its valid pairs."""
        m = self.m[k]
        vals = out_host[self.offsets[k]:self.offsets[k] + m * (m + 1) // 2]
        M = np.zeros((m, m))
        i, j = np.triu_indices(m)
        M[i, j] = vals
        M[j, i] = vals
        return M

    def pair_indices(self):
        """(pa, pb): the functions of every pair, in output order."""
        return self._pairs if self._pairs is not None else pair_list(self.n_orb)

    @staticmethod
    def pairs(n_orb):
        return pair_list(n_orb)


def run_batch(plans, Ps, Qs, outs=None, stream=None):
    """Runs several plans (different grids, orbital counts and k ranges, one
    precision mode) as one set of five launches.  Returns the list of packed
    outputs; each equals bitwise what ``plan.run`` alone returns."""
    if not plans:
        return []
    if len(Ps) != len(plans) or len(Qs) != len(plans):
        raise ValueError("one (P, Q) per plan")
    if outs is None:
        outs = [dv.empty((p.size,)) for p in plans]
    ins = [p.check_inputs(P, Q) for p, P, Q in zip(plans, Ps, Qs)]
    nb = len(plans)
    arr = ctypes.c_void_p * nb
    h = arr(*[p._h.value for p in plans])
    pp = arr(*[dv.ptr(P) for P, _ in ins])
    qq = arr(*[dv.ptr(Q) if Q is not None else None for _, Q in ins])
    oo = arr(*[dv.ptr(o) for o in outs])
    call("ax_rk_tensor_run_batch", h, nb, pp, qq, oo, dv.stream_handle(stream))
    return outs


def run_batch_raw(plans, P_ptrs, Q_ptrs, out_ptrs, stream=None):
    """``run_batch`` on raw device addresses (int, or None for a missing Q)
    whose shapes the caller has checked: (n_orb, n) FP64 orbitals and a
    ``size``-long output per plan."""
    nb = len(plans)
    if nb == 0:
        return
    if not (len(P_ptrs) == len(Q_ptrs) == len(out_ptrs) == nb):
        raise ValueError("one (P, Q, out) per plan")
    arr = ctypes.c_void_p * nb
    call("ax_rk_tensor_run_batch", arr(*[p._h.value for p in plans]), nb, arr(*P_ptrs),
         arr(*Q_ptrs), arr(*out_ptrs), dv.stream_handle(stream))


def run_batch_arrays(handles, n, P_addr, Q_addr, out_addr, stream=None):
    """``run_batch_raw`` from prebuilt tables: a ctypes array of n plan
    handles and numpy uint64 arrays of n device addresses (0 for a missing
    Q); the caller has checked every shape."""
    if n == 0:
        return
    P_addr = np.ascontiguousarray(P_addr, dtype=np.uint64)
    Q_addr = np.ascontiguousarray(Q_addr, dtype=np.uint64)
    out_addr = np.ascontiguousarray(out_addr, dtype=np.uint64)
    if not (P_addr.size == Q_addr.size == out_addr.size == n):
        raise ValueError("one (P, Q, out) address per plan")
    vpp = ctypes.POINTER(ctypes.c_void_p)
    call("ax_rk_tensor_run_batch", handles, n, P_addr.ctypes.data_as(vpp),
         Q_addr.ctypes.data_as(vpp), out_addr.ctypes.data_as(vpp), dv.stream_handle(stream))
