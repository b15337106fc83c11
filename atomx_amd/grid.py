"""Device-resident grid tables shared by every (Y)-path call on one grid.

A grid is identified by the content of its flattened nodes and weights
(like the reference's content-keyed bank registry, convol.py:1415-1475): the
tables (r, w, r^k, r^-(k+1) for k <= kmax, partial matrices) are uploaded and
built once and reused by later calls.
"""

from __future__ import annotations

import hashlib
from collections import OrderedDict

import numpy as np

from . import device as dv
from ._lib import call, load

_CACHE: "OrderedDict[tuple, GridTables]" = OrderedDict()
_CACHE_LIMIT = 8

_PARTIAL: dict = {}


def partial_matrices(n_gauss: int):
    """(mo, mi) host arrays of the interpolatory partial-interval rule."""
    mats = _PARTIAL.get(n_gauss)
    if mats is None:
        import ctypes
        mo = np.empty((n_gauss, n_gauss))
        mi = np.empty((n_gauss, n_gauss))
        dp = ctypes.POINTER(ctypes.c_double)
        call("ax_yk_partial_matrices", int(n_gauss), mo.ctypes.data_as(dp), mi.ctypes.data_as(dp))
        mats = (mo, mi)
        _PARTIAL[n_gauss] = mats
    return mats


class GridTables:
    def __init__(self, r, w, nv, ng, kmax):
        dv.require_gpu()
        load()
        self.nv, self.ng, self.n = int(nv), int(ng), int(nv) * int(ng)
        self.kmax = -1
        self.r = dv.f64(r)
        self.w = dv.f64(w)
        mo, mi = partial_matrices(self.ng)
        self.mo = dv.f64(mo)
        self.mi = dv.f64(mi)
        self.rk = self.rkm = None
        self.ensure_k(kmax)

    def ensure_k(self, kmax):
        if kmax <= self.kmax:
            return
        kmax = max(int(kmax), 2 * self.kmax + 1 if self.kmax >= 0 else int(kmax))
        rk = dv.empty((kmax + 1, self.n))
        rkm = dv.empty((kmax + 1, self.n))
        call("ax_grid_rpow", dv.ptr(self.r), self.n, kmax, dv.ptr(rk), dv.ptr(rkm),
             dv.stream_handle())
        self.rk, self.rkm, self.kmax = rk, rkm, kmax


def get(r_flat, w_flat, nv, ng, kmax=0) -> GridTables:
    r_flat = np.ascontiguousarray(r_flat, dtype=np.float64).ravel()
    w_flat = np.ascontiguousarray(w_flat, dtype=np.float64).ravel()
    if r_flat.size != int(nv) * int(ng) or w_flat.size != r_flat.size:
        raise ValueError(f"grid has {r_flat.size} nodes, expected nv*n_gauss = {nv}*{ng}")
    key = (int(nv), int(ng), hashlib.sha1(r_flat.tobytes()).hexdigest(),
           hashlib.sha1(w_flat.tobytes()).hexdigest())
    tab = _CACHE.get(key)
    if tab is None:
        tab = GridTables(r_flat, w_flat, nv, ng, kmax)
        _CACHE[key] = tab
        while len(_CACHE) > _CACHE_LIMIT:
            _CACHE.popitem(last=False)
    else:
        _CACHE.move_to_end(key)
        tab.ensure_k(kmax)
    return tab
