"""Radial integrals of the exact-CI Hamiltonian on the GPU (drop-in for the
B-spline branch of physics/csf/matrix.py:191-318 and :703-817).

* ``fk`` / ``gk`` / ``rk``: memoised F^k, G^k, R^k of orbital coefficient
  arrays (matrix.py:191-255), keyed by the arrays' identity for the life of
  a CI build (the arrays are kept alive by the memo).
* ``bspline_pair_potential``: the ordinary CI cell operator on one banded
  pair density — four case-1 contractions (matrix.py:703-720).
* ``RadialEvaluator`` / ``radial_evaluator``: exact-CI radial keys to values
  (matrix.py:723-817).  ``("h", n_a, kappa, n_c)`` is <a|h_D|c>/<c|c> from
  the FP64 one-body Dirac blocks; ``("r", k, a, c, b, d)`` is R^k(ac, bd) =
  the banded dot of the pair potential of the sorted source pair (b, d)
  with the banded density of the sorted target pair (a, c).
  ``evaluate_many`` takes a whole key list: per rank it builds every new
  source density and pair potential in one batched launch set and every
  value in one gathered banded dot, so a CI build's integrals cost a few
  launches per rank instead of one evaluation per key.
* ``breit_rk`` / ``breit_rk_many``: component-resolved Slater integrals of
  the Gaunt operator on the Gauss grid (matrix.py:266-318), symmetrised over
  electron interchange.
"""

from __future__ import annotations

from collections import OrderedDict

import numpy as np

from . import device as dv
from . import grid as _grid
from ._lib import call
from .convol import ConvolCache, axpy, density_sym_device
from .dirac import dirac_matrices
from .slater_yk import eval_orbitals_device
from .yk import yk_rows

__all__ = ["MissingOrbitalError", "fk", "gk", "rk", "clear_memo", "bspline_pair_potential",
           "pair_potentials_device", "RadialEvaluator", "radial_evaluator", "breit_rk",
           "breit_rk_many"]


class MissingOrbitalError(ValueError):
    """A CI element needs an orbital that is not in the orbital set."""


def _s():
    return dv.stream_handle()


# ---------------------------------------------------------------------------
# memoised F^k / G^k / R^k (matrix.py:191-255)
# ---------------------------------------------------------------------------
_MEMO = {}
_KEEP = []


def clear_memo():
    """Drop the memo (call when the orbitals of a CI build change)."""
    _MEMO.clear()
    _KEEP.clear()


def _memo_cache(bsp_data):
    cache = _MEMO.get("_cache")
    if cache is None or not cache.same_grid(bsp_data):
        clear_memo()
        cache = ConvolCache(bsp_data)
        _MEMO["_cache"] = cache
    return cache


def fk(Pa, Qa, Pb, Qb, bsp_data, k):
    from .convol import compute_slater_fk_cell
    cache = _memo_cache(bsp_data)
    ia, ib = id(Pa), id(Pb)
    key = ("f", k) + ((ia, ib) if ia <= ib else (ib, ia))
    v = _MEMO.get(key)
    if v is None:
        _KEEP.extend((Pa, Pb))
        v = compute_slater_fk_cell(Pa, Qa, Pb, Qb, bsp_data, k, cache=cache)
        _MEMO[key] = v
    return v


def gk(Pa, Qa, Pb, Qb, bsp_data, k):
    from .convol import compute_slater_gk_cell
    cache = _memo_cache(bsp_data)
    ia, ib = id(Pa), id(Pb)
    key = ("g", k) + ((ia, ib) if ia <= ib else (ib, ia))
    v = _MEMO.get(key)
    if v is None:
        _KEEP.extend((Pa, Pb))
        v = compute_slater_gk_cell(Pa, Qa, Pb, Qb, bsp_data, k, cache=cache)
        _MEMO[key] = v
    return v


def rk(Pa, Qa, Pb, Qb, Pc, Qc, Pd, Qd, bsp_data, k):
    """R^k(ab, cd): the first two orbitals form the first pair."""
    from .convol import compute_slater_rk_cell
    cache = _memo_cache(bsp_data)
    key = ("r", k, id(Pa), id(Pb), id(Pc), id(Pd))
    v = _MEMO.get(key)
    if v is None:
        _KEEP.extend((Pa, Pb, Pc, Pd))
        v = compute_slater_rk_cell(Pa, Qa, Pb, Qb, Pc, Qc, Pd, Qd, bsp_data, k, cache=cache)
        _MEMO[key] = v
    return v


# ---------------------------------------------------------------------------
# pair potentials (matrix.py:703-720)
# ---------------------------------------------------------------------------
def pair_potentials_device(cache, rank, dens_p, dens_q):
    """Cell potentials of nb source pair densities (banded, device
    (nb, ns, ksp) and (nb, ns, ksq)): (V_p (nb, ns, ksp), V_q (nb, ns, ksq)),
    V_left = sum over right components of case1(left right left right)."""
    vp = cache.case1(rank, "pppp", dens_p)
    axpy(vp, cache.case1(rank, "pqpq", dens_q))
    vq = cache.case1(rank, "qpqp", dens_p)
    axpy(vq, cache.case1(rank, "qqqq", dens_q))
    return vp, vq


def bspline_pair_potential(cache, ns, orders, rank, source):
    """The ordinary CI cell operator on one explicit pair-density tuple
    (host (ns, ksp), (ns, ksq) arrays) -> host potentials."""
    dp = dv.f64(np.asarray(source[0], dtype=np.float64)[None])
    dq = dv.f64(np.asarray(source[1], dtype=np.float64)[None])
    vp, vq = pair_potentials_device(cache, int(rank), dp, dq)
    return dv.host(vp)[0], dv.host(vq)[0]


# ---------------------------------------------------------------------------
# the CI radial evaluator (matrix.py:723-817)
# ---------------------------------------------------------------------------
def _ptr_table(tensors):
    """Device table of the base addresses of a list of device tensors."""
    addr = np.array([t.data_ptr() for t in tensors], dtype=np.int64)
    return dv.to_dev(addr)


class RadialEvaluator:
    """Exact-CI radial keys -> values.  The orbitals, geometry, Z and
    nuclear model are fixed for the evaluator's lifetime (copies are
    taken)."""

    def __init__(self, orbitals_dict, bsp_data, Z, ns, *, nuclear_potential=None,
                 pair_potential_cache_size=None):
        cap = 1024 if pair_potential_cache_size is None else pair_potential_cache_size
        if not isinstance(cap, int) or isinstance(cap, bool) or cap < 1:
            raise ValueError("pair potential cache capacity must be a positive integer")
        self.ns = int(ns)
        self.Z = Z
        self.bsp_data = bsp_data
        self.orders = (int(bsp_data.get("ksp", bsp_data["ks"])),
                       int(bsp_data.get("ksq", bsp_data["ks"])))
        self._keys = sorted(orbitals_dict)
        self._index = {key: i for i, key in enumerate(self._keys)}
        n = len(self._keys)
        P = np.zeros((max(n, 1), self.ns))
        Q = np.zeros((max(n, 1), self.ns))
        for key, i in self._index.items():
            p, q = (np.asarray(x, dtype=np.float64).ravel() for x in orbitals_dict[key])
            P[i, :min(p.size, self.ns)] = p[:self.ns]
            Q[i, :min(q.size, self.ns)] = q[:self.ns]
        self._P, self._Q = P, Q
        self._one_body = {}
        for _, kappa in self._keys:
            if kappa not in self._one_body:
                self._one_body[kappa] = dirac_matrices(kappa, Z, bsp_data, nuclear_potential)
        self.cache = ConvolCache(bsp_data)
        self._Pd, self._Qd = dv.f64(P), dv.f64(Q)
        self._cap = cap
        self._pot = OrderedDict()       # (k, i, j) -> (V_p row, V_q row) device views
        self._dens = OrderedDict()      # (i, j) -> (d_p row, d_q row) device views
        self._dens_cap = 512
        self._values = OrderedDict()
        self._values_cap = 8192
        self.hits = self.misses = 0

    def _idx(self, key):
        try:
            return self._index[key]
        except KeyError:
            raise MissingOrbitalError(
                f"CI matrix needs orbital (n={key[0]}, kappa={key[1]}), which was not solved; "
                f"available orbitals: {self._keys}") from None

    def pair_cache_info(self):
        return {"entries": len(self._pot), "capacity": self._cap}

    def __call__(self, key):
        return self.evaluate_many([key])[0]

    def _one(self, key):
        _, n_a, kappa, n_c = key
        ia, ic = self._idx((n_a, kappa)), self._idx((n_c, kappa))
        h0, overlap = self._one_body[kappa]
        bra = np.concatenate((self._P[ia], self._Q[ia]))
        ket = np.concatenate((self._P[ic], self._Q[ic]))
        h = bra @ h0 @ ket
        norm_sq = ket @ overlap @ ket
        return h / norm_sq if norm_sq > 0.0 else float("nan")

    def _densities(self, pairs):
        """Device banded densities of orbital-index pairs (cached, LRU 512)."""
        new = [p for p in dict.fromkeys(pairs) if p not in self._dens]
        if new:
            dp = density_sym_device(self._Pd, new, self.ns, self.orders[0])
            dq = density_sym_device(self._Qd, new, self.ns, self.orders[1])
            for r, p in enumerate(new):
                self._dens[p] = (dp[r], dq[r])
        out = []
        for p in pairs:
            self._dens.move_to_end(p)
            out.append(self._dens[p])
        while len(self._dens) > max(self._dens_cap, len(set(pairs))):
            self._dens.popitem(last=False)
        return out

    def _potentials(self, k, pairs):
        new = [p for p in dict.fromkeys(pairs) if (k,) + p not in self._pot]
        if new:
            dens = self._densities(new)
            torch = dv.torch()
            dp = torch.stack([d[0] for d in dens])
            dq = torch.stack([d[1] for d in dens])
            vp, vq = pair_potentials_device(self.cache, k, dp, dq)
            for r, p in enumerate(new):
                self._pot[(k,) + p] = (vp[r], vq[r])
        out = []
        for p in pairs:
            self._pot.move_to_end((k,) + p)
            out.append(self._pot[(k,) + p])
        while len(self._pot) > max(self._cap, len(set(pairs))):
            self._pot.popitem(last=False)
        return out

    def evaluate_many(self, keys):
        """Values of a list of keys (np.ndarray, float)."""
        keys = [tuple(k) for k in keys]
        out = np.empty(len(keys))
        todo = {}
        for i, key in enumerate(keys):
            v = self._values.get(key)
            if v is not None:
                self.hits += 1
                self._values.move_to_end(key)
                out[i] = v
                continue
            self.misses += 1
            if key[0] == "h":
                out[i] = self._remember(key, self._one(key))
            elif key[0] == "r":
                todo.setdefault(int(key[1]), []).append(i)
            else:
                raise ValueError(f"unknown radial key {key!r}")
        for k, idxs in todo.items():
            if k < 0:
                raise ValueError("multipole rank must be a nonnegative integer")
            src, tgt = [], []
            for i in idxs:
                _, _, ka, kc, kb, kd = keys[i]
                a, c, b, d = (self._idx(x) for x in (ka, kc, kb, kd))
                src.append(tuple(sorted((b, d), key=lambda t: self._keys[t])))
                tgt.append(tuple(sorted((a, c), key=lambda t: self._keys[t])))
            pots = self._potentials(k, src)
            dens = self._densities(tgt)
            vals = dv.empty((len(idxs),))
            tables = []
            for comp in range(2):
                ta = _ptr_table([p[comp] for p in pots])
                tb = _ptr_table([d[comp] for d in dens])
                tables.append((ta, tb))
                call("ax_sum_amb_ptr", dv.ptr(ta), dv.ptr(tb), len(idxs), self.ns,
                     self.orders[comp], comp, dv.ptr(vals), _s())
            host = dv.host(vals)
            for r, i in enumerate(idxs):
                if not np.isfinite(host[r]):
                    raise FloatingPointError(f"non-finite Slater integral {keys[i]}")
                out[i] = self._remember(keys[i], float(host[r]))
        return out

    def _remember(self, key, value):
        self._values[key] = value
        while len(self._values) > self._values_cap:
            self._values.popitem(last=False)
        return value


def radial_evaluator(orbitals_dict, bsp_data, Z, ns, *, nuclear_potential=None,
                     pair_potential_cache_size=None):
    """The evaluator of matrix.py:723 (callable on one key; ``evaluate_many``
    for a list)."""
    return RadialEvaluator(orbitals_dict, bsp_data, Z, ns, nuclear_potential=nuclear_potential,
                           pair_potential_cache_size=pair_potential_cache_size)


# ---------------------------------------------------------------------------
# Breit (Gaunt) component-resolved R^L (matrix.py:266-318)
# ---------------------------------------------------------------------------
def breit_rk_many(items, bsp_data):
    """items: sequence of (f1, f2, f3, f4, comps, L) with coefficient vectors
    f and comps a 4-string/tuple of "P"/"Q";
    value = 1/2 sum_x w [rho1 Y^L(rho2) + rho2 Y^L(rho1)],
    rho1 = f1 f3, rho2 = f2 f4 on the Gauss grid.  Returns np.ndarray."""
    bd = bsp_data
    gr, grw = np.asarray(bd["gr"]), np.asarray(bd["grw"])
    nv, ng = gr.shape
    ksp, nsp = bd.get("ksp", bd["ks"]), bd.get("nsp", bd["ns"])
    ksq, nsq = bd.get("ksq", bd["ks"]), bd.get("nsq", bd["ns"])
    # distinct component functions, large components first
    seen = {}
    for it in items:
        for x, c in zip(it[:4], it[4]):
            if c not in ("P", "Q"):
                raise ValueError(f"component must be 'P' or 'Q', got {c!r}")
            seen.setdefault((id(x), c), x)
    p_keys = [key for key in seen if key[1] == "P"]
    q_keys = [key for key in seen if key[1] == "Q"]
    funcs = {key: i for i, key in enumerate(p_keys + q_keys)}
    F = dv.empty((len(funcs), nv * ng))
    for keys_c, lo, table, ks, ns in ((p_keys, 0, bd["bsp_p"], ksp, nsp),
                                      (q_keys, len(p_keys), bd.get("bsp_q", bd["bsp_p"]), ksq,
                                       nsq)):
        if keys_c:
            coef = np.zeros((len(keys_c), ns))
            for r, key in enumerate(keys_c):
                v = np.asarray(seen[key], dtype=np.float64).ravel()
                coef[r, :min(v.size, ns)] = v[:ns]
            F[lo:lo + len(keys_c)].copy_(eval_orbitals_device(coef, table, nv, ks, ns))
    # densities rho1 = f1 f3 and rho2 = f2 f4 per item
    pa, pb, Ls = [], [], []
    for it in items:
        f, comps, L = it[:4], it[4], int(it[5])
        ix = [funcs[(id(x), c)] for x, c in zip(f, comps)]
        pa += [ix[0], ix[1]]
        pb += [ix[2], ix[3]]
        Ls += [L, L]
    n_it = len(items)
    rho = dv.empty((2 * n_it, nv * ng))
    pad, pbd = dv.i32(pa), dv.i32(pb)
    call("ax_pair_density", dv.ptr(F), None, nv * ng, nv * ng, 2 * n_it, dv.ptr(pad),
         dv.ptr(pbd), None, dv.ptr(rho), nv * ng, _s())
    tables = _grid.get(gr.ravel(), grw.ravel(), nv, ng, max(Ls) if Ls else 0)
    Y = yk_rows(rho, tables, Ls)
    D = dv.empty((2 * n_it, nv * ng))
    call("ax_mul_rows", dv.ptr(D), dv.ptr(rho), dv.ptr(tables.w), 2 * n_it, nv * ng, _s())
    # rho1 . Y(rho2) and rho2 . Y(rho1)
    p1 = np.arange(0, 2 * n_it, 2, dtype=np.int32)
    p2 = p1 + 1
    d12, d21 = dv.empty((n_it,)), dv.empty((n_it,))
    for out, p, q in ((d12, p1, p2), (d21, p2, p1)):
        pd_, qd_ = dv.i32(p), dv.i32(q)
        call("ax_rk_dot", dv.ptr(D), nv * ng, dv.ptr(Y), nv * ng, nv * ng, n_it, dv.ptr(pd_),
             dv.ptr(qd_), dv.ptr(out), _s())
    val = dv.zeros((n_it,))
    axpy(val, d12, 0.5)
    axpy(val, d21, 0.5)
    return dv.host(val)


def breit_rk(f1, f2, f3, f4, comps, bsp_data, L):
    """int int f1 f3 (r_<^L / r_>^(L+1)) f2 f4, symmetric in the electrons."""
    return float(breit_rk_many([(f1, f2, f3, f4, tuple(comps), L)], bsp_data)[0])
