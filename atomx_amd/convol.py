"""DBSR_HF cell algorithm for R^k on the GPU — drop-in for numerics/bspline/convol.py.

The pair x pair R^k banks (and their semiseparable factors) are built on the
device from cell moments and diagonal-cell double integrals and stay resident
in HBM, shared process-wide per grid content (the reference's content-keyed
bank registry, convol.py:1380-1475).  Every contraction (direct case 1,
left-pair case 2, exchange case 4, and the semiseparable forms) runs as a HIP
kernel batched over densities that share a bank.

Public names and argument orders follow the reference module; the
numpy-in/numpy-out private kernels (``_convol_case1_t`` etc.) upload their
operands, run the device kernel and return host arrays.
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict
from collections.abc import Mapping

import numpy as np

from . import device as dv
from ._lib import call, load

SEMISEPARABLE_DIRECT_NS = 96          # convol.py:1411
SEMISEPARABLE_EXCHANGE_NS = 320       # convol.py:1403
SEMISEPARABLE_MIXED_EXCHANGE_NS = 320  # convol.py:1406
_TRANSPOSE_ONLY_NS = 256              # convol.py:1400
_BANK_CACHE_LIMIT = 4                 # convol.py:1394
_RKB_BANKS: "OrderedDict[tuple, dict]" = OrderedDict()

_COMPONENTS = {"pppp": ("p", "p"), "qqqq": ("q", "q"), "pqpq": ("p", "q"), "qpqp": ("q", "p")}

__all__ = [
    "ConvolCache", "grid_bank_key", "shared_banks",
    "compute_slater_rk_cell", "compute_slater_fk_cell", "compute_slater_gk_cell",
    "compute_slater_rk_cell_batch", "compute_direct_potential_cell",
    "compute_exchange_matrix_cell", "build_slater_rk_bilinear", "eval_slater_rk_bilinear",
    "compute_slater_rk_table",
    "SEMISEPARABLE_DIRECT_NS", "SEMISEPARABLE_EXCHANGE_NS", "SEMISEPARABLE_MIXED_EXCHANGE_NS",
]


def _s():
    return dv.stream_handle()


# ---------------------------------------------------------------------------
# bank registry
# ---------------------------------------------------------------------------

def grid_bank_key(bsp_data):
    """Content key of the banks built on ``bsp_data`` (convol.py:1415-1452)."""
    t = bsp_data.get("t") if isinstance(bsp_data, Mapping) else None
    if t is None:
        raise TypeError(
            "ConvolCache needs a B-spline workspace mapping carrying a 't' knot vector "
            f"(got {type(bsp_data).__name__})")
    t = np.ascontiguousarray(t)
    tp = bsp_data.get("tp")
    tp_bytes = b"" if tp is None or np.array_equal(tp, t) else np.ascontiguousarray(tp).tobytes()
    gr = bsp_data.get("gr")
    return (int(bsp_data["nv"]), int(bsp_data.get("ks", bsp_data.get("ksq", 0))),
            int(bsp_data.get("ksp", bsp_data.get("ks", 0))),
            int(bsp_data.get("ksq", bsp_data.get("ks", 0))),
            int(bsp_data.get("nsp", bsp_data.get("ns", 0))),
            int(bsp_data.get("nsq", bsp_data.get("ns", 0))), int(bsp_data.get("ns", 0)),
            t.shape, t.tobytes(), tp_bytes, None if gr is None else np.shape(gr))


def shared_banks(bsp_data, key=None):
    """Process-wide bank dict of ``bsp_data``'s geometry, LRU over grids (convol.py:1455-1475)."""
    if key is None:
        key = grid_bank_key(bsp_data)
    banks = _RKB_BANKS.get(key)
    if banks is None:
        banks = {}
        _RKB_BANKS[key] = banks
        while len(_RKB_BANKS) > _BANK_CACHE_LIMIT:
            _RKB_BANKS.popitem(last=False)
    else:
        _RKB_BANKS.move_to_end(key)
    return banks


class ConvolCache:
    """Device-resident R^k banks and factors for one grid (convol.py:1478-1681).

    ``get_rkb`` / ``get_rkb_t`` / ``get_coulomb_factors`` return host arrays
    in the reference layouts; ``bank_device`` / ``factors_device`` return the
    resident device tensors used by the contractions.
    """

    def __init__(self, bsp_data):
        dv.require_gpu()
        load()
        self.bsp_data = bsp_data
        self.grid_key = grid_bank_key(bsp_data)
        self._cache = shared_banks(bsp_data, self.grid_key)
        bd = bsp_data
        self.ns = int(bd["ns"])
        self.nv = int(bd["nv"])
        self.ng = int(bd["gr"].shape[1])
        self.ksp = int(bd.get("ksp", bd["ks"]))
        self.ksq = int(bd.get("ksq", bd["ks"]))
        self.nsp = int(bd.get("nsp", bd["ns"]))
        self.nsq = int(bd.get("nsq", bd["ns"]))
        geo = self._cache.get("geometry")
        if geo is None:
            geo = {
                "gr": dv.f64(bd["gr"]), "grw": dv.f64(bd["grw"]),
                "bsp_p": dv.f64(bd["bsp_p"]), "bsp_q": dv.f64(bd.get("bsp_q", bd["bsp_p"])),
                "tp": dv.f64(bd.get("tp", bd["t"])), "t": dv.f64(bd["t"]),
            }
            self._cache["geometry"] = geo
        self._geo = geo

    def same_grid(self, bsp_data) -> bool:
        try:
            return grid_bank_key(bsp_data) == self.grid_key
        except (KeyError, TypeError):
            return False

    def order(self, c):
        return self.ksp if c == "p" else self.ksq

    def _comp(self, c):
        g = self._geo
        return (g["bsp_p"], self.ksp, g["tp"]) if c == "p" else (g["bsp_q"], self.ksq, g["t"])

    # -- device builders ----------------------------------------------------
    def _moments(self, c, k):
        key = ("mom", c, k)
        m = self._cache.get(key)
        if m is None:
            bsp, ks, _ = self._comp(c)
            m = dv.empty((ks * (ks + 1) // 2, self.nv))
            call("ax_cell_moments", dv.ptr(bsp), dv.ptr(bsp), dv.ptr(self._geo["gr"]),
                 dv.ptr(self._geo["grw"]), self.nv, self.ng, ks, ks, int(k), 1, dv.ptr(m), _s())
            self._cache[key] = m
        return m

    def _diag(self, left, right, k):
        key = ("diag", left, right, k)
        d = self._cache.get(key)
        if d is None:
            bl, ksl, tl = self._comp(left)
            br, ksr, tr = self._comp(right)
            n_sub = max(ksl, ksr) + (k // 2 if k >= 0 else 0)
            x = np.empty(n_sub)
            w = np.empty(n_sub)
            dp = ctypes.POINTER(ctypes.c_double)
            call("ax_gauss_legendre_01", n_sub, x.ctypes.data_as(dp), w.ctypes.data_as(dp))
            xd, wd = dv.f64(x), dv.f64(w)
            d = dv.empty((ksl * (ksl + 1) // 2, ksr * (ksr + 1) // 2, self.nv))
            call("ax_cell_triang", dv.ptr(bl), dv.ptr(br), dv.ptr(self._geo["gr"]),
                 dv.ptr(self._geo["grw"]), dv.ptr(tl), dv.ptr(tr), self.nv, self.ng, ksl, ksr,
                 int(k), n_sub, dv.ptr(xd), dv.ptr(wd), dv.ptr(d), _s())
            self._cache[key] = d
        return d

    def _pair(self, itype, k, mode):
        left, right = _COMPONENTS[itype]
        ksl, ksr = self.order(left), self.order(right)
        lo_l, hi_l = self._moments(left, k), self._moments(left, -(k + 1))
        lo_r, hi_r = self._moments(right, k), self._moments(right, -(k + 1))
        diag = self._diag(left, right, k)
        W = self.ns if mode == 0 else ksl + ksr - 1
        out = dv.empty((ksl, ksr, self.ns, W))
        call("ax_cell_pair_integrals", dv.ptr(lo_l), dv.ptr(hi_r), dv.ptr(lo_r), dv.ptr(hi_l),
             dv.ptr(diag), self.nv, ksl, ksr, self.ns, mode, dv.ptr(out), _s())
        return out

    def bank_device(self, k, itype="pppp"):
        """(bankT (ksl, ksr, ns, ns) device tensor, ns, ksl, ksr)."""
        if itype not in _COMPONENTS:
            raise ValueError(f"Unknown integral type: {itype}")
        key = ("bank", int(k), itype)
        b = self._cache.get(key)
        if b is None:
            b = self._pair(itype, int(k), 0)
            self._cache[key] = b
        left, right = _COMPONENTS[itype]
        return b, self.ns, self.order(left), self.order(right)

    def _global(self, c, k):
        key = ("glob", c, k)
        g = self._cache.get(key)
        if g is None:
            ks = self.order(c)
            g = dv.empty((ks, self.ns))
            mom = self._moments(c, k)
            call("ax_cell_global_moments", dv.ptr(mom), self.nv, ks, self.ns, dv.ptr(g), _s())
            self._cache[key] = g
        return g

    def factors_device(self, k, itype="pppp"):
        """(near, rising_left, falling_left, rising_right, falling_right) device tensors."""
        if itype not in _COMPONENTS:
            raise ValueError(f"Unknown integral type: {itype}")
        key = ("semi", int(k), itype)
        f = self._cache.get(key)
        if f is None:
            left, right = _COMPONENTS[itype]
            near = self._pair(itype, int(k), 1)
            f = (near, self._global(left, k), self._global(left, -(k + 1)),
                 self._global(right, k), self._global(right, -(k + 1)))
            self._cache[key] = f
        return f

    # -- reference-layout host views ------------------------------------------
    def get_rkb(self, k, itype="pppp"):
        """(rkb (ns, ns, ks_left, ks_right) host array, ns, ks_left) (convol.py:1507-1570).
        The array is read-only and remembers its resident device bank."""
        b, ns, ksl, _ = self.bank_device(k, itype)
        return _resident(dv.host(b).transpose(2, 3, 0, 1), bankT=b), ns, ksl

    def get_rkb_t(self, k, itype="pppp"):
        """Transposed bank rkbT[ip, jp, i, j] (convol.py:1572-1591), read-only,
        remembering its resident device bank."""
        b, ns, ksl, _ = self.bank_device(k, itype)
        return _resident(dv.host(b), device=b, bankT=b), ns, ksl

    def get_coulomb_factors(self, k, itype="pppp"):
        """Host copies of (near, rising_l, falling_l, rising_r, falling_r)
        (convol.py:1613-1658), read-only, remembering their device tensors."""
        return tuple(_resident(dv.host(x), device=x) for x in self.factors_device(k, itype))

    def get_symmetric_exchange_factors(self, k, itype="pppp"):
        if itype not in ("pppp", "qqqq"):
            raise ValueError("symmetric exchange factors need equal components")
        near, rising, falling, _, _ = self.get_coulomb_factors(k, itype)
        return near, rising, falling

    def release_shared_banks(self):
        geo = self._cache.get("geometry")
        self._cache.clear()
        if geo is not None:
            self._cache["geometry"] = geo

    # -- device contractions --------------------------------------------------
    def case1(self, k, itype, d, force=None):
        """Direct contraction of nb banded right densities d (nb, ns, ks_right)
        -> (nb, ns, ks_left); semiseparable when ns >= SEMISEPARABLE_DIRECT_NS
        unless ``force`` names the path ("bank" / "semiseparable")."""
        left, right = _COMPONENTS[itype]
        ksl, ksr = self.order(left), self.order(right)
        nb = d.shape[0]
        out = dv.empty((nb, self.ns, ksl))
        semi = (self.ns >= SEMISEPARABLE_DIRECT_NS) if force is None else force == "semiseparable"
        if semi:
            near, rl, fl, rr, fr = self.factors_device(k, itype)
            ws = dv.empty((max(1, load().ax_convol_semi_workspace(nb, self.ns, ksr)),))
            call("ax_convol_case1_semi", dv.ptr(near), dv.ptr(rl), dv.ptr(fl), dv.ptr(rr),
                 dv.ptr(fr), self.ns, ksl, ksr, dv.ptr(d), self.ns * ksr, nb, dv.ptr(ws),
                 dv.ptr(out), self.ns * ksl, _s())
        else:
            bank, ns, _, _ = self.bank_device(k, itype)
            call("ax_convol_case1", dv.ptr(bank), ns, ksl, ksr, dv.ptr(d), ns * ksr, nb,
                 dv.ptr(out), ns * ksl, _s())
        return out

    def case2(self, k, itype, d, n_left):
        """Left-pair contraction of left densities d (nb, n_left, ks_left)."""
        left, right = _COMPONENTS[itype]
        ksl, ksr = self.order(left), self.order(right)
        bank, ns, _, _ = self.bank_device(k, itype)
        nb = d.shape[0]
        out = dv.empty((nb, ns, ksr))
        call("ax_convol_case2", dv.ptr(bank), ns, ksl, ksr, int(n_left), dv.ptr(d),
             int(n_left) * ksl, nb, dv.ptr(out), ns * ksr, _s())
        return out

    def case4(self, k, itype, xx, symmetric, force=None):
        """Exchange contraction of full densities xx (nb, ns, ns)."""
        left, right = _COMPONENTS[itype]
        ksl, ksr = self.order(left), self.order(right)
        nb = xx.shape[0]
        out = dv.empty((nb, self.ns, self.ns))
        thr = SEMISEPARABLE_EXCHANGE_NS if left == right else SEMISEPARABLE_MIXED_EXCHANGE_NS
        semi = (self.ns >= thr) if force is None else force == "semiseparable"
        if semi:
            near, rl, fl, rr, fr = self.factors_device(k, itype)
            call("ax_convol_case4_semi", dv.ptr(near), dv.ptr(rl), dv.ptr(fl), dv.ptr(rr),
                 dv.ptr(fr), self.ns, ksl, ksr, dv.ptr(xx), self.ns * self.ns, nb,
                 1 if symmetric else 0, dv.ptr(out), self.ns * self.ns, _s())
        else:
            bank, ns, _, _ = self.bank_device(k, itype)
            call("ax_convol_case4", dv.ptr(bank), ns, ksl, ksr, dv.ptr(xx), ns * ns, nb,
                 1 if symmetric else 0, dv.ptr(out), ns * ns, _s())
        return out


# ---------------------------------------------------------------------------
# densities and small helpers on the device
# ---------------------------------------------------------------------------

def _pad_rows(vectors, n):
    out = np.zeros((len(vectors), n))
    for r, v in enumerate(vectors):
        v = np.asarray(v, dtype=np.float64).ravel()
        m = min(v.size, n)
        out[r, :m] = v[:m]
    return out


def density_sym_device(A, pairs, ns, ks, n1=None, n2=None):
    """Banded densities of coefficient rows A (device (n_rows, >= ns)) for the
    row pairs ``pairs`` (list of (a, b)) -> (np, ns, ks)."""
    pa = dv.i32(np.array([p[0] for p in pairs], dtype=np.int32))
    pb = dv.i32(np.array([p[1] for p in pairs], dtype=np.int32))
    out = dv.empty((len(pairs), ns, ks))
    n1 = ns if n1 is None else n1
    n2 = ns if n2 is None else n2
    call("ax_density_sym", dv.ptr(A), dv.ld(A), dv.ptr(pa), dv.ptr(pb), len(pairs), int(n1),
         int(n2), ns, ks, dv.ptr(out), _s())
    return out


def axpy(y, x, alpha=1.0):
    """y += alpha x on the device (contiguous tensors of equal size)."""
    call("ax_axpy", dv.ptr(y), dv.ptr(x), y.numel(), float(alpha), _s())
    return y


def sum_amb_device(a, b, ns, ks):
    """sum_{i, ip < ks, i+ip < ns} a[., i, ip] b[., i, ip] for batched banded operands."""
    nb = a.shape[0]
    out = dv.empty((nb,))
    call("ax_sum_amb_banded", dv.ptr(a), a[0].numel(), a.shape[2], dv.ptr(b), b[0].numel(),
         b.shape[2], nb, ns, ks, dv.ptr(out), _s())
    return out


# ---------------------------------------------------------------------------
# scalar and batched cell R^k
# ---------------------------------------------------------------------------

def _cell_rk_many(cache, k, quartets, force=None):
    """R^k(ab, cd) for many (Pa, Qa, Pb, Qb, Pc, Qc, Pd, Qd) tuples on one rank,
    the four-block cell formula of convol.py:1788-1882 (semiseparable for
    ns >= 96 unless ``force`` = "classic" / "semiseparable")."""
    ns, nsp, nsq, ksp, ksq = cache.ns, cache.nsp, cache.nsq, cache.ksp, cache.ksq
    nq = len(quartets)
    P = dv.f64(_pad_rows([v for qt in quartets for v in (qt[0], qt[2], qt[4], qt[6])], ns))
    Q = dv.f64(_pad_rows([v for qt in quartets for v in (qt[1], qt[3], qt[5], qt[7])], ns))
    ab = [(4 * t, 4 * t + 1) for t in range(nq)]
    cd = [(4 * t + 2, 4 * t + 3) for t in range(nq)]
    semi = ns >= SEMISEPARABLE_DIRECT_NS if force is None else force == "semiseparable"
    total = None
    if semi:
        lp, lq = density_sym_device(P, ab, ns, ksp), density_sym_device(Q, ab, ns, ksq)
        rp, rq = density_sym_device(P, cd, ns, ksp), density_sym_device(Q, cd, ns, ksq)
        for itype, left, right, ksl in (("pppp", lp, rp, ksp), ("qqqq", lq, rq, ksq),
                                        ("pqpq", lp, rq, ksp), ("qpqp", lq, rp, ksq)):
            con = cache.case1(k, itype, right, force="semiseparable")
            v = sum_amb_device(con, left, ns, ksl)
            total = v if total is None else axpy(total, v)
        return total
    # classic: bank case 2 then trace (convol.py:1815-1880), per-block extents
    for itype, lcoef, lks, ln, rcoef, rks, rn, ext in (
            ("pppp", P, ksp, nsp, P, ksp, nsp, nsp), ("qqqq", Q, ksq, nsq, Q, ksq, nsq, nsq),
            ("pqpq", P, ksp, nsp, Q, ksq, nsq, min(ns, nsq)),
            ("qpqp", Q, ksq, nsq, P, ksp, nsp, min(ns, nsp))):
        dl = density_sym_device(lcoef, ab, ln, lks)
        dr = density_sym_device(rcoef, cd, rn, rks)
        con = cache.case2(k, itype, dl, ln)  # (nq, ns, rks)
        v = sum_amb_device(con, dr, ext, rks)
        total = v if total is None else axpy(total, v)
    return total


def _finite_or_raise(values, k, what="compute_slater_rk_cell"):
    if not np.all(np.isfinite(values)):
        raise FloatingPointError(f"{what}: non-finite R^k (k={k})")
    return values


def compute_slater_rk_cell(Pa, Qa, Pb, Qb, Pc, Qc, Pd, Qd, bsp_data, k, cache=None):
    """Scalar R^k(ab, cd) by the cell algorithm, all four component blocks
    (convol.py:1756-1882); FloatingPointError after three non-finite results."""
    if cache is None:
        cache = ConvolCache(bsp_data)
    last = None
    for _ in range(3):
        v = float(dv.host(_cell_rk_many(cache, int(k), [(Pa, Qa, Pb, Qb, Pc, Qc, Pd, Qd)]))[0])
        if np.isfinite(v):
            return v
        last = v
    raise FloatingPointError(
        f"compute_slater_rk_cell: non-finite R^k on 3 consecutive evaluations (k={k}, value={last})")


def compute_slater_fk_cell(Pa, Qa, Pb, Qb, bsp_data, k, cache=None):
    """F^k(a,b) = R^k(aa, bb) (convol.py:1885-1888)."""
    return compute_slater_rk_cell(Pa, Qa, Pa, Qa, Pb, Qb, Pb, Qb, bsp_data, k, cache)


def compute_slater_gk_cell(Pa, Qa, Pb, Qb, bsp_data, k, cache=None):
    """G^k(a,b) = R^k(ab, ba) (convol.py:1891-1894)."""
    return compute_slater_rk_cell(Pa, Qa, Pb, Qb, Pb, Qb, Pa, Qa, bsp_data, k, cache)


def compute_slater_rk_cell_batch(orbital_arrays, keys, bsp_data, cache=None):
    """Batched R^k(ab, cd) over many (a, b, c, d, k) keys (convol.py:2084-2144):
    grouped by k, the full-bank contraction for every key (as the reference's
    batch kernels, independent of the ns threshold)."""
    if cache is None:
        cache = ConvolCache(bsp_data)
    by_k = {}
    for idx, key in enumerate(keys):
        by_k.setdefault(int(key[4]), []).append(idx)
    result = [0.0] * len(keys)
    for k, idxs in by_k.items():
        quartets = []
        for i in idxs:
            a, b, c, d, _ = keys[i]
            quartets.append((*orbital_arrays[a], *orbital_arrays[b], *orbital_arrays[c],
                             *orbital_arrays[d]))
        vals = dv.host(_cell_rk_many(cache, k, quartets, force="classic"))
        for r, i in enumerate(idxs):
            result[i] = float(vals[r])
    return result


def compute_direct_potential_cell(orb_P, orb_Q, bsp_data, k, cache=None):
    """Direct matrices (PP, QQ) from the complete density P^2 + Q^2 (convol.py:1694-1717)."""
    if cache is None:
        cache = ConvolCache(bsp_data)
    ns = cache.ns
    C = {"p": dv.f64(_pad_rows([orb_P], ns)), "q": dv.f64(_pad_rows([orb_Q], ns))}
    dens = {c: density_sym_device(C[c], [(0, 0)], ns, cache.order(c)) for c in "pq"}
    blocks = []
    for a in "pq":
        banded = None
        for b in "pq":
            v = cache.case1(int(k), a + b + a + b, dens[b], force="bank")
            banded = v if banded is None else axpy(banded, v)
        full = dv.zeros((1, ns, ns))
        banded = banded.contiguous()
        call("ax_banded_to_full_add", dv.ptr(full), ns * ns, dv.ptr(banded),
             ns * cache.order(a), 1, ns, cache.order(a), 1.0, _s())
        blocks.append(dv.host(full)[0])
    return tuple(blocks)


def compute_exchange_matrix_cell(orb_a_P, orb_a_Q, orb_b_P, orb_b_Q, bsp_data, k, cache=None):
    """Exchange blocks (PP, QQ, PQ, QP) (convol.py:1722-1751)."""
    if cache is None:
        cache = ConvolCache(bsp_data)
    ns = cache.ns
    left = {"p": orb_a_P, "q": orb_a_Q}
    right = {"p": orb_b_P, "q": orb_b_Q}
    diagonal = orb_a_P is orb_b_P and orb_a_Q is orb_b_Q
    blocks = {}
    for a, b in (("p", "p"), ("q", "q"), ("p", "q"), ("q", "p")):
        if diagonal and (a, b) == ("q", "p"):
            blocks[a, b] = blocks["p", "q"].T.copy()
            continue
        la = _pad_rows([left[a]], ns)[0]
        rb = _pad_rows([right[b]], ns)[0]
        xx = dv.f64(np.outer(la, rb)[None, :, :])
        blocks[a, b] = dv.host(cache.case4(int(k), a + b + a + b, xx,
                                           symmetric=diagonal and a == b, force="bank"))[0]
    return tuple(blocks[x] for x in (("p", "p"), ("q", "q"), ("p", "q"), ("q", "p")))


def build_slater_rk_bilinear(Pa, Qa, Pb, Qb, bsp_data, k, cache=None):
    """Matrices M with R^k(a r, b s) = r^T M_PP s + ... for fixed (a, b)
    (convol.py:1900-1978): stage 1 is the case-2 contraction of the banded
    densities sym(wa, e_m) for every m (one batched launch), stage 2 the
    gather over the right band pair (ax_bilinear_stage2)."""
    if cache is None:
        cache = ConvolCache(bsp_data)
    ns = cache.ns
    mats = []
    for itype, wa, wb, nl, nr in (("pppp", Pa, Pb, cache.nsp, cache.nsp),
                                  ("qqqq", Qa, Qb, cache.nsq, cache.nsq),
                                  ("pqpq", Pa, Qb, min(ns, cache.nsp), min(ns, cache.nsq)),
                                  ("qpqp", Qa, Pb, min(ns, cache.nsq), min(ns, cache.nsp))):
        left, right = _COMPONENTS[itype]
        ksl, ksr = cache.order(left), cache.order(right)
        rows = np.zeros((ns + 1, ns))
        rows[:ns, :] = np.eye(ns)
        rows[ns] = _pad_rows([wa], ns)[0]
        A = dv.f64(rows)
        dl = density_sym_device(A, [(ns, m) for m in range(ns)], nl, ksl)
        T = cache.case2(int(k), itype, dl, nl)  # (ns, ns, ksr)
        wbd = dv.f64(_pad_rows([wb], ns)[0])
        M = dv.empty((ns, ns))
        call("ax_bilinear_stage2", dv.ptr(T), ns, ksr, int(nr), dv.ptr(wbd), ns, dv.ptr(M), _s())
        mats.append(dv.host(M))
    return tuple(mats)


def eval_slater_rk_bilinear(M, Pl, Ql, Pr, Qr):
    """Single R^k(a l, b r) from the bilinear matrices (convol.py:1981-1985)."""
    M_PP, M_QQ, M_PQ, M_QP = M
    return float(Pl @ (M_PP @ Pr) + Ql @ (M_QQ @ Qr) + Pl @ (M_PQ @ Qr) + Ql @ (M_QP @ Pr))


def compute_slater_rk_table(M, CP_l, CQ_l, CP_r, CQ_r):
    """Table R^k(a l, b r) over coefficient columns (convol.py:1988-1996)."""
    M_PP, M_QQ, M_PQ, M_QP = M
    return (CP_l.T @ M_PP @ CP_r + CQ_l.T @ M_QQ @ CQ_r
            + CP_l.T @ M_PQ @ CQ_r + CQ_l.T @ M_QP @ CP_r)


# ---------------------------------------------------------------------------
# private kernels of the reference module, by name and argument order
# (convol.py:609-1373): numpy in, numpy out (or in place, as the reference),
# every step on the device.  Banks and factors handed out by
# ConvolCache.get_rkb / get_rkb_t / get_coulomb_factors carry their resident
# device tensor, so passing them back costs no upload.
# ---------------------------------------------------------------------------

class _Resident(np.ndarray):
    """Read-only host copy of a device-resident bank or factor array that
    remembers the device tensor it came from (views and copies do not)."""

    def __array_finalize__(self, obj):
        self._ax_device = None
        self._ax_bankT = None


def _resident(host, device=None, bankT=None):
    h = np.asarray(host).view(_Resident)
    h._ax_device = device
    h._ax_bankT = bankT
    h.flags.writeable = False
    return h


def _dev(a):
    """Device FP64 copy of a host operand, or the tensor a resident array came from."""
    d = getattr(a, "_ax_device", None)
    if d is not None:
        return d
    return dv.f64(np.ascontiguousarray(a, dtype=np.float64))


def _bankT_dev(bank, transposed):
    """Device bankT (ksl, ksr, ns, ns) of a bank given as rkbT[ip, jp, i, j]
    (transposed) or rkb[i, j, ip, jp]."""
    d = getattr(bank, "_ax_bankT", None)
    if d is not None:
        return d
    if dv._is_tensor(bank):  # already on the device (a wrapper passing its upload on)
        return dv.f64(bank if transposed else bank.permute(2, 3, 0, 1))
    b = np.asarray(bank, dtype=np.float64)
    if not transposed:
        b = b.transpose(2, 3, 0, 1)
    return dv.f64(np.ascontiguousarray(b))


def _rows(vectors, width):
    """Coefficient vectors as zero-padded rows of one device matrix."""
    out = np.zeros((len(vectors), max(width, 1)))
    for r, v in enumerate(vectors):
        v = np.asarray(v, dtype=np.float64).ravel()
        m = min(v.size, out.shape[1])
        out[r, :m] = v[:m]
    return dv.f64(out)


def _check_bank(b, ksl, ksr, ns, name):
    if tuple(b.shape) != (ksl, ksr, ns, ns):
        raise ValueError(f"{name}: bank of shape {tuple(b.shape)}, expected "
                         f"({ksl}, {ksr}, {ns}, {ns}) in the transposed layout")


def _into(dst, dev_result):
    """Writes a device result into the caller's array in place."""
    dst[...] = dv.host(dev_result).reshape(dst.shape)


def _density_sym(p1, p2, ns, ks):
    """convol.py:712-753: banded symmetric density (ns, ks), reads past
    either vector's length are zero."""
    p1, p2 = np.asarray(p1).ravel(), np.asarray(p2).ravel()
    A = _rows([p1, p2], max(p1.size, p2.size, ns))
    return dv.host(density_sym_device(A, [(0, 1)], ns, ks, p1.size, p2.size))[0]


def _density_full(p1, p2, ns):
    """convol.py:756-771: d[i, j] = p1[i] p2[j] (ns, ns), zero past the lengths."""
    A, B = _rows([p1], ns), _rows([p2], ns)
    out = dv.empty((1, ns, ns))
    one = dv.f64(np.ones((1, 1)))
    call("ax_exchange_density", dv.ptr(A), dv.ptr(B), dv.ld(A), 1, ns, dv.ptr(one), 1,
         dv.ptr(out), _s())
    return dv.host(out)[0]


def _accumulate_density_sym(dd, p1, p2, ns, ks, coeff):
    """convol.py:609-630: dd += coeff * _density_sym(p1, p2) (in place)."""
    d = dv.f64(np.ascontiguousarray(dd, dtype=np.float64))
    axpy(d, dv.f64(_density_sym(p1, p2, ns, ks)).reshape(d.shape), float(coeff))
    _into(dd, d)


def _accumulate_density_full(xx, p1, p2, ns, coeff):
    """convol.py:633-643: xx += coeff * outer(p1, p2) (in place)."""
    x = dv.f64(np.ascontiguousarray(xx, dtype=np.float64))
    axpy(x, dv.f64(_density_full(p1, p2, ns)).reshape(x.shape), float(coeff))
    _into(xx, x)


def _accumulate_direct_density_batch(dd, comp_stack, coeffs, target_i, ns, ks):
    """convol.py:646-685: dd += sum_j c_j _density_sym(p_j, p_j), the
    target's own coefficient doubled (in place)."""
    comp = np.asarray(comp_stack, dtype=np.float64)
    c = np.array(coeffs, dtype=np.float64).ravel()
    if 0 <= int(target_i) < c.size:
        c[int(target_i)] = c[int(target_i)] + c[int(target_i)]
    n_src = comp.shape[0]
    A = _rows(list(comp), max(comp.shape[1], ns))
    dens = density_sym_device(A, [(j, j) for j in range(n_src)], ns, ks,
                              min(comp.shape[1], ns), min(comp.shape[1], ns))
    acc = dv.empty((1, ns, ks))
    call("ax_combine_rows", dv.ptr(dens), ns * ks, n_src, dv.ptr(dv.f64(c[None, :])), 1,
         dv.ptr(acc), _s())
    d = dv.f64(np.ascontiguousarray(dd, dtype=np.float64))
    axpy(d, acc.reshape(d.shape))
    _into(dd, d)


def _accumulate_exchange_density_batch(xx, comp_ii_stack, comp_jj_stack, coeffs, ns):
    """convol.py:688-709: xx += sum_j (b_j A_j) B_j^T (in place)."""
    A = _rows(list(np.asarray(comp_ii_stack)), ns)
    B = _rows(list(np.asarray(comp_jj_stack)), ns)
    c = dv.f64(np.asarray(coeffs, dtype=np.float64).ravel()[None, :])
    out = dv.empty((1, ns, ns))
    call("ax_exchange_density", dv.ptr(A), dv.ptr(B), dv.ld(A), A.shape[0], ns, dv.ptr(c), 1,
         dv.ptr(out), _s())
    x = dv.f64(np.ascontiguousarray(xx, dtype=np.float64))
    axpy(x, out.reshape(x.shape))
    _into(xx, x)


def _banded_to_full_add(V, v_banded, ns, ks, coeff):
    """convol.py:814-825: V[i, i+s] and V[i+s, i] += coeff v_banded[i, s] (in place)."""
    if np.shape(V) != (ns, ns):
        raise ValueError(f"_banded_to_full_add: V must be ({ns}, {ns})")
    Vd = dv.f64(np.ascontiguousarray(V, dtype=np.float64))
    vb = dv.f64(np.ascontiguousarray(v_banded, dtype=np.float64)[:ns, :ks])
    call("ax_banded_to_full_add", dv.ptr(Vd), ns * ns, dv.ptr(vb), ns * ks, 1, ns, ks,
         float(coeff), _s())
    _into(V, Vd)


def _convol_case1_t(rkbT, d_right, ns, ks_left, ks_right):
    """convol.py:971-997: a[i, ip] = sum rkbT[ip, jp, i, j] d[j, jp] -> (ns, ks_left)."""
    b = _bankT_dev(rkbT, True)
    _check_bank(b, ks_left, ks_right, ns, "_convol_case1_t")
    d = dv.f64(np.ascontiguousarray(d_right, dtype=np.float64))
    out = dv.empty((1, ns, ks_left))
    call("ax_convol_case1", dv.ptr(b), ns, ks_left, ks_right, dv.ptr(d), ns * ks_right, 1,
         dv.ptr(out), ns * ks_left, _s())
    return dv.host(out)[0]


def _convol_case1(rkb, d, ns, ks):
    """convol.py:788-811 (bank rkb[i, j, ip, jp])."""
    return _convol_case1_t(_bankT_dev(rkb, False), d, ns, ks, ks)


def _convol_case1_mixed(rkb, d_right, ns, ks_left, ks_right):
    """convol.py:828-849 (bank rkb[i, j, ip, jp])."""
    return _convol_case1_t(_bankT_dev(rkb, False), d_right, ns, ks_left, ks_right)


def _convol_case2_mixed(rkb, d_left, ns, ks_left, ks_right):
    """convol.py:1310-1345: a[j, jp] = sum rkb[i, j, ip, jp] d_left[i, ip] over
    i + ip < n_left (the density's own row count) -> (ns, ks_right)."""
    b = _bankT_dev(rkb, False)
    _check_bank(b, ks_left, ks_right, ns, "_convol_case2_mixed")
    dl = np.ascontiguousarray(d_left, dtype=np.float64)
    n_left = dl.shape[0]
    d = dv.f64(dl)
    out = dv.empty((1, ns, ks_right))
    call("ax_convol_case2", dv.ptr(b), ns, ks_left, ks_right, n_left, dv.ptr(d),
         n_left * ks_left, 1, dv.ptr(out), ns * ks_right, _s())
    return dv.host(out)[0]


def _convol_case2(rkb, d, ns, ks):
    """convol.py:852-874."""
    return _convol_case2_mixed(rkb, np.asarray(d)[:ns], ns, ks, ks)


def _convol_case4_t(rkbT, d, ns, ks, ks_right=None, symmetric_output=False):
    """convol.py:1000-1058: exchange contraction of a full density (ns, ns)."""
    ksr = ks if ks_right is None else ks_right
    b = _bankT_dev(rkbT, True)
    _check_bank(b, ks, ksr, ns, "_convol_case4_t")
    x = dv.f64(np.ascontiguousarray(np.asarray(d, dtype=np.float64)[:ns, :ns]))
    out = dv.empty((1, ns, ns))
    call("ax_convol_case4", dv.ptr(b), ns, ks, ksr, dv.ptr(x), ns * ns, 1,
         1 if symmetric_output else 0, dv.ptr(out), ns * ns, _s())
    return dv.host(out)[0]


def _convol_case4(rkb, d, ns, ks, ks_right=None):
    """convol.py:877-948 (bank rkb[i, j, ip, jp])."""
    ksr = ks if ks_right is None else ks_right
    return _convol_case4_t(_bankT_dev(rkb, False), d, ns, ks, ksr, False)


def _convol_case1_semiseparable_t(near, rising_left, falling_left, rising_right,
                                  falling_right, density, ns, ks_left, ks_right):
    """convol.py:1125-1162: the direct contraction from pair moments and the
    overlap strip -> (ns, ks_left)."""
    f = [_dev(x) for x in (near, rising_left, falling_left, rising_right, falling_right)]
    d = dv.f64(np.ascontiguousarray(density, dtype=np.float64))
    ws = dv.empty((max(1, int(load().ax_convol_semi_workspace(1, ns, ks_right))),))
    out = dv.empty((1, ns, ks_left))
    call("ax_convol_case1_semi", *[dv.ptr(x) for x in f], ns, ks_left, ks_right, dv.ptr(d),
         ns * ks_right, 1, dv.ptr(ws), dv.ptr(out), ns * ks_left, _s())
    return dv.host(out)[0]


def _convol_case4_semi_common(factors, density, ns, ksl, ksr, symmetric):
    f = [_dev(x) for x in factors]
    x = dv.f64(np.ascontiguousarray(np.asarray(density, dtype=np.float64)[:ns, :ns]))
    out = dv.empty((1, ns, ns))
    call("ax_convol_case4_semi", *[dv.ptr(t) for t in f], ns, ksl, ksr, dv.ptr(x), ns * ns, 1,
         1 if symmetric else 0, dv.ptr(out), ns * ns, _s())
    return dv.host(out)[0]


def _convol_case4_semiseparable_t(near, rising, falling, density, ns, ks):
    """convol.py:1187-1222: symmetric exchange from pair moments and local cells."""
    return _convol_case4_semi_common((near, rising, falling, rising, falling), density, ns, ks,
                                     ks, True)


def _convol_case4_semiseparable_mixed_t(near, rising_left, falling_left, rising_right,
                                        falling_right, density, ns, ks_left, ks_right):
    """convol.py:1250-1303: mixed-component exchange from pair moments."""
    return _convol_case4_semi_common((near, rising_left, falling_left, rising_right,
                                      falling_right), density, ns, ks_left, ks_right, False)


def _sum_amb_banded(a, b, ns, ks):
    """convol.py:1348-1356: sum_{i, s; i+s < ns} a[i, s] b[i, s]."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    ad, bd = dv.f64(a[None]), dv.f64(b[None])
    out = dv.empty((1,))
    call("ax_sum_amb_banded", dv.ptr(ad), a.size, a.shape[1], dv.ptr(bd), b.size, b.shape[1], 1,
         ns, ks, dv.ptr(out), _s())
    return float(dv.host(out)[0])


def _sum_amb_full(a, b, ns):
    """convol.py:1359-1373: sum_{i, j < ns} a[i, j] b[i, j]."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64)[:ns, :ns]).reshape(-1, 1)
    b = np.ascontiguousarray(np.asarray(b, dtype=np.float64)[:ns, :ns]).reshape(-1, 1)
    ad, bd = dv.f64(a), dv.f64(b)
    idx = dv.i32(np.zeros(1, dtype=np.int32))
    out = dv.empty((1,))
    call("ax_feat_dots", dv.ptr(ad), 1, dv.ptr(bd), 1, a.shape[0], 1, dv.ptr(idx), dv.ptr(idx), 0,
         dv.ptr(out), _s())
    return float(dv.host(out)[0])
