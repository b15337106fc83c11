"""Second-order many-body perturbation theory on a closed-shell Dirac-Fock
reference: drop-in for methods/mbpt.py (``MBPT2Result``, ``_allowed_k_range``,
``reduced_C_jj``, ``build_virtual_basis``, ``mbpt2_energy``).

Device path of ``mbpt2_energy``:
  1. the closed-shell Fock operator of every kappa (V^N by default): the cell
     direct potential (k = 0) and exchange matrices of each occupied orbital,
     computed once and combined per kappa with its coefficients, then the
     FP64 host eigensolve;
  2. for every occupied pair (a, b) and needed rank k the bilinear form
     M^k_ab (2ns x 2ns, R^k(a r, b s) = c_r^T M c_s with c = [P; Q]): stage 1
     (the case-2 contraction of the band densities (a, e_m)) once per
     (rank, component block) for all occupied orbitals together, stage 2 for
     every (pair, rank, block) in one launch (``ax_bilinear_stage2_batch``),
     written side by side as [M_1 | M_2 | ...];
  3. the Slater tables T^k_ab[r][s] = R^k(a r, b s) of all virtual pairs:
     two plain FP64 GEMMs per group of pairs, C^T [M_1 | M_2 | ...] and then
     (N B x 2ns) times C, laid out [r][item][s];
  4. ``ax_mbpt2_pair_sum``: the J-coupled amplitudes and the energy
     denominators of every (pair, r <= s) with per-channel coefficients (6j,
     reduced C^k, phases) built on the host once per pair type, fixed-order
     sums per pair.
The reference's state object (SCFResult: subshells, orbitals_raw,
orbital_energies, _bsp_data, Z, converged) and this package's DiracFockSCF
(shells, P, Q, eps) are both accepted."""

from __future__ import annotations

import logging
import time
from dataclasses import dataclass, field
from functools import lru_cache
from typing import Any, Dict, List, Optional, Tuple

import numpy as np

from . import device as dv
from ._lib import call, load
from .angular import exchange_lambda, l_of_kappa, reduced_c_jj, twoj_of_kappa, wigner_6j
from .convol import (ConvolCache, compute_direct_potential_cell, compute_exchange_matrix_cell,
                     density_sym_device)
from .dirac import C_SQ, dirac_matrices, solve_dirac

__all__ = ["MBPT2Result", "build_virtual_basis", "mbpt2_energy", "reduced_C_jj",
           "_allowed_k_range"]

logger = logging.getLogger(__name__)

reduced_C_jj = reduced_c_jj  # Grant's phase convention (angular/coefficients.py:29-65)

_MAX_J = 8            # J values of one occupied pair type the pair-sum kernel holds
_TABLE_BYTES = 1 << 30  # Slater tables resident at once (pairs are grouped to fit)


@dataclass
class MBPT2Result:
    """methods/mbpt.py:96-127."""
    correlation_energy: float = 0.0
    pair_contributions: Dict[Tuple[str, str], float] = field(default_factory=dict)
    n_virtuals: Dict[int, int] = field(default_factory=dict)
    kappa_max: int = 0
    n_max: Optional[int] = None
    cpu_seconds: float = 0.0
    provenance: Dict[str, Any] = field(default_factory=dict)


@lru_cache(maxsize=None)
def _allowed_k_range(kappa_a: int, kappa_b: int) -> List[int]:
    """Ranks k with |j_a - j_b| <= k <= j_a + j_b and l_a + k + l_b even
    (mbpt.py:138-150; the cached list must not be mutated)."""
    ta, tb = twoj_of_kappa(kappa_a), twoj_of_kappa(kappa_b)
    la, lb = l_of_kappa(kappa_a), l_of_kappa(kappa_b)
    return [k for k in range(abs(ta - tb) // 2, (ta + tb) // 2 + 1) if (k + la + lb) % 2 == 0]


# ---------------------------------------------------------------------------
# the reference state
# ---------------------------------------------------------------------------
@dataclass
class _Shell:
    n: int
    kappa: int
    occupation: float
    label: str


class _Reference:
    """Closed-shell reference as this module reads it."""

    def __init__(self, scf):
        if hasattr(scf, "shells") and hasattr(scf, "P") and hasattr(scf, "eps"):
            self.Z, self.bd = scf.Z, scf.bd
            self.nuclear, self.A = scf.nuclear, scf.A
            self.shells = [_Shell(s.n, s.kappa, float(s.occupation), s.label) for s in scf.shells]
            self.P = [np.asarray(p, dtype=np.float64) for p in scf.P]
            self.Q = [np.asarray(q, dtype=np.float64) for q in scf.Q]
            self.energies = {s.label: float(e) for s, e in zip(scf.shells, scf.eps)}
            self.converged = bool(getattr(scf, "converged", True))
            return
        bd = getattr(scf, "_bsp_data", None)
        if bd is None:
            raise ValueError("SCFResult is missing _bsp_data — re-run the SCF solve.")
        if getattr(scf, "Z", None) is None:
            raise ValueError("SCFResult is missing Z.")
        self.Z, self.bd, self.A = scf.Z, bd, None
        nuc = getattr(scf, "_nuclear_potential", None)
        if nuc is None and hasattr(scf, "provenance"):
            nuc = scf.provenance.get("nuclear_potential_func")
        self.nuclear = nuc
        self.shells, self.P, self.Q = [], [], []
        raw = scf.orbitals_raw
        for sh in scf.subshells:
            same = [s for s in scf.subshells if s.kappa == sh.kappa]
            idx = same.index(sh)
            rows = raw.get(sh.kappa, [])
            if idx >= len(rows):
                continue
            P, Q, _ = rows[idx]
            self.shells.append(_Shell(sh.n, sh.kappa, float(sh.occupation), sh.label))
            self.P.append(np.asarray(P, dtype=np.float64))
            self.Q.append(np.asarray(Q, dtype=np.float64))
        self.energies = dict(scf.orbital_energies)
        self.converged = bool(scf.converged)


class _CoreParts:
    """Cell direct potentials (k = 0) and exchange matrices of the occupied
    orbitals, each computed once on the device for every kappa's operator."""

    def __init__(self, ref, cache):
        self.ref, self.cache = ref, cache
        self._direct, self._exchange = {}, {}

    def direct(self, i):
        if i not in self._direct:
            self._direct[i] = compute_direct_potential_cell(self.ref.P[i], self.ref.Q[i],
                                                            self.ref.bd, 0, self.cache)
        return self._direct[i]

    def exchange(self, i, k):
        if (i, k) not in self._exchange:
            P, Q = self.ref.P[i], self.ref.Q[i]
            self._exchange[i, k] = compute_exchange_matrix_cell(P, Q, P, Q, self.ref.bd, k,
                                                                self.cache)
        return self._exchange[i, k]


def _closed_shell_fock(ref, kappa, parts, potential_mode="VN"):
    """H0(kappa) + sum_b q_b V_d^0(b) - sum_b sum_k q_b Lambda^k /
    ((2j+1)(2j_b+1)) K^k(b) (mbpt.py:158-296); "VN-1" takes one electron
    from the outermost subshell (largest (n, l, -|kappa|))."""
    bd = ref.bd
    ns = bd["ns"]
    H, S = dirac_matrices(kappa, ref.Z, bd, ref.nuclear, ref.A)
    if potential_mode == "VN-1":
        outer = max(ref.shells, key=lambda s: (s.n, l_of_kappa(s.kappa), -abs(s.kappa)))
        reduce_key = (outer.n, outer.kappa)
    elif potential_mode == "VN":
        reduce_key = None
    else:
        raise ValueError(f"Unknown potential_mode {potential_mode!r}; expected 'VN-1' or 'VN'.")
    tj = twoj_of_kappa(kappa)
    for i, sh in enumerate(ref.shells):
        q = sh.occupation - (1.0 if (sh.n, sh.kappa) == reduce_key else 0.0)
        if q <= 0.0:
            continue
        vpp, vqq = parts.direct(i)
        H[:vpp.shape[0], :vpp.shape[0]] += q * vpp
        H[ns:ns + vqq.shape[0], ns:ns + vqq.shape[0]] += q * vqq
        for k in _allowed_k_range(kappa, sh.kappa):
            lam = exchange_lambda(kappa, sh.kappa, k)
            if abs(lam) < 1e-15:
                continue
            c = -q * lam / ((tj + 1) * (twoj_of_kappa(sh.kappa) + 1))
            kpp, kqq, kpq, kqp = parts.exchange(i, k)
            H[:kpp.shape[0], :kpp.shape[0]] += c * kpp
            H[ns:ns + kqq.shape[0], ns:ns + kqq.shape[0]] += c * kqq
            H[:ns, ns:] += c * kpq
            H[ns:, :ns] += c * kqp
    return H, S


def _virtual_basis(ref, kappa, parts, e_max=None, potential_mode="VN",
                   orthogonal_to_occupied=False):
    bd = ref.bd
    ns = bd["ns"]
    nsp, nsq = bd.get("nsp", ns), bd.get("nsq", ns)
    H, S = _closed_shell_fock(ref, kappa, parts, potential_mode)
    stored = [i for i, sh in enumerate(ref.shells) if sh.kappa == kappa and sh.occupation > 0.01]
    cons = (np.column_stack([np.r_[ref.P[i], ref.Q[i]] for i in stored])
            if orthogonal_to_occupied and stored else None)
    w, v = solve_dirac(H, S, ns, nsp, nsq, orthogonal_to=cons)
    keep = w > -1.5 * C_SQ
    if e_max is not None:
        keep &= w < e_max
    idx = np.flatnonzero(keep)
    idx = idx[np.argsort(w[idx], kind="stable")]
    if kappa > 0 and bd.get("ksp", bd["ks"]) == bd.get("ksq", bd["ks"]):
        idx = idx[1:]  # equal P/Q orders carry a spurious lowest positive-kappa root
    states = []
    for i in idx:
        P = np.zeros(ns)
        Q = np.zeros(ns)
        P[:nsp] = v[:nsp, i]
        Q[:nsq] = v[ns:ns + nsq, i]
        states.append((P, Q, float(w[i])))
    if orthogonal_to_occupied:
        return [(ref.P[i], ref.Q[i], ref.energies[ref.shells[i].label]) for i in stored], states
    used = set()
    for i, sh in enumerate(ref.shells):
        if sh.kappa != kappa or sh.occupation <= 0:
            continue
        a = np.r_[ref.P[i], ref.Q[i]]
        Sa = S @ a
        best, best_j = -1.0, -1
        for j, (P, Q, _) in enumerate(states):
            if j in used:
                continue
            ov = abs(float(np.r_[P, Q] @ Sa))
            if ov > best:
                best, best_j = ov, j
        if best_j >= 0 and best > 0.5:
            used.add(best_j)
    return ([states[j] for j in sorted(used)],
            [s for j, s in enumerate(states) if j not in used])


def build_virtual_basis(scf_result, kappa, *, cache=None, e_max=None, potential_mode="VN",
                        orthogonal_to_occupied=False):
    """(occupied, virtuals) eigenstates [(P, Q, energy)] of the closed-shell
    Fock operator of ``kappa`` (mbpt.py:303-408): physical roots above
    -1.5 c^2 (and below ``e_max``) in ascending order; occupied states are
    those overlapping a stored occupied orbital by more than 0.5 in the S
    metric, or, with ``orthogonal_to_occupied``, the stored orbitals
    themselves with the virtuals solved in their metric complement."""
    ref = _Reference(scf_result)
    cache = ConvolCache(ref.bd) if cache is None else cache
    return _virtual_basis(ref, kappa, _CoreParts(ref, cache), e_max, potential_mode,
                          orthogonal_to_occupied)


# ---------------------------------------------------------------------------
# J-coupled channel coefficients (host, once per pair type)
# ---------------------------------------------------------------------------
def _channel(ka, kb, kr, ks, multipole_max):
    """(first 2J, cd, ce, ranks) of virtual channel (kr, ks) for an occupied
    pair of kappas (ka, kb), or None where the reference adds nothing
    (mbpt.py:652-790): cd[J][k], ce[J][k] as dicts keyed by (j index, k)."""
    if not _allowed_k_range(ka, kr) and not _allowed_k_range(kb, kr):
        return None
    ta, tb, tr, ts = (twoj_of_kappa(x) for x in (ka, kb, kr, ks))
    lo, hi = max(abs(ta - tb), abs(tr - ts)), min(ta + tb, tr + ts)
    if lo > hi:
        return None
    if (l_of_kappa(ka) + l_of_kappa(kb) + l_of_kappa(kr) + l_of_kappa(ks)) % 2:
        return None
    kd = [k for k in _allowed_k_range(ka, kr) if k in _allowed_k_range(kb, ks)]
    ke = [k for k in _allowed_k_range(ka, ks) if k in _allowed_k_range(kb, kr)]
    if multipole_max is not None:
        kd = [k for k in kd if k <= multipole_max]
        ke = [k for k in ke if k <= multipole_max]
    if not kd and not ke:
        return None
    rd = [k for k in kd if abs(reduced_C_jj(ka, kr, k)) >= 1e-15
          and abs(reduced_C_jj(kb, ks, k)) >= 1e-15]
    re = [k for k in ke if abs(reduced_C_jj(ka, ks, k)) >= 1e-15
          and abs(reduced_C_jj(kb, kr, k)) >= 1e-15]
    if not rd and not re:
        return None
    cd, ce = {}, {}
    for j, tJ in enumerate(range(lo, hi + 1, 2)):
        ph_d = -1.0 if ((tb + tr + tJ) // 2) % 2 else 1.0
        ph_e = -1.0 if ((tb + ts + tJ) // 2) % 2 else 1.0
        swap = -1.0 if ((tr + ts - tJ) // 2) % 2 else 1.0
        for k in rd:
            sj = wigner_6j(ta, tb, tJ, ts, tr, 2 * k)
            if abs(sj) >= 1e-15:
                cd[j, k] = ph_d * sj * (reduced_C_jj(ka, kr, k) * reduced_C_jj(kb, ks, k))
        for k in re:
            sj = wigner_6j(ta, tb, tJ, tr, ts, 2 * k)
            if abs(sj) >= 1e-15:
                ce[j, k] = swap * ph_e * sj * (reduced_C_jj(ka, ks, k) * reduced_C_jj(kb, kr, k))
    return lo, (hi - lo) // 2 + 1, cd, ce, set(rd) | set(re)


# ---------------------------------------------------------------------------
# the device pair sums
# ---------------------------------------------------------------------------
_BLOCKS = (("pppp", "p", "p", 0, 0), ("qqqq", "q", "q", 1, 1),
           ("pqpq", "p", "q", 0, 1), ("qpqp", "q", "p", 1, 0))


class _PairSums:
    """Pair energies of occupied orbital pairs over one virtual space."""

    def __init__(self, bd, cache, kappa_list, virtuals, multipole_max, table_bytes=None):
        self.bd, self.cache, self.kl = bd, cache, list(kappa_list)
        self.ns = cache.ns
        self.multipole_max = multipole_max
        self.table_bytes = int(_TABLE_BYTES if table_bytes is None else table_bytes)
        vk, veps, cols, self.where = [], [], [], []
        for ci, kappa in enumerate(self.kl):
            for i, (P, Q, e) in enumerate(virtuals.get(kappa, [])):
                vk.append(ci)
                veps.append(float(e))
                cols.append(np.r_[self._pad(P), self._pad(Q)])
                self.where.append((kappa, i))
        self.n_virt = len(vk)
        self.vk_host = np.asarray(vk, dtype=np.int32)
        self.veps_host = np.asarray(veps, dtype=np.float64)
        self.count = {kappa: len(virtuals.get(kappa, [])) for kappa in self.kl}
        if self.n_virt:
            self.C = dv.f64(np.column_stack(cols))           # (2ns, N)
            self.Ct = self.C.T.contiguous()
            self.vk = dv.i32(self.vk_host)
            self.veps = dv.f64(self.veps_host)
        self.types = {}

    def _pad(self, x):
        x = np.asarray(x, dtype=np.float64).ravel()
        out = np.zeros(self.ns)
        out[:min(x.size, self.ns)] = x[:self.ns]
        return out

    def _type(self, ka, kb):
        """Channel tables of one occupied pair type."""
        if (ka, kb) in self.types:
            return self.types[ka, kb]
        nk = len(self.kl)
        chans = {}
        for ir, kr in enumerate(self.kl):
            for is_ in range(ir, nk):
                ks = self.kl[is_]
                if not self.count[kr] or not self.count[ks]:
                    continue
                c = _channel(ka, kb, kr, ks, self.multipole_max)
                if c is not None:
                    if c[1] > _MAX_J:
                        raise ValueError(f"occupied pair ({ka}, {kb}) couples {c[1]} J values; "
                                         f"the pair-sum kernel holds {_MAX_J}")
                    chans[ir, is_] = c
        ranks = sorted(set().union(*[c[4] for c in chans.values()])) if chans else []
        self.types[ka, kb] = (chans, ranks)
        return chans, ranks

    def energies(self, occupied, pairs, want_terms=False):
        """occupied: [(P, Q, eps, kappa)]; pairs: [(ia, ib, same)] ->
        (pair energies, terms or None) with terms[p] = (mask, jsum) host
        arrays of the counted r <= s entries."""
        out = np.zeros(len(pairs))
        terms = [None] * len(pairs) if want_terms else None
        if not pairs or not self.n_virt:
            return out, terms
        types = [self._type(occupied[a][3], occupied[b][3]) for a, b, _ in pairs]
        ranks_all = sorted(set().union(*[set(t[1]) for t in types]))
        if not ranks_all:
            return out, terms
        n_rank = ranks_all[-1] + 1
        tkeys = list(dict.fromkeys((occupied[a][3], occupied[b][3]) for a, b, _ in pairs))
        chan_nj, chan_lo, cd, ce = self._coefficients(tkeys, n_rank)
        type_index = {t: i for i, t in enumerate(tkeys)}
        stage1 = self._stage1(occupied, ranks_all)
        N = self.n_virt
        per_pair = max(1, sum(1 for _ in ranks_all))
        group = max(1, min(len(pairs), self.table_bytes // max(1, per_pair * N * N * 8)))
        for g0 in range(0, len(pairs), group):
            sel = list(range(g0, min(len(pairs), g0 + group)))
            items, pair_item = [], np.full((len(sel), n_rank), -1, dtype=np.int32)
            for li, p in enumerate(sel):
                for k in types[p][1]:
                    pair_item[li, k] = len(items)
                    items.append((pairs[p][0], pairs[p][1], k))
            e, js = self._group(occupied, [pairs[p] for p in sel], items, pair_item, stage1,
                                [type_index[(occupied[pairs[p][0]][3], occupied[pairs[p][1]][3])]
                                 for p in sel], chan_nj, chan_lo, cd, ce, n_rank, want_terms)
            out[sel] = e
            if want_terms:
                for li, p in enumerate(sel):
                    terms[p] = js[li]
        return out, terms

    def _coefficients(self, tkeys, n_rank):
        nk = len(self.kl)
        chan_nj = np.zeros((len(tkeys), nk, nk), dtype=np.int32)
        chan_lo = np.zeros((len(tkeys), nk, nk), dtype=np.int32)
        cd = np.zeros((len(tkeys), nk, nk, _MAX_J, n_rank))
        ce = np.zeros((len(tkeys), nk, nk, _MAX_J, n_rank))
        for t, key in enumerate(tkeys):
            for (ir, is_), (lo, nj, d, e, _) in self.types[key][0].items():
                chan_nj[t, ir, is_] = nj
                chan_lo[t, ir, is_] = lo
                for (j, k), v in d.items():
                    cd[t, ir, is_, j, k] = v
                for (j, k), v in e.items():
                    ce[t, ir, is_, j, k] = v
        return chan_nj, chan_lo, cd, ce

    def _stage1(self, occupied, ranks):
        """Case-2 contractions of the band densities (a, e_m) of every
        occupied orbital a: {(k, block): (n_occ ns, ns, ks_right)}."""
        ns, cache = self.ns, self.cache
        rows = {}
        for comp, ci in (("p", 0), ("q", 1)):
            A = np.zeros((ns + len(occupied), ns))
            A[:ns] = np.eye(ns)
            for a, o in enumerate(occupied):
                A[ns + a] = self._pad(o[ci])
            rows[comp] = dv.f64(A)
        dens = {}
        out = {}
        for itype, left, right, _, _ in _BLOCKS:
            nl = {"pppp": cache.nsp, "qqqq": cache.nsq, "pqpq": min(ns, cache.nsp),
                  "qpqp": min(ns, cache.nsq)}[itype]
            key = (left, nl)
            if key not in dens:
                pairs = [(ns + a, m) for a in range(len(occupied)) for m in range(ns)]
                dens[key] = density_sym_device(rows[left], pairs, nl, cache.order(left))
            for k in ranks:
                out[k, itype] = cache.case2(int(k), itype, dens[key], nl)
        self._rows = rows  # the right coefficient rows of stage 2
        return out

    def _group(self, occupied, pairs, items, pair_item, stage1, ptype, chan_nj, chan_lo, cd, ce,
               n_rank, want_terms):
        torch = dv.torch()
        ns, cache = self.ns, self.cache
        B, N, n2 = len(items), self.n_virt, 2 * self.ns
        Mh = dv.empty((n2, B * n2))   # [M_1 | M_2 | ...]
        base = dv.ptr(Mh)
        nright = {"pppp": cache.nsp, "qqqq": cache.nsq, "pqpq": min(ns, cache.nsq),
                  "qpqp": min(ns, cache.nsp)}
        jobs = []
        for it, (a, b, k) in enumerate(items):
            for itype, left, right, ro, co in _BLOCKS:
                T1 = stage1[k, itype]
                ksr = T1.shape[2]
                t_addr = dv.ptr(T1) + a * ns * ns * ksr * 8
                wb = dv.ptr(self._rows[right]) + (ns + b) * dv.ld(self._rows[right]) * 8
                m_addr = base + (ro * ns * B * n2 + it * n2 + co * ns) * 8
                jobs.append((t_addr, wb, m_addr, ksr, nright[itype]))
        jobs = np.asarray(jobs, dtype=np.int64)
        jobs_dev = []
        for j0 in range(0, len(jobs), 65535):
            jd = dv.to_dev(np.ascontiguousarray(jobs[j0:j0 + 65535]))
            jobs_dev.append(jd)  # referenced until the group's work is queued
            call("ax_bilinear_stage2_batch", ns, ns, B * n2, len(jd), dv.ptr(jd),
                 dv.stream_handle())
        X = torch.matmul(self.Ct, Mh)                      # (N, B 2ns): C^T M_b side by side
        T = torch.matmul(X.view(N * B, n2), self.C)         # (N B, N): [r][item][s]
        P = len(pairs)
        pe = dv.f64([occupied[a][2] + occupied[b][2] for a, b, _ in pairs])
        pn = dv.f64([2.0 if same else 1.0 for _, _, same in pairs])
        pt = dv.i32(np.asarray(ptype, dtype=np.int32))
        pi = dv.i32(pair_item)
        cnj, clo = dv.i32(chan_nj), dv.i32(chan_lo)
        dcd, dce = dv.f64(cd), dv.f64(ce)
        ws = dv.empty((max(1, int(load().ax_mbpt2_workspace_doubles(N, P))),))
        energy = dv.empty((P,))
        js = dv.zeros((P, N, N)) if want_terms else None
        call("ax_mbpt2_pair_sum", N, len(self.kl), n_rank, P, _MAX_J, dv.ptr(self.vk),
             dv.ptr(self.veps), dv.ptr(T), N, B * N, dv.ptr(pi), dv.ptr(pe), dv.ptr(pn),
             dv.ptr(pt), dv.ptr(cnj), dv.ptr(clo), dv.ptr(dcd), dv.ptr(dce), dv.ptr(energy),
             dv.ptr(js) if js is not None else None, dv.ptr(ws), dv.stream_handle())
        e = dv.host(energy)   # synchronises: the group's buffers are released after this
        if not want_terms:
            return e, None
        jh = dv.host(js)
        out = []
        for p, (a, b, _) in enumerate(pairs):
            de = self.veps_host[:, None] + self.veps_host[None, :] - pe_host(occupied, a, b)
            t = ptype[p]
            active = chan_nj[t][self.vk_host[:, None], self.vk_host[None, :]] > 0
            mask = np.triu(active & (np.abs(de) >= 1e-10))
            out.append((mask, jh[p]))
        return e, out


def pe_host(occupied, a, b):
    return occupied[a][2] + occupied[b][2]


# ---------------------------------------------------------------------------
# entry point
# ---------------------------------------------------------------------------
def _kappa_list(kappa_max):
    out = []
    for k in range(1, kappa_max + 2):
        out += [-k, k]
    return [k for k in out if l_of_kappa(k) <= kappa_max]


def mbpt2_energy(scf_result, *, kappa_max: int = 4, n_virt_max: Optional[int] = None,
                 multipole_max: Optional[int] = None, e_virt_max: Optional[float] = None,
                 verbose: bool = False, term_recorder=None) -> MBPT2Result:
    """Second-order correlation energy of a closed-shell Dirac-Fock reference
    (mbpt.py:411-588): V^N virtuals of every kappa with l <= ``kappa_max``
    (the lowest ``n_virt_max`` per kappa, below ``e_virt_max``), occupied
    energies from the SCF, and the J-coupled antisymmetrised pair sum over
    unordered virtual channels, ranks capped by ``multipole_max``.
    ``term_recorder`` is called once per counted virtual pair of every
    occupied pair with the reference's record keys."""
    t0 = time.time()
    ref = _Reference(scf_result)
    if not ref.converged:
        raise ValueError("SCFResult must be converged.")
    for sh in ref.shells:
        if abs(sh.occupation - (twoj_of_kappa(sh.kappa) + 1)) > 0.01:
            raise ValueError(f"MBPT(2) requires a closed-shell reference; subshell {sh.label} has "
                             f"occupation {sh.occupation} (expected "
                             f"{twoj_of_kappa(sh.kappa) + 1}).")
    cache = ConvolCache(ref.bd)
    parts = _CoreParts(ref, cache)
    kl = _kappa_list(kappa_max)
    occ_basis, virt_basis = {}, {}
    for kappa in kl:
        occ, virt = _virtual_basis(ref, kappa, parts, e_virt_max, "VN")
        subs = [sh for sh in ref.shells if sh.kappa == kappa]
        rows = []
        for i, (P, Q, e) in enumerate(occ):
            if i < len(subs):
                lab = subs[i].label
                eps = ref.energies.get(lab)
                rows.append((P, Q, e if eps is None else float(eps), lab))
            else:
                rows.append((P, Q, e, f"occ-{kappa}-{i}"))
        occ_basis[kappa] = rows
        virt_basis[kappa] = virt[:n_virt_max] if n_virt_max is not None else virt
    occupied, labels, pairs = [], [], []
    okl = [k for k in kl if occ_basis[k]]
    index = {}
    for kappa in okl:
        for i, (P, Q, e, lab) in enumerate(occ_basis[kappa]):
            index[kappa, i] = len(occupied)
            occupied.append((P, Q, e, kappa))
            labels.append(lab)
    for ia, ka in enumerate(okl):
        for kb in okl[ia:]:
            for i in range(len(occ_basis[ka])):
                for j in range(i if ka == kb else 0, len(occ_basis[kb])):
                    pairs.append((index[ka, i], index[kb, j], ka == kb and i == j))
    engine = _PairSums(ref.bd, cache, kl, virt_basis, multipole_max)
    energies, terms = engine.energies(occupied, pairs, want_terms=term_recorder is not None)
    contributions = {}
    total = 0.0
    for p, (a, b, same) in enumerate(pairs):
        total += float(energies[p])
        contributions[(labels[a], labels[b])] = float(energies[p])
        if verbose:
            logger.debug(f"  ({labels[a]}, {labels[b]})  dE(2) = {energies[p]:.6e} Ha")
        if term_recorder is not None:
            _record(term_recorder, engine, occupied, labels, a, b, same, terms[p])
    return MBPT2Result(
        correlation_energy=total, pair_contributions=contributions,
        n_virtuals={k: len(v) for k, v in virt_basis.items()}, kappa_max=kappa_max,
        n_max=n_virt_max, cpu_seconds=time.time() - t0,
        provenance={"multipole_max": multipole_max, "e_virt_max": e_virt_max,
                    "kappa_list": kl, "Z": ref.Z, "n_subshells": len(ref.shells)})


def _record(recorder, engine, occupied, labels, a, b, same, term):
    """The reference's per-term records (mbpt.py:772-788), in its channel
    and index order."""
    mask, jsum = term
    ea, eb = occupied[a][2], occupied[b][2]
    norm_ab = 2.0 if same else 1.0
    order = sorted(zip(*np.nonzero(mask)),
                   key=lambda rs: (engine.vk_host[rs[0]], engine.vk_host[rs[1]], rs[0], rs[1]))
    for r, s in order:
        kr, ir = engine.where[r]
        ks, is_ = engine.where[s]
        er, es = float(engine.veps_host[r]), float(engine.veps_host[s])
        de = er + es - ea - eb
        norm = norm_ab * (2.0 if r == s else 1.0)
        num = float(jsum[r, s])
        recorder({"kappa_a": occupied[a][3], "kappa_b": occupied[b][3], "kappa_r": kr,
                  "kappa_s": ks, "i_r": ir, "i_s": is_, "eps_a": ea, "eps_b": eb, "eps_r": er,
                  "eps_s": es, "delta_eps": de, "numerator": num, "norm": norm,
                  "contribution": -num / (norm * de), "a": labels[a], "b": labels[b]})
