"""Signed atomic spinor integrals for tensor correlation engines (CCSD /
FSCC): drop-in for physics/csf/spinor_integrals.py (``packed_size``,
``normal_ordered_one_body``, ``SpinorIntegrals``, ``OnDemandSpinorIntegrals``,
``build_spinor_integrals``).

The Hamiltonian is E_frozen + h_eff[p,q] a+_p a_q + 1/4 g[p,q,r,s] a+_p a+_q
a_s a_r over the complete magnetic multiplets of the retained radial
orbitals (occupied, then virtual); frozen electrons enter the constant and
the one-body operator, correlated occupied spinors the Fock operator.

Device path (one export):
  1. the radial table R^k(target pair, source pair) of every rank over all
     unordered pairs of the retained and frozen radial orbitals: the banded
     pair densities and cell pair potentials of the CI evaluator
     (``ci_radial``), then per rank and component one FP64 GEMM
     densities x potentials^T (the evaluator's banded dots, all at once);
  2. per-electron angular factors T_k[a][c] (host, exact-enough 3j and
     reduced C^k, Grant's convention) uploaded once;
  3. ``ax_spinor_eri_packed`` fills the packed signed integrals (one wave
     per (p, q) block, written coalesced), ``ax_spinor_eri_list`` the
     quartets of the Fock, frozen-core and reference-energy sums;
  4. the one-body Dirac matrix elements <a|h|b>/<b|b> on the host (FP64,
     as the CI evaluator's "h" keys).
Only the complex packed copy the interface returns lives on the host.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass
from math import comb
from types import MappingProxyType

import numpy as np

from . import device as dv
from ._lib import call

__all__ = ["packed_size", "normal_ordered_one_body", "SpinorIntegrals",
           "OnDemandSpinorIntegrals", "build_spinor_integrals"]


# ---------------------------------------------------------------------------
# packed layout (spinor_integrals.py:20-50)
# ---------------------------------------------------------------------------
def packed_size(n):
    """Values of the packed signed-integral array of n spinors: block (p, q),
    p < q, holds every (r, s) with p <= r < s < n."""
    return sum((n - p - 1) * comb(n - p, 2) for p in range(n - 1))


def _blocks(n):
    """(p, q, offset) of every block, in storage order."""
    p, q = np.triu_indices(n, 1)
    sizes = np.array([comb(n - int(x), 2) for x in range(n)], dtype=np.int64)[p]
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    return p.astype(np.int32), q.astype(np.int32), off


def _lookup(n, offsets, p, q, r, s):
    """Positions, signs and conjugation flags of integrals (p, q, r, s)
    (arrays) in the packed layout; position -1 for an index pair p == q or
    r == s (the value is zero)."""
    p, q, r, s = (np.asarray(x, dtype=np.int64) for x in (p, q, r, s))
    zero = (p == q) | (r == s)
    sign = np.ones(p.shape)
    sw = p > q
    p, q = np.where(sw, q, p), np.where(sw, p, q)
    sign[sw] *= -1
    sw = r > s
    r, s = np.where(sw, s, r), np.where(sw, r, s)
    sign[sw] *= -1
    conj = p > r
    p, q, r, s = (np.where(conj, r, p), np.where(conj, s, q), np.where(conj, p, r),
                  np.where(conj, q, s))
    pq = n * p + q - (p + 1) * (p + 2) // 2
    m, rx, sx = n - p, r - p, s - p
    rs = m * rx + sx - (rx + 1) * (rx + 2) // 2
    pos = np.where(zero, -1, offsets[np.where(zero, 0, pq)] + rs)
    return pos, sign, conj


def _gather(packed, n, offsets, p, q, r, s):
    pos, sign, conj = _lookup(n, offsets, p, q, r, s)
    v = packed[np.maximum(pos, 0)]
    v = np.where(conj, np.conj(v), v) * sign
    return np.where(pos < 0, 0.0, v)


def _check_index(n, *idx):
    for x in idx:
        if np.any(np.asarray(x) < 0) or np.any(np.asarray(x) >= n):
            raise IndexError("spinor integral index out of range")


def normal_ordered_one_body(fock, packed, noccupied):
    """h_eff = F - sum_i g[p,i,q,i] over the first ``noccupied`` spinors."""
    fock = np.asarray(fock, dtype=complex)
    n = len(fock)
    offsets = _blocks(n)[2]
    p, q = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    h = fock.copy()
    for i in range(noccupied):
        h -= _gather(np.asarray(packed), n, offsets, p, np.full_like(p, i), q,
                     np.full_like(p, i))
    return h


def _owned(a, shape):
    a = np.asarray(a, dtype=np.complex128)
    if a.shape != shape or not np.all(np.isfinite(a)):
        raise ValueError("nonfinite or malformed integral arrays")
    return np.frombuffer(a.tobytes(), dtype=a.dtype).reshape(shape)  # read-only


@dataclass(frozen=True)
class SpinorIntegrals:
    """Integrals resident as one packed array (spinor_integrals.py:64-127):
    construction checks the Fock operator and reference energy against the
    packed integrals and the conjugacy of the duplicated entries."""
    labels: tuple
    noccupied: int
    one_body_Ha: np.ndarray
    fock_Ha: np.ndarray
    packed_eri_Ha: np.ndarray
    frozen_energy_Ha: float
    reference_energy_Ha: float
    identity: object

    def __post_init__(self):
        n = len(self.labels)
        if not 0 < self.noccupied < n or len(set(self.labels)) != n:
            raise ValueError("require unique spinors and occupied/virtual spaces")
        for name, shape in (("one_body_Ha", (n, n)), ("fock_Ha", (n, n)),
                            ("packed_eri_Ha", (packed_size(n),))):
            a = _owned(getattr(self, name), shape)
            if a.ndim == 2 and not np.array_equal(a, a.conj().T):
                raise ValueError("one-body and Fock operators must be Hermitian")
            object.__setattr__(self, name, a)
        if not np.isfinite(self.frozen_energy_Ha + self.reference_energy_Ha):
            raise ValueError("nonfinite energy constants")
        object.__setattr__(self, "_offsets", _blocks(n)[2])
        object.__setattr__(self, "identity", MappingProxyType(dict(self.identity)))
        o = self.noccupied
        p, q = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        fock = self.one_body_Ha.copy()
        for i in range(o):
            fock += self._eri_arrays(p, np.full_like(p, i), q, np.full_like(p, i))
        err = float(np.max(np.abs(fock - self.fock_Ha)))
        i, j = np.meshgrid(np.arange(o), np.arange(o), indexing="ij")
        energy = (self.frozen_energy_Ha + np.trace(self.one_body_Ha[:o, :o])
                  + 0.5 * np.sum(self._eri_arrays(i, j, i, j)))
        if err > 1e-9 or abs(energy - self.reference_energy_Ha) > 1e-9:
            raise ValueError("Fock/reference energy do not match signed normal-ordered Hamiltonian")
        # entries (p, q, p, s) and (p, s, p, q) are both stored: conjugates
        pp, qq, ss = [], [], []
        for a in range(n - 1):
            b = np.arange(a + 1, n)
            qv, sv = np.meshgrid(b, b, indexing="ij")
            pp.append(np.full(qv.size, a))
            qq.append(qv.ravel())
            ss.append(sv.ravel())
        if pp:
            pp, qq, ss = (np.concatenate(x) for x in (pp, qq, ss))
            d = np.abs(self._eri_arrays(pp, qq, pp, ss) - np.conj(self._eri_arrays(pp, ss, pp, qq)))
            if d.size and float(d.max()) > 1e-10:
                raise ValueError("packed two-body duplicate entries are not Hermitian")

    def _eri_arrays(self, p, q, r, s):
        return _gather(self.packed_eri_Ha, len(self.labels), self._offsets, p, q, r, s)

    def eri(self, p, q, r, s):
        _check_index(len(self.labels), p, q, r, s)
        return complex(self._eri_arrays([p], [q], [r], [s])[0])

    @property
    def dense_eri_bytes(self):
        return self.packed_eri_Ha.nbytes

    def diagnostics(self):
        return dict(representation="packed",
                    resident_matrix_bytes=self.one_body_Ha.nbytes + self.fock_Ha.nbytes,
                    resident_eri_bytes=self.packed_eri_Ha.nbytes,
                    dense_eri_bytes=self.dense_eri_bytes)

    def iter_packed_eri(self, *, max_chunk_bytes=4 * 1024 ** 2):
        if max_chunk_bytes < 16:
            raise ValueError("integral chunk must hold at least one complex value")
        count = int(max_chunk_bytes) // 16
        for begin in range(0, len(self.packed_eri_Ha), count):
            yield self.packed_eri_Ha[begin:begin + count]


def _pair_symmetry(labels, p, q):
    a, b = labels[p], labels[q]
    la = a[1] if a[1] > 0 else -a[1] - 1
    lb = b[1] if b[1] > 0 else -b[1] - 1
    return a[2] + b[2], (la + lb) % 2


@dataclass(frozen=True)
class OnDemandSpinorIntegrals:
    """Only the two one-body matrices resident on the host
    (spinor_integrals.py:139-255): single integrals are evaluated from the
    radial and angular tables, and ``iter_packed_eri`` streams the packed
    order in bounded complex chunks computed block range by block range on
    the device."""
    labels: tuple
    noccupied: int
    one_body_Ha: np.ndarray
    fock_Ha: np.ndarray
    frozen_energy_Ha: float
    reference_energy_Ha: float
    identity: object
    _evaluate: object
    _diagnostics: object

    def __post_init__(self):
        n = len(self.labels)
        if not 0 < self.noccupied < n or len(set(self.labels)) != n:
            raise ValueError("require unique spinors and occupied/virtual spaces")
        for name in ("one_body_Ha", "fock_Ha"):
            a = _owned(getattr(self, name), (n, n))
            if not np.array_equal(a, a.conj().T):
                raise ValueError("nonfinite, malformed or non-Hermitian one-body operator")
            for p, row in enumerate(a):
                if any(self.labels[q][1:] != self.labels[p][1:] for q in np.flatnonzero(row)):
                    raise ValueError("native spherical one-body operator has nonscalar entries")
            object.__setattr__(self, name, a)
        if not np.isfinite(self.frozen_energy_Ha + self.reference_energy_Ha):
            raise ValueError("nonfinite energy constants")
        object.__setattr__(self, "identity", MappingProxyType(dict(self.identity)))
        radial = {}
        for i, label in enumerate(self.labels):
            radial.setdefault(label[:2], {})[label[2]] = i
        for key, multiplet in radial.items():
            if set(multiplet) != set(range(-2 * abs(key[1]) + 1, 2 * abs(key[1]), 2)):
                raise ValueError("native operator requires complete magnetic multiplets")
            occupied = [i < self.noccupied for i in multiplet.values()]
            if any(occupied) and not all(occupied):
                raise ValueError("native reference requires a closed spherical core")
        o = self.noccupied
        energy = self.frozen_energy_Ha + np.trace(self.one_body_Ha[:o, :o]) + 0.5 * sum(
            self.eri(i, j, i, j) for i in range(o) for j in range(o))
        if abs(energy - self.reference_energy_Ha) > 1e-9:
            raise ValueError("native reference energy failed")

    @property
    def dense_eri_bytes(self):
        return 16 * packed_size(len(self.labels))

    def diagnostics(self):
        return dict(self._diagnostics(), representation="on_demand",
                    resident_matrix_bytes=self.one_body_Ha.nbytes + self.fock_Ha.nbytes,
                    resident_eri_bytes=0, dense_eri_bytes=self.dense_eri_bytes)

    def eri(self, p, q, r, s):
        n = len(self.labels)
        _check_index(n, p, q, r, s)
        if p == q or r == s:
            return 0j
        sign = 1
        if p > q:
            p, q, sign = q, p, -sign
        if r > s:
            r, s, sign = s, r, -sign
        conj = p > r
        if conj:
            p, q, r, s = r, s, p, q
        if _pair_symmetry(self.labels, p, q) != _pair_symmetry(self.labels, r, s):
            return 0j
        value = complex(self._evaluate.one(p, q, r, s))
        return sign * (value.conjugate() if conj else value)

    def iter_packed_eri(self, *, max_chunk_bytes=4 * 1024 ** 2):
        """The packed order, block (p, q) after block, each block in chunks of
        at most max_chunk_bytes (complex values)."""
        if max_chunk_bytes < 16:
            raise ValueError("integral chunk must hold at least one complex value")
        count = int(max_chunk_bytes) // 16
        n = len(self.labels)
        for blocks in self._evaluate.block_batches(max(count, 1 << 16)):
            for length, values in blocks:
                for begin in range(0, length, count):
                    yield values[begin:begin + count].astype(complex)


# ---------------------------------------------------------------------------
# the export (spinor_integrals.py:258-466)
# ---------------------------------------------------------------------------
def _multiplets(keys):
    return tuple((n, k, m) for n, k in keys for m in range(-2 * abs(k) + 1, 2 * abs(k), 2))


class _Tables:
    """Device tables of one export: spinors (retained, then frozen), their
    radial orbitals and 2m, the radial pair table and the angular factors."""

    def __init__(self, selected, spinors, retained, bsp_data, ns):
        import torch

        from .angular import reduced_c_jj, twoj_of_kappa, wigner_3j
        from .ci_radial import pair_potentials_device
        from .convol import ConvolCache, density_sym_device

        keys = sorted({s[:2] for s in spinors})
        index = {k: i for i, k in enumerate(keys)}
        nr = len(keys)
        P = np.zeros((nr, ns))
        Q = np.zeros((nr, ns))
        for k, i in index.items():
            p, q = (np.asarray(x, dtype=np.float64).ravel() for x in selected[k][:2])
            P[i, :min(p.size, ns)] = p[:ns]
            Q[i, :min(q.size, ns)] = q[:ns]
        self.n = len(spinors)
        self.retained = int(retained)  # packed over the first `retained` spinors
        self.n_rad = nr
        ii, jj = np.triu_indices(nr)
        pairs = list(zip(ii.tolist(), jj.tolist()))
        self.n_rpair = len(pairs)
        pid = np.zeros((nr, nr), dtype=np.int32)
        pid[ii, jj] = np.arange(len(pairs))
        pid[jj, ii] = np.arange(len(pairs))
        ksp = int(bsp_data.get("ksp", bsp_data["ks"]))
        ksq = int(bsp_data.get("ksq", bsp_data["ks"]))
        tj = [twoj_of_kappa(s[1]) for s in spinors]
        self.kmax = max(tj)  # (j_a + j_c) <= 2 max j
        # radial table: per rank, R = D_p V_p^T + D_q V_q^T over banded
        # entries i + s < ns (the evaluator's banded dots)
        cache = ConvolCache(bsp_data)
        Pd, Qd = dv.f64(P), dv.f64(Q)
        dp = density_sym_device(Pd, pairs, ns, ksp)
        dq = density_sym_device(Qd, pairs, ns, ksq)
        band = lambda ks: dv.f64((np.arange(ns)[:, None] + np.arange(ks)[None, :] < ns)  # noqa: E731
                                 .astype(np.float64))
        mp, mq = band(ksp), band(ksq)
        dpf = (dp * mp).reshape(len(pairs), -1)
        dqf = (dq * mq).reshape(len(pairs), -1)
        R = dv.empty((self.kmax + 1, len(pairs), len(pairs)))
        for k in range(self.kmax + 1):
            vp, vq = pair_potentials_device(cache, k, dp, dq)
            R[k] = torch.matmul(dpf, (vp * mp).reshape(len(pairs), -1).T)
            R[k] += torch.matmul(dqf, (vq * mq).reshape(len(pairs), -1).T)
        self.R = R.contiguous()
        # angular factors
        n = self.n
        T = np.zeros((self.kmax + 1, n, n))
        for a, (na, ka, ma) in enumerate(spinors):
            for c, (nc, kc, mc) in enumerate(spinors):
                ph = -1.0 if ((tj[a] - ma) // 2) % 2 else 1.0
                for k in range(self.kmax + 1):
                    ck = reduced_c_jj(ka, kc, k)
                    if ck == 0.0:
                        continue
                    w = wigner_3j(tj[a], 2 * k, tj[c], -ma, ma - mc, mc)
                    if w != 0.0:
                        T[k, a, c] = ph * w * ck
        self.T_host = T
        self.R_host = None
        self.T = dv.f64(T)
        self.rad_host = np.array([index[s[:2]] for s in spinors], dtype=np.int32)
        self.m2_host = np.array([s[2] for s in spinors], dtype=np.int32)
        self.pid_host = pid
        self.rad = dv.i32(self.rad_host)
        self.m2 = dv.i32(self.m2_host)
        self.pid = dv.i32(pid)
        self.err = dv.zeros((1,))

    def _args(self):
        return (self.n, self.kmax, self.n_rad, self.n_rpair, dv.ptr(self.rad), dv.ptr(self.m2),
                dv.ptr(self.pid), dv.ptr(self.T), dv.ptr(self.R))

    def packed(self, n, blocks=None, out=None):
        """Device packed integrals of the first n spinors (all blocks, or the
        given (p, q, offset) arrays written from their first offset)."""
        bp, bq, boff = _blocks(n) if blocks is None else blocks
        total = packed_size(n) if blocks is None else None
        base = int(boff[0]) if len(boff) else 0
        if out is None:
            out = dv.zeros((max(total, 1),))
        # the uploaded index arrays stay referenced until the launch is queued
        # (a released block is handed to the next upload on the same stream)
        dp, dq, do = dv.i32(bp), dv.i32(bq), dv.to_dev(np.asarray(boff, dtype=np.int64))
        call("ax_spinor_eri_packed", *self._args(), int(n), len(bp), dv.ptr(dp), dv.ptr(dq),
             dv.ptr(do), base, dv.ptr(out), dv.ptr(self.err), dv.stream_handle())
        self._keep = (dp, dq, do)
        return out

    def quartets(self, quad):
        quad = np.ascontiguousarray(quad, dtype=np.int32).reshape(-1, 4)
        out = dv.empty((max(len(quad), 1),))
        dquad = dv.i32(quad)
        call("ax_spinor_eri_list", *self._args(), len(quad), dv.ptr(dquad), dv.ptr(out),
             dv.ptr(self.err), dv.stream_handle())
        host = dv.host(out)[:len(quad)]  # synchronises: dquad may go after this
        return host

    def max_mismatch(self):
        return float(dv.host(self.err)[0])

    def one(self, p, q, r, s):
        """g(p, q, r, s) on the host from the table copies."""
        if self.R_host is None:
            self.R_host = dv.host(self.R)
        T, R, rad, m2, pid = self.T_host, self.R_host, self.rad_host, self.m2_host, self.pid_host

        def direct(a, b, c, d):
            sign = -1.0 if ((m2[a] - m2[c]) // 2) % 2 else 1.0
            return sign * float(np.sum(T[:, a, c] * T[:, b, d]
                                       * R[:, pid[rad[a], rad[c]], pid[rad[b], rad[d]]]))

        if m2[p] + m2[q] != m2[r] + m2[s]:
            return 0.0
        x = direct(p, q, r, s) - direct(p, q, s, r)
        y = direct(r, s, p, q) - direct(r, s, q, p)
        if abs(x - y) > 1e-10:
            raise ValueError(f"two-electron Hermitian gate failed: {abs(x - y)}")
        return 0.5 * (x + y)

    def block_batches(self, max_values):
        """(length, host values) of every packed block of the retained
        spinors, computed on the device in batches of blocks of about
        max_values values."""
        n = self.retained
        bp, bq, boff = _blocks(n)
        sizes = np.array([comb(n - int(p), 2) for p in bp], dtype=np.int64)
        b = 0
        while b < len(bp):
            e, acc = b, 0
            while e < len(bp) and (e == b or acc + sizes[e] <= max_values):
                acc += int(sizes[e])
                e += 1
            vals = dv.host(self.packed(n, (bp[b:e], bq[b:e], boff[b:e]),
                                       out=dv.zeros((max(acc, 1),))))
            if self.max_mismatch() > 1e-10:
                raise ValueError(f"two-electron Hermitian gate failed: {self.max_mismatch()}")
            out, o = [], 0
            for s in sizes[b:e]:
                out.append((int(s), vals[o:o + int(s)]))
                o += int(s)
            yield out
            b = e


def build_spinor_integrals(orbitals, workspace, Z, ns, *, occupied_keys, virtual_keys,
                           frozen_keys=(), nuclear_potential=None, max_integral_bytes=2 * 1024 ** 3,
                           metric_tolerance=1e-8, representation="packed",
                           radial_pair_cache_size=1024, include_breit=False,
                           breit_orbital_energies=None, transverse_cache_size=128):
    """Signed integrals of the complete declared orthonormal B-spline atomic
    space (spinor_integrals.py:258-466): retained spinors are the complete
    multiplets of ``occupied_keys`` then ``virtual_keys`` ((n, kappa) keys of
    ``orbitals``: (P, Q) coefficient pairs); ``frozen_keys`` enter the
    constant and one-body operator.  Every retained or frozen orbital must be
    orthonormal in the one-body metric of its kappa (``metric_tolerance``);
    ``max_integral_bytes`` caps the complex payload (2 n^2 one-body values,
    plus the packed integrals for ``representation="packed"``).  The
    Dirac-Coulomb operator only: the transverse (Breit) terms of the
    reference's ``include_breit`` are not built here and are refused."""
    from .dirac import dirac_matrices

    if representation not in ("packed", "on_demand"):
        raise ValueError("integral representation must be packed or on_demand")
    if breit_orbital_energies is not None and not include_breit:
        raise ValueError("breit_orbital_energies requires include_breit=True")
    if include_breit:
        raise NotImplementedError("the transverse (Breit) terms of the spinor export are not "
                                  "built by this engine (Dirac-Coulomb only)")
    groups = [tuple(tuple(k) for k in keys) for keys in (occupied_keys, virtual_keys, frozen_keys)]
    occ, vir, frozen = groups
    keys = occ + vir + frozen
    if (not occ or not vir or len(set(keys)) != len(keys)
            or any(k not in orbitals or len(k) != 2 or k[1] == 0 for k in keys)):
        raise ValueError("require disjoint occupied/virtual/frozen orbital keys")
    labels = _multiplets(occ + vir)
    core = _multiplets(occ + frozen)
    frozen_labels = _multiplets(frozen)
    n = len(labels)
    required = 16 * (2 * n * n + (packed_size(n) if representation == "packed" else 0))
    if required > max_integral_bytes or max_integral_bytes < 1:
        raise MemoryError(f"complete integral space needs {required} bytes; cap "
                          f"{max_integral_bytes}")
    selected = {k: tuple(orbitals[k]) for k in keys}
    # metric: retained, frozen and virtual orbitals of each kappa orthonormal
    one_body = {}
    metric_error = 0.0
    for kappa in sorted({k for _, k in keys}):
        H0, S = dirac_matrices(kappa, Z, workspace, nuclear_potential)
        one_body[kappa] = (H0, S)
        subset = [key for key in keys if key[1] == kappa]
        m = S.shape[0] // 2
        cols = []
        for key in subset:
            p, q = (np.asarray(x, dtype=np.float64).ravel() for x in selected[key][:2])
            v = np.zeros(2 * m)
            v[:min(p.size, m)] = p[:m]
            v[m:m + min(q.size, m)] = q[:m]
            cols.append(v)
        a = np.column_stack(cols)
        metric_error = max(metric_error, float(np.max(np.abs(a.T @ S @ a - np.eye(len(subset))))))
    if metric_error > metric_tolerance:
        raise ValueError(f"common orbital metric failed: {metric_error}")

    spinors = labels + frozen_labels  # device numbering: retained, then frozen
    pos = {s: i for i, s in enumerate(spinors)}
    tab = _Tables(selected, spinors, n, workspace, int(ns))
    # one-body Dirac elements <a|h|b>/<b|b> (the CI evaluator's "h" keys)
    vec = {}
    for key in keys:
        H0, S = one_body[key[1]]
        m = S.shape[0] // 2
        p, q = (np.asarray(x, dtype=np.float64).ravel() for x in selected[key][:2])
        v = np.zeros(2 * m)
        v[:min(p.size, m)] = p[:m]
        v[m:m + min(q.size, m)] = q[:m]
        vec[key] = v
    hcache = {}

    def h(a, b):
        if a[1:] != b[1:]:
            return 0.0
        key = (a[:2], b[:2])
        if key not in hcache:
            H0, S = one_body[a[1]]
            ket = vec[b[:2]]
            nrm = ket @ S @ ket
            hcache[key] = (vec[a[:2]] @ H0 @ ket) / nrm if nrm > 0.0 else float("nan")
        return hcache[key]

    # Fock / frozen / reference sums: one list launch
    core_idx = [pos[c] for c in core]
    frozen_idx = [pos[c] for c in frozen_labels]
    A, B = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    quads = [np.stack([A.ravel(), np.full(A.size, i), B.ravel(), np.full(A.size, i)], 1)
             for i in core_idx]
    ci, cj = np.meshgrid(core_idx, core_idx, indexing="ij")
    fi, fj = np.meshgrid(frozen_idx, frozen_idx, indexing="ij")
    quads.append(np.stack([ci.ravel(), cj.ravel(), ci.ravel(), cj.ravel()], 1))
    quads.append(np.stack([fi.ravel(), fj.ravel(), fi.ravel(), fj.ravel()], 1))
    allq = np.concatenate([q for q in quads if q.size]) if any(q.size for q in quads) else \
        np.zeros((0, 4), dtype=np.int32)
    vals = tab.quartets(allq)
    nf = n * n
    gsum_core = np.zeros((n, n))
    gsum_frozen = np.zeros((n, n))
    o = 0
    for i in core_idx:
        g = vals[o:o + nf].reshape(n, n)
        gsum_core += g
        if i >= n:  # frozen spinors follow the retained ones
            gsum_frozen += g
        o += nf
    e_core = 0.5 * float(np.sum(vals[o:o + len(core_idx) ** 2]))
    o += len(core_idx) ** 2
    e_frozen = 0.5 * float(np.sum(vals[o:o + len(frozen_idx) ** 2]))
    hmat = np.array([[h(a, b) for b in labels] for a in labels], dtype=np.float64)
    frozen_energy = sum(h(a, a) for a in frozen_labels) + e_frozen
    reference_energy = sum(h(a, a) for a in core) + e_core
    heff = (hmat + gsum_frozen).astype(complex)
    fock = (hmat + gsum_core).astype(complex)
    packed_dev = tab.packed(n) if representation == "packed" else None
    mismatch = tab.max_mismatch()
    if mismatch > 1e-10:
        raise ValueError(f"two-electron Hermitian gate failed: {mismatch}")
    max_projection = 0.5 * mismatch
    error = max(float(np.max(np.abs(heff - heff.T.conj()))),
                float(np.max(np.abs(fock - fock.T.conj()))))
    if error > 1e-10:
        raise ValueError("one-body Hermitian gate failed")
    max_projection = max(max_projection, 0.5 * error)
    heff, fock = 0.5 * (heff + heff.T.conj()), 0.5 * (fock + fock.T.conj())
    if representation == "on_demand":  # the reference zeroes off-multiplet one-body entries
        same = np.array([[a[1:] == b[1:] for b in labels] for a in labels])
        heff, fock = np.where(same, heff, 0.0), np.where(same, fock, 0.0)
    digest = hashlib.sha256()
    for key in keys:
        digest.update(repr(key).encode())
        for x in selected[key][:2]:
            digest.update(np.ascontiguousarray(x, dtype=np.float64).tobytes())
    digest.update(np.ascontiguousarray(workspace["gr"], dtype=np.float64).tobytes())
    digest.update(repr((Z, int(ns))).encode())
    identity = dict(hamiltonian="Dirac-Coulomb", binding=digest.hexdigest(),
                    occupied_keys=occ, virtual_keys=vir, frozen_keys=frozen,
                    metric_error=metric_error, hermitian_projection_max_Ha=max_projection,
                    integral_bytes=required, radial_pair_cache_size=radial_pair_cache_size,
                    radial_pairs=tab.n_rpair, ranks=tab.kmax + 1)
    if representation == "on_demand":
        def diagnostics():
            return dict(radial_pairs=tab.n_rpair, ranks=tab.kmax + 1,
                        table_bytes=int(tab.R.numel() + tab.T.numel()) * 8,
                        hermitian_projection_max_Ha=max_projection)
        return OnDemandSpinorIntegrals(labels, len(_multiplets(occ)), heff, fock, frozen_energy,
                                       reference_energy,
                                       dict(identity, representation="device radial/angular "
                                                                     "tables, Dirac-Coulomb"),
                                       tab, diagnostics)
    packed = dv.host(packed_dev)[:packed_size(n)].astype(complex)
    return SpinorIntegrals(labels, len(_multiplets(occ)), heff, fock, packed, frozen_energy,
                           reference_energy, identity)
