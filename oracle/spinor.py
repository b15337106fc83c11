"""The signed spinor-integral export restated on the CPU — TEST
INFRASTRUCTURE ONLY (the checker of atomx_amd.spinor_integrals; imported by
tests/ only, never by the product).

Follows physics/csf/spinor_integrals.py:258-466 (build_spinor_integrals,
packed representation, Dirac-Coulomb) with
  * R^k(ac, bd) = the oracle's scalar cell R^k (compute_slater_rk_cell with
    orbitals a, c, b, d: the CI evaluator's ("r", k, a, c, b, d) keys,
    physics/csf/matrix.py:723-817), memoised;
  * the angular terms of angular/csf_hamiltonian.py:874-920
    (two_body_angular) with the exact rational 3j of this oracle and the
    reduced C^k of angular/coefficients.py:29-65 (Grant's phase);
  * g(a,b,c,d) = ((D(abcd) - D(abdc)) + (D(cdab) - D(cdba))) / 2 and the
    one-body, frozen and reference sums of :401-432, in plain loops;
  * the one-body Dirac elements <a|h|b>/<b|b> from the product's host
    matrices (atomx_amd.dirac.dirac_matrices, pinned elsewhere by the Dirac
    1s closed form), as oracle/scf.py does.
Small spinor sets only (pure Python)."""

from __future__ import annotations

import math
from math import comb

import numpy as np

from . import Banks, compute_slater_rk_cell, wigner_3j_exact


def _l(kappa):
    return kappa if kappa > 0 else -kappa - 1


def _tj(kappa):
    return 2 * abs(kappa) - 1


def reduced_c(ka, kb, k):
    """angular/coefficients.py:29-65 with the exact 3j."""
    if (_l(ka) + k + _l(kb)) % 2:
        return 0.0
    ta, tb = _tj(ka), _tj(kb)
    if 2 * k < abs(ta - tb) or 2 * k > ta + tb:
        return 0.0
    w = wigner_3j_exact(ta, 2 * k, tb, -1, 0, 1)
    if abs(w) < 1e-15:
        return 0.0
    phase = 1.0 if ((ta + 1) // 2) % 2 == 0 else -1.0
    return phase * math.sqrt((ta + 1) * (tb + 1)) * w


def two_body_angular(ka, ma, kb, mb, kc, mc, kd, md):
    """angular/csf_hamiltonian.py:874-920: (k, coefficient) terms."""
    if ma + mb != mc + md:
        return ()
    q2 = ma - mc
    if q2 % 2:
        return ()
    q = q2 // 2
    ta, tb, tc, td = _tj(ka), _tj(kb), _tj(kc), _tj(kd)
    out = []
    for k in range(min((ta + tc) // 2, (tb + td) // 2) + 1):
        ca = reduced_c(ka, kc, k)
        if ca == 0.0:
            continue
        cb = reduced_c(kb, kd, k)
        if cb == 0.0:
            continue
        t1 = wigner_3j_exact(ta, 2 * k, tc, -ma, q2, mc)
        if t1 == 0.0:
            continue
        t2 = wigner_3j_exact(tb, 2 * k, td, -mb, -q2, md)
        if t2 == 0.0:
            continue
        e = (ta - ma) // 2 + (tb - mb) // 2 + q
        out.append((k, (-1.0 if e % 2 else 1.0) * t1 * t2 * ca * cb))
    return tuple(out)


def packed_size(n):
    return sum((n - p - 1) * comb(n - p, 2) for p in range(n - 1))


class SignedIntegrals:
    """g(a, b, c, d) of spinor labels (n, kappa, 2m) over the radial
    orbitals ``orbitals[(n, kappa)] = (P, Q)``: direct integrals from
    two_body_angular and the memoised scalar cell R^k(ac, bd), antisymmetrised
    and Hermitian-projected as spinor_integrals.py:389-399."""

    def __init__(self, orbitals, bsp_data, banks=None):
        self.orbitals, self.bsp_data = orbitals, bsp_data
        self.banks = Banks(bsp_data) if banks is None else banks
        self.memo = {}

    def radial(self, k, a, c, b, d):
        key = (k, a, c, b, d)
        if key not in self.memo:
            o = self.orbitals
            self.memo[key] = compute_slater_rk_cell(*o[a][:2], *o[c][:2], *o[b][:2], *o[d][:2],
                                                    self.bsp_data, k, self.banks)
        return self.memo[key]

    def direct(self, a, b, c, d):
        return sum(v * self.radial(k, a[:2], c[:2], b[:2], d[:2])
                   for k, v in two_body_angular(a[1], a[2], b[1], b[2], c[1], c[2], d[1], d[2]))

    def g(self, a, b, c, d):
        if a[2] + b[2] != c[2] + d[2]:
            return 0.0
        x = self.direct(a, b, c, d) - self.direct(a, b, d, c)
        y = self.direct(c, d, a, b) - self.direct(c, d, b, a)
        return 0.5 * (x + y)


def spinor_integrals(orbitals, bsp_data, Z, ns, occupied_keys, virtual_keys, frozen_keys=(),
                     nuclear_potential=None):
    """(labels, noccupied, heff, fock, packed, frozen_energy, reference_energy)."""
    from atomx_amd.dirac import dirac_matrices

    occ, vir, frozen = (tuple(tuple(k) for k in ks) for ks in
                        (occupied_keys, virtual_keys, frozen_keys))
    keys = occ + vir + frozen

    def multiplets(sel):
        return tuple((n, k, m) for n, k in sel for m in range(-2 * abs(k) + 1, 2 * abs(k), 2))

    labels = multiplets(occ + vir)
    core = multiplets(occ + frozen)
    frozen_labels = multiplets(frozen)
    banks = Banks(bsp_data)
    one = {k: dirac_matrices(k, Z, bsp_data, nuclear_potential) for k in {x[1] for x in keys}}

    def vec(key):
        H0, S = one[key[1]]
        m = S.shape[0] // 2
        p, q = (np.asarray(x, dtype=np.float64).ravel() for x in orbitals[key][:2])
        v = np.zeros(2 * m)
        v[:min(p.size, m)] = p[:m]
        v[m:m + min(q.size, m)] = q[:m]
        return v

    vecs = {k: vec(k) for k in keys}
    g = SignedIntegrals(orbitals, bsp_data, banks).g

    def h(a, b):
        if a[1:] != b[1:]:
            return 0.0
        H0, S = one[a[1]]
        ket = vecs[b[:2]]
        return (vecs[a[:2]] @ H0 @ ket) / (ket @ S @ ket)

    frozen_energy = sum(h(a, a) for a in frozen_labels) + 0.5 * sum(
        g(a, b, a, b) for a in frozen_labels for b in frozen_labels)
    reference_energy = sum(h(a, a) for a in core) + 0.5 * sum(
        g(a, b, a, b) for a in core for b in core)
    n = len(labels)
    heff = np.empty((n, n))
    fock = np.empty((n, n))
    for p, a in enumerate(labels):
        for q, b in enumerate(labels):
            heff[p, q] = h(a, b) + sum(g(a, i, b, i) for i in frozen_labels)
            fock[p, q] = h(a, b) + sum(g(a, i, b, i) for i in core)
    heff, fock = 0.5 * (heff + heff.T), 0.5 * (fock + fock.T)
    packed = np.zeros(packed_size(n))
    off = 0
    for p in range(n - 1):
        for q in range(p + 1, n):
            m = n - p
            for r in range(p, n):
                for s in range(r + 1, n):
                    rx, sx = r - p, s - p
                    packed[off + m * rx + sx - (rx + 1) * (rx + 2) // 2] = g(
                        labels[p], labels[q], labels[r], labels[s])
            off += comb(n - p, 2)
    return labels, len(multiplets(occ)), heff, fock, packed, frozen_energy, reference_energy
