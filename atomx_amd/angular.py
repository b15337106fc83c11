"""Angular coefficients of the average-level Dirac-Fock functional.

Integer selection rules and Wigner 3j values turned into the FP64 weights the
Fock build and the two-electron energy multiply R^k contractions with
(angular/coefficients.py:153-210; methods/dirac_fock.py:714-760,
1070-1126, 1321-1387).  Host-side table setup, computed once per solve.
"""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

__all__ = ["wigner_3j", "wigner_6j", "exchange_lambda", "exchange_k_range", "l_of_kappa",
           "twoj_of_kappa", "reduced_c_jj", "AngularTables"]


def l_of_kappa(kappa: int) -> int:
    return kappa if kappa > 0 else -kappa - 1


def twoj_of_kappa(kappa: int) -> int:
    return 2 * abs(kappa) - 1


@lru_cache(maxsize=None)
def _lf(n: int) -> float:
    return math.lgamma(n + 1)


@lru_cache(maxsize=None)
def wigner_3j(tj1, tj2, tj3, tm1, tm2, tm3) -> float:
    """Wigner 3j symbol, all arguments twice the angular momenta (Racah sum in
    log-factorials)."""
    if tm1 + tm2 + tm3 != 0:
        return 0.0
    if abs(tm1) > tj1 or abs(tm2) > tj2 or abs(tm3) > tj3:
        return 0.0
    if (tj1 + tm1) % 2 or (tj2 + tm2) % 2 or (tj3 + tm3) % 2:
        return 0.0
    if tj3 < abs(tj1 - tj2) or tj3 > tj1 + tj2 or (tj1 + tj2 + tj3) % 2:
        return 0.0
    a = (tj1 + tj2 - tj3) // 2
    b = (tj1 - tj2 + tj3) // 2
    c = (-tj1 + tj2 + tj3) // 2
    d = (tj1 + tj2 + tj3) // 2 + 1
    log_pre = 0.5 * (_lf(a) + _lf(b) + _lf(c) - _lf(d)
                     + _lf((tj1 + tm1) // 2) + _lf((tj1 - tm1) // 2)
                     + _lf((tj2 + tm2) // 2) + _lf((tj2 - tm2) // 2)
                     + _lf((tj3 + tm3) // 2) + _lf((tj3 - tm3) // 2))
    total = 0.0
    for t in range(0, a + b + c + 1):
        args = (t, a - t, (tj1 - tm1) // 2 - t, (tj2 + tm2) // 2 - t,
                (tj3 - tj2 + tm1) // 2 + t, (tj3 - tj1 - tm2) // 2 + t)
        if min(args) < 0:
            continue
        term = math.exp(log_pre - sum(_lf(x) for x in args))
        total += -term if t % 2 else term
    phase = -1.0 if ((tj1 - tj2 - tm3) // 2) % 2 else 1.0
    return phase * total


def _closes(ta, tb, tc):
    return (ta + tb + tc) % 2 == 0 and abs(ta - tb) <= tc <= ta + tb


@lru_cache(maxsize=None)
def wigner_6j(tj1, tj2, tj3, tj4, tj5, tj6) -> float:
    """Wigner 6j symbol {j1 j2 j3; j4 j5 j6}, all arguments twice the angular
    momenta: Racah's single sum over t of (-1)^t (t+1)! over the factorials
    of (t - triad sums) and (tetrad sums - t), times the four triangle
    coefficients of (j1 j2 j3), (j1 j5 j6), (j4 j2 j6), (j4 j5 j3), in
    log-factorials.  Zero unless the four triads close."""
    triads = ((tj1, tj2, tj3), (tj1, tj5, tj6), (tj4, tj2, tj6), (tj4, tj5, tj3))
    if not all(_closes(*t) for t in triads):
        return 0.0
    log_tri = 0.5 * sum(_lf((a + b - c) // 2) + _lf((a - b + c) // 2) + _lf((b + c - a) // 2)
                        - _lf((a + b + c) // 2 + 1) for a, b, c in triads)
    lo = [sum(t) // 2 for t in triads]
    hi = ((tj1 + tj2 + tj4 + tj5) // 2, (tj2 + tj3 + tj5 + tj6) // 2,
          (tj3 + tj1 + tj6 + tj4) // 2)
    total = 0.0
    for t in range(max(lo), min(hi) + 1):
        lg = _lf(t + 1) - sum(_lf(t - x) for x in lo) - sum(_lf(x - t) for x in hi)
        term = math.exp(log_tri + lg)
        total += -term if t % 2 else term
    return total


@lru_cache(maxsize=None)
def reduced_c_jj(kappa_a, kappa_b, k) -> float:
    """<kappa_a || C^k || kappa_b> in Grant's phase convention
    (angular/coefficients.py:29-65): (-1)^(j_a + 1/2) sqrt((2j_a+1)(2j_b+1))
    (j_a k j_b; -1/2 0 1/2), zero unless l_a + k + l_b is even and
    |j_a - j_b| <= k <= j_a + j_b.  At k = 0 it is -sqrt(2j + 1)."""
    la, lb = l_of_kappa(kappa_a), l_of_kappa(kappa_b)
    if (la + k + lb) % 2:
        return 0.0
    ta, tb = twoj_of_kappa(kappa_a), twoj_of_kappa(kappa_b)
    if not abs(ta - tb) <= 2 * k <= ta + tb:
        return 0.0
    w = wigner_3j(ta, 2 * k, tb, -1, 0, 1)
    if abs(w) < 1e-15:
        return 0.0
    sign = -1.0 if ((ta + 1) // 2) % 2 else 1.0
    return sign * math.sqrt((ta + 1) * (tb + 1)) * w


def exchange_lambda(kappa_a, kappa_b, k) -> float:
    """(2ja+1)(2jb+1) (ja k jb; -1/2 0 1/2)^2 (angular/coefficients.py:153-183)."""
    ta, tb = twoj_of_kappa(kappa_a), twoj_of_kappa(kappa_b)
    w = wigner_3j(ta, 2 * k, tb, -1, 0, 1)
    if abs(w) < 1e-15:
        return 0.0
    return (ta + 1) * (tb + 1) * w * w


def exchange_k_range(kappa_a, kappa_b):
    """Allowed exchange ranks: triangle in j and parity in l (coefficients.py:186-210)."""
    la, lb = l_of_kappa(kappa_a), l_of_kappa(kappa_b)
    ta, tb = twoj_of_kappa(kappa_a), twoj_of_kappa(kappa_b)
    return [k for k in range(abs(ta - tb) // 2, (ta + tb) // 2 + 1) if (k + la + lb) % 2 == 0]


class AngularTables:
    """Direct a(i, j, k) and exchange b(i, j, k) coefficients of target i's Fock
    operator, the divisors q_i, and the energy weights WW(i, j).

    ``mean_field_coeffs`` / ``mean_field_normalization`` reproduce the
    reference's optional asymmetric mean-field seam (dirac_fock.py:735-760).
    """

    def __init__(self, kappas, occupations, mean_field_coeffs=None,
                 mean_field_normalization=None, ww=None):
        self.kappas = [int(x) for x in kappas]
        self.occ = [float(x) for x in occupations]
        self.n = len(self.kappas)
        self._ww = None if ww is None else np.asarray(ww, dtype=float)
        self._mfc = None if mean_field_coeffs is None else np.asarray(mean_field_coeffs, float)
        self._mfn = (None if mean_field_normalization is None
                     else np.asarray(mean_field_normalization, float))
        self.kmax = max(abs(a) + abs(b) for a in self.kappas for b in self.kappas)
        n, K = self.n, self.kmax + 1
        self.direct = np.zeros((n, n, K))
        self.exchange = np.zeros((n, n, K))
        self.k_for_target = []
        for i in range(n):
            ks = set()
            for j in range(n):
                for k in range(K):
                    dc = self._direct_coef(i, j, k)
                    ec = self._exchange_coef(i, j, k)
                    if abs(dc) > 1e-15 or abs(ec) > 1e-15:
                        self.direct[i, j, k] = dc
                        self.exchange[i, j, k] = ec
                        ks.add(k)
            self.k_for_target.append(tuple(sorted(ks)))

    def ww(self, i, j):
        """Occupation-product weight WW(i, j) (dirac_fock.py:714-733)."""
        if self._ww is not None:
            return float(self._ww[i, j])
        qi = self.occ[i]
        return qi * (qi - 1.0) if i == j else qi * self.occ[j]

    def mean_field_weight(self, i, j):
        return float(self._mfc[i, j]) if self._mfc is not None else self.ww(i, j)

    def divisor(self, i):
        return float(self._mfn[i]) if self._mfn is not None else self.occ[i]

    def _direct_coef(self, i, j, k):
        """dirac_fock.py:1321-1355."""
        ki, kj = self.kappas[i], self.kappas[j]
        tji, tjj = twoj_of_kappa(ki), twoj_of_kappa(kj)
        if k % 2 or k > (tji + tjj) // 2:
            return 0.0
        if i == j:
            c = self.mean_field_weight(i, j) / 2.0
            if k == 0:
                return c
            lam = exchange_lambda(ki, ki, k)
            if abs(lam) < 1e-15:
                return 0.0
            return -c * lam / (tji * (tji + 1.0))
        if k != 0:
            return 0.0
        return self.mean_field_weight(i, j)

    def _exchange_coef(self, i, j, k):
        """dirac_fock.py:1357-1387."""
        if i == j:
            return 0.0
        ki, kj = self.kappas[i], self.kappas[j]
        li, lj = l_of_kappa(ki), l_of_kappa(kj)
        tji, tjj = twoj_of_kappa(ki), twoj_of_kappa(kj)
        if (k + li + lj) % 2 or k < abs(tji - tjj) // 2 or k > (tji + tjj) // 2:
            return 0.0
        w = self.mean_field_weight(i, j)
        if abs(w) < 1e-15:
            return 0.0
        lam = exchange_lambda(ki, kj, k)
        if abs(lam) < 1e-15:
            return 0.0
        return -w * lam / ((tji + 1.0) * (tjj + 1.0))

    def energy_terms(self):
        """The F^k / G^k terms of E_2e = E_J - E_K (dirac_fock.py:1618-1661):
        list of (weight, kind, i, j, k, is_direct) with kind 'F' = R^k(ii;jj)
        and 'G' = R^k(ij;ji)."""
        out = []
        for i in range(self.n):
            ka = self.kappas[i]
            tja, la = twoj_of_kappa(ka), l_of_kappa(ka)
            for j in range(i + 1):
                kb = self.kappas[j]
                tjb = twoj_of_kappa(kb)
                if i == j:
                    c = self.ww(i, i) / 2.0
                    if abs(c) < 1e-15:
                        continue
                    out.append((c, "F", i, i, 0, True))
                    for k in range(1, tja + 1):
                        if (k + 2 * la) % 2:
                            continue
                        lam = exchange_lambda(ka, ka, k)
                        if lam < 1e-15:
                            continue
                        out.append((c * lam / (tja * (tja + 1.0)), "F", i, i, k, False))
                else:
                    w = self.ww(i, j)
                    if abs(w) < 1e-15:
                        continue
                    out.append((w, "F", i, j, 0, True))
                    for k in exchange_k_range(ka, kb):
                        lam = exchange_lambda(ka, kb, k)
                        if lam < 1e-15:
                            continue
                        out.append((w * lam / ((tja + 1.0) * (tjb + 1.0)), "G", i, j, k, False))
        return out
