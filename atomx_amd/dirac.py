"""Dirac-Fock SCF around the GPU Fock build (host side, FP64).

The one-body Dirac matrices, the nuclear potential and the generalised
eigensolve are host work in the reference too (numerics/bspline/
dirac_assembly.py, potentials/nuclear.py); the eigensolve stays FP64 on the
CPU.  Each SCF sweep updates the orbitals one at a time in subshell order
(the reference's Gauss-Seidel order, methods/dirac_fock.py:2057-2168): the
Fock matrix of orbital i is the one-body matrix of its kappa plus the
two-electron blocks of ``_build_hf_matrix_dbsr`` built on the GPU from the
current orbitals (``FockBuilder``), then the reduced generalised eigenproblem
is solved under orthogonality to the other orbitals of the same kappa and
the root of the subshell's radial node count is taken.

The two-electron backend is pluggable (``fock(i, P, Q)`` and
``energy(P, Q)``) so that the same driver can run on another
implementation of those two functions.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import linalg

from .workloads import SPEED_OF_LIGHT

__all__ = ["fermi_params", "nuclear_potential", "assemble_integral", "dirac_matrices",
           "solve_dirac", "Subshell", "neon_subshells", "DiracFockSCF", "BOHR_RADIUS_FM"]

C = SPEED_OF_LIGHT
C_SQ = C * C

_THREADPOOLS = None


def one_blas_thread():
    """Context manager: BLAS/LAPACK on one thread.  The SCF's dense algebra
    is on matrices of a few hundred rows, where a multi-threaded BLAS spends
    more time waking threads than computing (a 120-row neon eigensolve
    takes ~15x longer on 8 threads than on one); systems solved in parallel
    get their parallelism from host threads instead."""
    global _THREADPOOLS
    import contextlib
    try:
        from threadpoolctl import ThreadpoolController
    except ImportError:  # pragma: no cover - threadpoolctl ships with scikit-learn
        return contextlib.nullcontext()
    if _THREADPOOLS is None:
        _THREADPOOLS = ThreadpoolController()
    return _THREADPOOLS.limit(limits=1, user_api="blas")
BOHR_RADIUS_FM = 5.29177210544e-11 * 1e15   # CODATA 2022 Bohr radius in fm

# DBSR_HF RMS charge radii (fm) and mass numbers of the light elements
# (potentials/nuclear.py:17-34)
_RRMS_FM = {1: 0.8783, 2: 1.6755, 3: 2.5890, 4: 2.5190, 5: 2.4277, 6: 2.4702, 7: 2.5582,
            8: 2.6991, 9: 2.8976, 10: 3.0055}
_MASS = {1: 1, 2: 4, 3: 7, 4: 9, 5: 11, 6: 12, 7: 14, 8: 16, 9: 19, 10: 20}


# ---------------------------------------------------------------------------
# Nuclear potential (potentials/nuclear.py:90-290)
# ---------------------------------------------------------------------------
def _mass(Z, A):
    return A if A is not None else _MASS.get(Z, int(round(2.5 * max(Z, 1))))


def _rrms_fm(Z, A):
    if Z in _RRMS_FM:
        return _RRMS_FM[Z]
    return 0.836 * _mass(Z, A) ** (1.0 / 3.0) + 0.570


def _sk(k, x):
    """S_k(x) = sum_{n>=1} (-1)^n e^{n x} / n^k (nuclear.py:160-181)."""
    if x < -500.0:
        return 0.0
    eps = 10.0 * np.finfo(float).eps
    base = -np.exp(x)
    term = base
    out = base
    n = 1
    while True:
        n += 1
        term *= base
        delta = term / float(n) ** k
        out += delta
        if abs(delta) <= eps * abs(out) or n > 1000:
            break
    return out


def _rms_of(a, c):
    """RMS radius of the Fermi density (nuclear.py:184-196)."""
    A = a / c
    P = np.pi * A
    X = -c / a
    dn = 1.0 + 10.0 / 3.0 * P ** 2 + 7.0 / 3.0 * P ** 4 - 120.0 * A ** 5 * _sk(5, X)
    dd = 1.0 + P ** 2 - 6.0 * A ** 3 * _sk(3, X)
    return c * np.sqrt(3.0 * dn / 5.0 / dd)


def _c_from_rms(rrms, a):
    """Half-density radius for a given RMS radius by bisection (nuclear.py:199-226)."""
    accy = 2.0 * np.finfo(float).eps
    lo = rrms
    while _rms_of(a, lo) > rrms:
        lo *= 0.5
    hi = rrms
    while _rms_of(a, hi) < rrms:
        hi *= 2.0
    mid = 0.5 * (hi + lo)
    for _ in range(200):
        mid = 0.5 * (hi + lo)
        r = _rms_of(a, mid)
        if r > rrms:
            hi = mid
        else:
            lo = mid
        if (hi - lo) / (hi + lo) <= accy or abs(r - rrms) / rrms <= accy:
            break
    return mid


def fermi_params(Z, A=None):
    """(a, c) of the Fermi charge distribution in bohr (nuclear.py:90-131):
    skin thickness t = 2.30 fm, a = t / (4 ln 3), RMS radius clamped to >= 2 fm."""
    a_fm = 2.30 / (4.0 * np.log(3.0))
    rrms = max(_rrms_fm(Z, _mass(Z, A)), 2.0)
    c_fm = _c_from_rms(rrms, a_fm)
    return a_fm / BOHR_RADIUS_FM, c_fm / BOHR_RADIUS_FM


def nuclear_potential(Z, r, model="point", A=None):
    """V(r) on an array of radii: -Z/r, or the closed-form potential of the
    Fermi distribution (nuclear.py:254-290, Parpia & Mohanty 1992)."""
    r = np.asarray(r, dtype=float)
    if model == "point":
        return -Z / r
    if model != "fermi":
        raise ValueError(f"Unknown nuclear model: {model!r}")
    a, c = fermi_params(Z, A)
    ac = a / c
    ac2, ac3 = ac * ac, ac * ac * ac
    p = np.pi ** 2 * ac2
    s = 6.0 * ac3 * _sk(3, -c / a)
    N = 1.0 + p - s
    out = np.empty_like(r)
    for idx, x in np.ndenumerate(r):
        if x <= 0.0:
            out[idx] = -Z * (3.0 + p) / (2.0 * N * c)
            continue
        rc = x / c
        if x <= c:
            ra = (x - c) / a
            s2, s3 = ac2 * _sk(2, ra), ac3 * _sk(3, ra)
            v = (3.0 - rc * rc + p) / 2.0 - 3.0 * s2 - s / rc + 6.0 / rc * s3
            out[idx] = -Z * v / (N * c)
        else:
            ra = (c - x) / a
            s2, s3 = ac2 * _sk(2, ra), ac3 * _sk(3, ra)
            v = N + 6.0 * s3 + 3.0 * rc * s2
            out[idx] = -Z / x * (v / N)
    return out


# ---------------------------------------------------------------------------
# One-body Dirac matrices (numerics/bspline/integrals.py, dirac_assembly.py)
# ---------------------------------------------------------------------------
def assemble_integral(nv, ks1, ks2, mm, bsp1, bsp2, weights):
    """<B1_i | w | B2_j> on an (mm, mm) matrix, cell by cell
    (numerics/bspline/integrals.py:18-56)."""
    mat = np.zeros((mm, mm))
    for iv in range(nv):
        blk = np.einsum("m,mi,mj->ij", weights[iv], bsp1[iv], bsp2[iv])
        i1 = min(ks1, mm - iv)
        j1 = min(ks2, mm - iv)
        if i1 > 0 and j1 > 0:
            mat[iv:iv + i1, iv:iv + j1] += blk[:i1, :j1]
    return mat


def dirac_matrices(kappa, Z, bd, nuclear="point", A=None):
    """(H, S) of the one-electron Dirac operator in the P/Q B-spline basis,
    (2 ns, 2 ns) with ns = max(nsp, nsq) (dirac_assembly.py:82-155): nuclear
    attraction on both diagonal blocks, -2c^2 S on the QQ block, the
    derivative and 1/r couplings, and under kinetic balance the Bloch
    boundary terms at the last knot.  ``nuclear`` is "point" (or None),
    "fermi", or a callable V(r)."""
    nv, ks = bd["nv"], bd["ks"]
    gr, grw = bd["gr"], bd["grw"]
    ksp, ksq = bd.get("ksp", ks), bd.get("ksq", ks)
    nsp, nsq = bd.get("nsp", bd["ns"]), bd.get("nsq", bd["ns"])
    bp, bq, bqd = bd["bsp_p"], bd["bsp_q"], bd["bsp_q_d"]
    bpd = bd.get("bsp_p_d")
    ns = max(nsp, nsq)
    if nuclear is None:
        nuclear = "point"
    if callable(nuclear):  # V(r) evaluated node by node (integrals.py:103-117)
        vw = grw * np.vectorize(lambda x: float(nuclear(x)))(gr)
    else:
        vw = grw * nuclear_potential(Z, gr, nuclear, A)
    sym = lambda M: 0.5 * (M + M.T)  # noqa: E731
    S_PP = sym(assemble_integral(nv, ksp, ksp, ns, bp, bp, grw))
    S_QQ = sym(assemble_integral(nv, ksq, ksq, ns, bq, bq, grw))
    V_PP = sym(assemble_integral(nv, ksp, ksp, ns, bp, bp, vw))
    V_QQ = sym(assemble_integral(nv, ksq, ksq, ns, bq, bq, vw))
    wr = grw / gr
    fpqbsd = assemble_integral(nv, ksp, ksq, ns, bp, bqd, grw)
    fpqbs = assemble_integral(nv, ksp, ksq, ns, bp, bq, wr)
    if bpd is not None and ksp != ksq:
        fqpbsd = assemble_integral(nv, ksq, ksp, ns, bq, bpd, grw)
        fqpbs = assemble_integral(nv, ksq, ksp, ns, bq, bp, wr)
    else:
        fqpbsd, fqpbs = fpqbsd, fpqbs
    H = np.zeros((2 * ns, 2 * ns))
    S = np.zeros((2 * ns, 2 * ns))
    H[:ns, :ns] = V_PP
    H[ns:, ns:] = V_QQ - 2.0 * C_SQ * S_QQ
    H[:ns, ns:] = -C * fpqbsd + kappa * C * fpqbs
    H[ns:, :ns] = C * fqpbsd + kappa * C * fqpbs
    S[:ns, :ns] = S_PP
    S[ns:, ns:] = S_QQ
    if ksp != ksq:
        tmax = bd["t"][nsq]
        if tmax > 0:
            rn = kappa / (2.0 * tmax * C)
            ip, iq = nsp - 1, ns + nsq - 1
            H[ip, ip] -= rn * C
            H[ip, iq] += C
    return H, S


def solve_dirac(H, S, ns, nsp, nsq, orthogonal_to=None, nb=1, ibzero=2):
    """Reduced generalised eigenproblem (dirac_assembly.py:455-575): the first
    ``nb`` and last ``ibzero`` splines of each component removed, optional
    S-orthogonality to the columns of ``orthogonal_to`` (full layout), the
    lower triangle of H used, LAPACK divide and conquer.  Returns the
    physical roots (above -c^2) in ascending order and their eigenvectors
    embedded in the full (2 ns) layout."""
    keep = np.concatenate([np.arange(nb, nsp - ibzero), np.arange(ns + nb, ns + nsq - ibzero)])
    Hr = H[np.ix_(keep, keep)]
    Sr = S[np.ix_(keep, keep)]
    if orthogonal_to is not None and np.size(orthogonal_to):
        cons = np.asarray(orthogonal_to)[keep, :]
        basis, _ = linalg.qr(Sr @ cons, mode="full", check_finite=False)
        comp = basis[:, cons.shape[1]:]
        hs = np.tril(Hr) + np.tril(Hr, -1).T
        w, v = linalg.eigh(comp.T @ hs @ comp, comp.T @ Sr @ comp, driver="gvd",
                           check_finite=False)
        v = comp @ v
    else:
        w, v = linalg.eigh(Hr, Sr, driver="gvd", check_finite=False)
    phys = w > -C_SQ
    full = np.zeros((2 * ns, int(phys.sum())))
    full[keep, :] = v[:, phys]
    return w[phys], full


# ---------------------------------------------------------------------------
# SCF
# ---------------------------------------------------------------------------
@dataclass
class Subshell:
    label: str
    n: int
    kappa: int
    occupation: float

    @property
    def l(self):  # noqa: E743
        return self.kappa if self.kappa > 0 else -self.kappa - 1


def neon_subshells():
    """1s2 2s2 2p-2 2p4 (methods/dirac_fock.py neon_subshells)."""
    return [Subshell("1s1/2", 1, -1, 2.0), Subshell("2s1/2", 2, -1, 2.0),
            Subshell("2p1/2", 2, 1, 2.0), Subshell("2p3/2", 2, -2, 4.0)]


def _phase(P, Q):
    i = int(np.argmax(np.abs(P)))
    return (-P, -Q) if P[i] < 0.0 else (P, Q)


class DiracFockSCF:
    """Average-level Dirac-Fock SCF on a B-spline basis.

    backend: None for the GPU (``FockBuilder``), or an object with
    ``fock(i, P, Q) -> (PP, QQ, PQ, QP)`` and ``energy(P, Q) -> (E2e, EJ, EK)``.
    """

    def __init__(self, Z, subshells, bsp_data, nuclear="point", A=None, backend=None,
                 mode="dbsr"):
        self.Z = Z
        self.shells = list(subshells)
        self.bd = bsp_data
        self.nuclear, self.A = nuclear, A
        bd = bsp_data
        self.nsp, self.nsq = bd.get("nsp", bd["ns"]), bd.get("nsq", bd["ns"])
        self.ns = max(self.nsp, self.nsq)
        self.kb = bd.get("ksp", bd["ks"]) != bd.get("ksq", bd["ks"])
        kappas = [s.kappa for s in self.shells]
        occ = [s.occupation for s in self.shells]
        if backend is None:
            from .dirac_fock import FockBuilder
            fb = FockBuilder(bsp_data, kappas, occ)
            if mode == "dbsr":
                backend = _EngineBackend(fb)
            elif mode == "orbital":
                from .orbital_mode import OrbitalModeFock
                backend = _OrbitalBackend(OrbitalModeFock(bsp_data, kappas, occ, fb.cache), fb)
            else:
                raise ValueError(f"mode must be 'dbsr' or 'orbital', got {mode!r}")
        self.backend = backend
        self.h0 = {k: dirac_matrices(k, Z, bd, nuclear, A) for k in sorted(set(kappas))}
        n = len(self.shells)
        self.P = np.zeros((n, self.ns))
        self.Q = np.zeros((n, self.ns))
        self.eps = np.zeros(n)
        self.iterations = 0

    # -- orbitals ----------------------------------------------------------------
    def _root(self, i, energies, vectors, constrained):
        sh = self.shells[i]
        idx = sh.n - sh.l - 1
        if sh.kappa > 0 and not self.kb:
            idx += 1
        if constrained:
            idx -= sum(1 for j, o in enumerate(self.shells)
                       if j != i and o.kappa == sh.kappa and o.n < sh.n)
            idx = max(0, idx)
        bound = np.flatnonzero(energies < 0)
        order = bound[np.argsort(energies[bound])]
        if idx >= order.size:
            raise RuntimeError(f"no bound state for {sh.label}")
        j = order[idx]
        v = vectors[:, j]
        P = np.zeros(self.ns)
        Q = np.zeros(self.ns)
        P[:self.nsp] = v[:self.nsp]
        Q[:self.nsq] = v[self.ns:self.ns + self.nsq]
        P, Q = _phase(P, Q)
        return P, Q, float(energies[j])

    def bare_nucleus_start(self):
        """Orbitals of the bare nucleus (one-body Dirac operator only)."""
        with one_blas_thread():
            for i, sh in enumerate(self.shells):
                H, S = self.h0[sh.kappa]
                w, v = solve_dirac(H, S, self.ns, self.nsp, self.nsq)
                self.P[i], self.Q[i], self.eps[i] = self._root(i, w, v, False)
        return self

    def virtual_orbitals(self, kappa, count):
        """The ``count`` lowest physical solutions of the one-body Dirac
        operator of ``kappa`` that are orthogonal, in that kappa's metric, to
        this SCF's orbitals of the same kappa: an orthonormal extension of
        the occupied space (bare-nucleus virtuals) for correlation bases.
        Returns [(P, Q, energy)]."""
        H, S = self.h0[kappa] if kappa in self.h0 else dirac_matrices(
            kappa, self.Z, self.bd, self.nuclear, self.A)
        occ = [np.concatenate((self.P[j], self.Q[j])) for j, o in enumerate(self.shells)
               if o.kappa == kappa]
        with one_blas_thread():
            w, v = solve_dirac(H, S, self.ns, self.nsp, self.nsq,
                               orthogonal_to=np.column_stack(occ) if occ else None)
        out = []
        for j in range(min(count, v.shape[1])):
            P = np.zeros(self.ns)
            Q = np.zeros(self.ns)
            P[:self.nsp] = v[:self.nsp, j]
            Q[:self.nsq] = v[self.ns:self.ns + self.nsq, j]
            P, Q = _phase(P, Q)
            out.append((P, Q, float(w[j])))
        return out

    def set_orbitals(self, P, Q):
        self.P[:] = 0.0
        self.Q[:] = 0.0
        for i in range(len(self.shells)):
            m = min(np.shape(P)[1], self.ns)
            self.P[i, :m] = np.asarray(P)[i, :m]
            self.Q[i, :m] = np.asarray(Q)[i, :m]
        return self

    def fock_matrix(self, i):
        """H0 of the orbital's kappa plus its two-electron blocks (PQ block in
        the strictly upper part, which the eigensolve does not read)."""
        sh = self.shells[i]
        H, S = self.h0[sh.kappa]
        H = H.copy()
        pp, qq, pq, qp = self.backend.fock(i, self.P, self.Q)
        ns = self.ns
        H[:ns, :ns] += pp
        H[ns:, ns:] += qq
        H[:ns, ns:] += pq
        H[ns:, :ns] += qp
        return H, S

    def sweep(self, diis=False):
        """One Gauss-Seidel pass over the orbitals; returns the largest
        change of a large-component coefficient relative to its maximum.
        diis: Pulay extrapolation of the swept orbitals, as the reference
        applies it (methods/dirac_fock.py:2061-2202): only for five or more
        orbitals from the fourth sweep on, history of 8, extrapolated when 3
        or more entries are held and the sweep changed the orbitals by
        between 1e-3 and 0.3, then Lowdin-orthonormalised per kappa."""
        with one_blas_thread():
            return self._sweep(diis)

    def _sweep(self, diis):
        use_diis = diis and len(self.shells) >= 5 and self.iterations + 1 >= 4
        x_in = self._flat() if use_diis else None
        change = 0.0
        for i, sh in enumerate(self.shells):
            H, S = self.fock_matrix(i)
            others = [np.concatenate((self.P[j], self.Q[j])) for j, o in enumerate(self.shells)
                      if j != i and o.kappa == sh.kappa]
            cons = np.column_stack(others) if others else None
            w, v = solve_dirac(H, S, self.ns, self.nsp, self.nsq, orthogonal_to=cons)
            P, Q, e = self._root(i, w, v, cons is not None)
            pmax = np.max(np.abs(self.P[i]))
            if pmax > 1e-15:
                if np.dot(P, self.P[i]) < 0:
                    P, Q = -P, -Q
                change = max(change, float(np.max(np.abs(P - self.P[i])) / pmax))
            self.P[i], self.Q[i], self.eps[i] = P, Q, e
        self.iterations += 1
        if use_diis:
            self.diis_step(x_in, change)
        return change

    # -- DIIS ----------------------------------------------------------------------
    def _flat(self):
        """All orbitals as one vector: [P_0, Q_0, P_1, Q_1, ...]."""
        return np.concatenate([np.concatenate((p, q)) for p, q in zip(self.P, self.Q)])

    def diis_step(self, x_in, change):
        """Records the sweep (input x_in, output the current orbitals) and
        replaces the orbitals by the extrapolation when it applies."""
        from .diis import DiisHistory, lowdin_by_kappa
        if getattr(self, "_diis", None) is None:
            self._diis = DiisHistory(8)
        x_out = self._flat()
        self._diis.push(x_out, x_out - x_in)
        if len(self._diis) < 3 or not 1e-3 < change < 0.3:
            return False
        x = self._diis.extrapolate()
        if x is None:
            return False
        ns, n = self.ns, len(self.shells)
        v = x.reshape(n, 2, ns)
        try:
            P, Q = lowdin_by_kappa(v[:, 0], v[:, 1], [s.kappa for s in self.shells],
                                   {k: self.h0[k][1] for k in self.h0})
        except (ValueError, np.linalg.LinAlgError):
            self._diis.clear()  # outside the independent-orbital manifold: keep the sweep
            return False
        self.P[:], self.Q[:] = P, Q
        return True

    def run(self, max_iter=200, tol=1e-8, energy_tol=1e-12, diis=False):
        """Sweeps until the orbitals change by less than ``tol`` (relative to
        each large component's maximum) and the total energy by less than
        ``energy_tol`` Ha; returns the energy.  The energy is evaluated only
        on sweeps that changed the orbitals by less than 1e3 tol (an energy
        change is stationary, second order in the orbital change, so it can
        only decide convergence there)."""
        e_old = None
        for _ in range(max_iter):
            change = self.sweep(diis=diis)
            if change >= 1e3 * tol:
                e_old = None
                continue
            e = self.energy()[0]
            if e_old is not None and change < tol and abs(e - e_old) < energy_tol:
                return e
            e_old = e
        raise RuntimeError(f"SCF did not converge in {max_iter} sweeps (change {change:.2e})")

    def energy(self):
        """(E_total, E_1e, E_2e) of the average-level functional
        (methods/dirac_fock.py:1553-1661)."""
        e1 = 0.0
        ns = self.ns
        for i, sh in enumerate(self.shells):
            H, S = self.h0[sh.kappa]
            v = np.concatenate((self.P[i, :ns], self.Q[i, :ns]))
            nrm = v @ S @ v
            if nrm > 1e-20:
                e1 += sh.occupation * (v @ H @ v) / nrm
        e2 = self.backend.energy(self.P, self.Q)[0]
        return e1 + e2, e1, e2


class _EngineBackend:
    """Fock blocks from one Fock-engine plan per target over resident
    coefficient stacks: a call uploads only the orbitals that changed since
    the previous call (one row after a Gauss-Seidel step), runs the target's
    plan and brings its four blocks back in one copy."""

    def __init__(self, fb):
        from . import device as dv
        self.fb = fb
        self.P = dv.zeros((fb.n_orb, fb.ns))
        self.Q = dv.zeros((fb.n_orb, fb.ns))
        self._hP = np.full((fb.n_orb, fb.ns), np.nan)  # what the device holds
        self._hQ = np.full((fb.n_orb, fb.ns), np.nan)
        self._plans = {}
        self.uploads = 0

    def _sync(self, P, Q):
        from . import device as dv
        P, Q = np.asarray(P, dtype=np.float64), np.asarray(Q, dtype=np.float64)
        m = min(P.shape[1], self.fb.ns)
        for b in range(self.fb.n_orb):
            if np.array_equal(P[b, :m], self._hP[b, :m]) and \
                    np.array_equal(Q[b, :m], self._hQ[b, :m]):
                continue
            self._hP[b] = 0.0
            self._hQ[b] = 0.0
            self._hP[b, :m], self._hQ[b, :m] = P[b, :m], Q[b, :m]
            self.P[b] = dv.f64(self._hP[b])
            self.Q[b] = dv.f64(self._hQ[b])
            self.uploads += 1

    def plan(self, i):
        from .dirac_fock import FockEngine
        eng = self._plans.get(i)
        if eng is None:
            eng = FockEngine([self.fb], skip_pq_exchange=True, targets=[[i]], P=[self.P],
                             Q=[self.Q])
            self._plans[i] = eng
        return eng

    def fock(self, i, P, Q):
        from . import device as dv
        self._sync(P, Q)
        eng = self.plan(i)
        eng.run()
        return tuple(dv.host(eng.H[0][0]))

    def rk_equivalents(self):
        return sum(e.rk_equivalents for e in self._plans.values())

    def energy(self, P, Q):
        return self.fb.energy(P, Q)


class _OrbitalBackend:
    """Orbital-mode Fock blocks; orbitals changed since the last call are
    re-uploaded one by one so the other sources' exchange contractions stay
    cached (as the reference drops only the updated orbital's entries)."""

    def __init__(self, om, fb):
        self.om = om
        self.fb = fb
        self._P = self._Q = None

    def fock(self, i, P, Q):
        P, Q = np.asarray(P), np.asarray(Q)
        if self._P is None or self._P.shape != P.shape:
            self.om.load(P, Q)
        else:
            for b in range(P.shape[0]):
                if not (np.array_equal(P[b], self._P[b]) and np.array_equal(Q[b], self._Q[b])):
                    self.om.update(b, P[b], Q[b])
        self._P, self._Q = P.copy(), Q.copy()
        return self.om.blocks(i)

    def energy(self, P, Q):
        return self.fb.energy(P, Q)

