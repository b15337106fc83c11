"""Pulay DIIS on flattened orbital vectors, as the reference's SCF applies it
(numerics/diis.py:24-88 for the extrapolation, methods/dirac_fock.py:
2040-2202 for when it is used).

Given iterates x_i = F(y_i) and residuals r_i = x_i - y_i, the coefficients
c minimise |sum_i c_i r_i| under sum_i c_i = 1 (a bordered Gram system,
ridge-regularised by 1e-12 of its mean diagonal); the extrapolated vector
is sum_i c_i x_i.  The solve is rejected (None) when it is singular or any
|c_i| exceeds 5, as the reference's is.  For an affine map F the
extrapolation lands on the fixed point once the history spans it.
"""

from __future__ import annotations

import numpy as np

COEFFICIENT_LIMIT = 5.0
RIDGE_FRACTION = 1e-12

__all__ = ["DiisHistory", "extrapolate", "lowdin_by_kappa"]


def extrapolate(iterates, residuals):
    """sum_i c_i iterates[i] with the DIIS coefficients, or None."""
    m = len(residuals)
    if m < 2:
        return None
    R = np.asarray(residuals, dtype=np.float64)
    A = np.zeros((m + 1, m + 1))
    A[:m, :m] = R @ R.T
    scale = float(np.mean(np.abs(np.diag(A[:m, :m]))))
    if scale > 0.0:
        A[np.arange(m), np.arange(m)] += RIDGE_FRACTION * scale
    A[m, :m] = A[:m, m] = -1.0
    rhs = np.zeros(m + 1)
    rhs[m] = -1.0
    try:
        c = np.linalg.solve(A, rhs)[:m]
    except np.linalg.LinAlgError:
        return None
    if not np.all(np.isfinite(c)) or np.any(np.abs(c) > COEFFICIENT_LIMIT):
        return None
    return c @ np.asarray(iterates, dtype=np.float64)


class DiisHistory:
    """The last ``depth`` (iterate, residual) pairs."""

    def __init__(self, depth=8):
        self.depth = depth
        self.x, self.r = [], []

    def __len__(self):
        return len(self.r)

    def push(self, x, r):
        self.x.append(np.array(x, dtype=np.float64))
        self.r.append(np.array(r, dtype=np.float64))
        del self.x[:-self.depth], self.r[:-self.depth]

    def clear(self):
        self.x.clear()
        self.r.clear()

    def extrapolate(self):
        return extrapolate(self.x, self.r)


def lowdin_by_kappa(P, Q, kappas, overlaps):
    """Symmetric (Lowdin) orthonormalisation of the orbitals of each kappa in
    the metric of that kappa's overlap matrix (2 ns x 2 ns over [P; Q]), as
    the reference retracts an extrapolated set (methods/dirac_fock.py:
    1691-1704).  Raises ValueError when a kappa's orbitals are linearly
    dependent (smallest Gram eigenvalue below 1e-12)."""
    P, Q = np.array(P, dtype=np.float64), np.array(Q, dtype=np.float64)
    ns = P.shape[1]
    for kappa in sorted(set(kappas)):
        rows = [i for i, k in enumerate(kappas) if k == kappa]
        V = np.concatenate((P[rows], Q[rows]), axis=1).T      # (2 ns, m)
        G = V.T @ overlaps[kappa] @ V
        w, U = np.linalg.eigh(0.5 * (G + G.T))
        if w.min() < 1e-12:
            raise ValueError(f"orbitals of kappa {kappa} are linearly dependent")
        V = V @ ((U / np.sqrt(w)) @ U.T)
        P[rows], Q[rows] = V[:ns].T, V[ns:].T
    return P, Q
