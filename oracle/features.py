"""CPU restatement of the auxiliary-node Coulomb action and the pair
contraction — TEST INFRASTRUCTURE ONLY (imported by tests/ only, as the
checker).

``apply`` restates numerics/bspline/coulomb_features.py:30-61 (_apply) as
plain loops in the reference's order; ``pair_contraction`` evaluates
pair_contraction.py's M[m,n] = sum_{r,s} R^k(mr,sn) rho[r,s] in operator
form, R^k(ab;cd) = (a*b) . K (c*d) with K applied by ``apply`` column by
column — the definition, not the moment/edge factorisation the reference
(and the GPU path) uses.

Pinned by tests/test_features_cpu.py: with the product's geometry,
rho_ab . apply(rho_cd) equals this package's scalar cell R^k
(oracle.compute_slater_rk_cell, itself pinned by the DBSR_HF neon fixture)
for every P/Q block, kinetic balance on and off and ranks 0-7, the origin
cell monomial double integral R^5/5 * 2/(3+k) holds, and the action is
symmetric.  ``pair_contraction`` is pinned by the brute-force sum over
oracle.compute_slater_rk_cell at small sizes.
"""

from __future__ import annotations

import numpy as np


def apply(rho, outer, rising, falling, left, right, weights):
    """K rho: completed-cell moments through prefix/suffix sums, then the
    symmetric same-cell edges (coulomb_features.py:30-61)."""
    rho = np.asarray(rho)
    out = np.zeros_like(rho)
    _, nv, ng = outer.shape
    rise = np.zeros(nv, dtype=rho.dtype)
    fall = np.zeros(nv, dtype=rho.dtype)
    for c in range(nv):
        for comp in range(2):
            for g in range(ng):
                i = outer[comp, c, g]
                rise[c] += rising[c, g] * rho[i]
                fall[c] += falling[c, g] * rho[i]
    pre = rho.dtype.type(0)
    for c in range(nv):
        for comp in range(2):
            for g in range(ng):
                out[outer[comp, c, g]] += falling[c, g] * pre
        pre += rise[c]
    suf = rho.dtype.type(0)
    for c in range(nv - 1, -1, -1):
        for comp in range(2):
            for g in range(ng):
                out[outer[comp, c, g]] += rising[c, g] * suf
        suf += fall[c]
    for e in range(len(left)):
        i, j, w = int(left[e]), int(right[e]), weights[e]
        out[i] += w * rho[j]
        out[j] += w * rho[i]
    return out


def apply_features(features, rho, rank):
    """``apply`` with the geometry arrays of a features object."""
    rising, falling, w = features.rank_arrays(rank)
    return apply(rho, np.asarray(features.outer_indices), rising, falling,
                 np.asarray(features.edge_left), np.asarray(features.edge_right), w)


def pair_contraction(features, A, B, C, D, rank, rho):
    """M[m,n] = sum_{r,s} rho[r,s] (A_m B_r) . K (C_s D_n)."""
    A, B, C, D = (np.asarray(x) for x in (A, B, C, D))
    dtype = np.result_type(A, B, C, D, rho, float)
    m, r = A.shape[1], B.shape[1]
    s, n = C.shape[1], D.shape[1]
    out = np.zeros((m, n), dtype=dtype)
    for si in range(s):
        for ni in range(n):
            v = apply_features(features, (C[:, si] * D[:, ni]).astype(dtype), rank)
            for mi in range(m):
                for ri in range(r):
                    out[mi, ni] += rho[ri, si] * np.dot(A[:, mi] * B[:, ri], v)
    return out
