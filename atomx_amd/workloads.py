"""Synthetic radial workloads of the Slater-integral engine.

``synthetic_grid`` / ``synthetic_orbitals`` build the batched all-quartet
workload: an exponential knot grid of ``nv`` intervals with ``ng`` Gauss
points (r from r_min to r_max, the first interval [0, r_min]), and radial
orbitals

    P(r) = r^(l+1) exp(-Z_eff r / n) (1 + sum_{j=1..4} c_j r^j),  Q = P'/(2c),

l ~ U{0..4}, n ~ U{1..6} redrawn until n > l, Z_eff ~ U[1, 30],
c_j ~ N(0, 0.1^2), jointly normalised so that sum grw (P^2 + Q^2) = 1 in FP64.
"""

from __future__ import annotations

import numpy as np

from .bspline import make_exponential_knots, setup_gauss_points

SPEED_OF_LIGHT = 137.03599917759013  # 1/alpha, CODATA 2022


def synthetic_grid(nv=400, ng=8, r_min=1e-5, r_max=60.0, ks=8):
    """(gr, grw) of shape (nv, ng) on make_exponential_knots(nv+ks-1, ks, r_min, r_max)."""
    ns = nv + ks - 1
    t = make_exponential_knots(ns, ks, r_min, r_max)
    return setup_gauss_points(t, ns, ks, ng)


def draw_orbital_params(n_orb, seed):
    rng = np.random.default_rng(seed)
    params = []
    for _ in range(n_orb):
        while True:
            l = int(rng.integers(0, 5))
            n = int(rng.integers(1, 7))
            if n > l:
                break
        z = float(rng.uniform(1.0, 30.0))
        c = rng.normal(0.0, 0.1, size=4)
        params.append((l, n, z, c))
    return params


def orbital_samples(r, l, n, z, c):
    """P(r), Q(r) = P'(r)/(2c) of one synthetic orbital (unnormalised)."""
    beta = z / n
    j = np.arange(1, 5)
    poly = 1.0 + (c[None, :] * r[:, None] ** j[None, :]).sum(axis=1)
    dpoly = (c[None, :] * j[None, :] * r[:, None] ** (j[None, :] - 1)).sum(axis=1)
    e = np.exp(-beta * r)
    P = r ** (l + 1) * e * poly
    dP = ((l + 1) * r ** l - beta * r ** (l + 1)) * e * poly + r ** (l + 1) * e * dpoly
    return P, dP / (2.0 * SPEED_OF_LIGHT)


def synthetic_orbitals(gr, grw, n_orb=40, seed=0):
    """Returns (P, Q, l) with P, Q of shape (n_orb, nv*ng) normalised in FP64."""
    r = np.ravel(gr)
    w = np.ravel(grw)
    P = np.empty((n_orb, r.size))
    Q = np.empty((n_orb, r.size))
    ls = np.empty(n_orb, dtype=np.int32)
    for i, (l, n, z, c) in enumerate(draw_orbital_params(n_orb, seed)):
        p, q = orbital_samples(r, l, n, z, c)
        norm = np.sqrt(np.sum(w * (p * p + q * q)))
        P[i] = p / norm
        Q[i] = q / norm
        ls[i] = l
    return P, Q, ls


def valid_pairs(ls, k):
    """Pair indices (a <= b, lexicographic numbering) valid at rank k:
    |l_a - l_b| <= k <= l_a + l_b and l_a + l_b + k even."""
    out = []
    n = len(ls)
    p = 0
    for a in range(n):
        for b in range(a, n):
            la, lb = int(ls[a]), int(ls[b])
            if abs(la - lb) <= k <= la + lb and (la + lb + k) % 2 == 0:
                out.append(p)
            p += 1
    return np.asarray(out, dtype=np.int64)


def pair_list(n_orb):
    a, b = np.triu_indices(n_orb)
    return a.astype(np.int32), b.astype(np.int32)


def count_rk(ls, kmax):
    """Number of R^k values of the all-quartet workload (pair-pairs once)."""
    total = 0
    for k in range(kmax + 1):
        m = len(valid_pairs(ls, k))
        total += m * (m + 1) // 2
    return total
