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
    ls = np.asarray(ls, dtype=np.int64)
    a, b = np.triu_indices(ls.size)
    la, lb = ls[a], ls[b]
    ok = (np.abs(la - lb) <= k) & (k <= la + lb) & ((la + lb + k) % 2 == 0)
    return np.nonzero(ok)[0].astype(np.int64)


def pair_list(n_orb):
    a, b = np.triu_indices(n_orb)
    return a.astype(np.int32), b.astype(np.int32)


def count_rk(ls, kmax):
    """Number of R^k values of the all-quartet workload (pair-pairs once),
    counted by angular momentum class: the pairs of orbitals with l1 < l2
    number c(l1) c(l2), those with l1 = l2 number c(c + 1) / 2, and a class
    is valid at rank k when |l1 - l2| <= k <= l1 + l2 and l1 + l2 + k is
    even (as ``valid_pairs``)."""
    cnt = {}
    for l in (ls.tolist() if isinstance(ls, np.ndarray) else list(ls)):
        cnt[l] = cnt.get(l, 0) + 1
    cls = sorted(cnt.items())
    kmax = int(kmax)
    m = [0] * (kmax + 1)
    for i, (l1, c1) in enumerate(cls):
        for l2, c2 in cls[i:]:
            n = c1 * (c1 + 1) // 2 if l1 == l2 else c1 * c2
            for k in range(l2 - l1, min(l1 + l2, kmax) + 1, 2):  # k = l2 - l1, ..., l1 + l2
                m[k] += n
    return sum(x * (x + 1) // 2 for x in m)


# ---------------------------------------------------------------------------
# Serving workload: a Poisson stream of Slater-integral and Fock-build requests
# ---------------------------------------------------------------------------
def _log_uniform(rng, lo, hi):
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


_BSP_CACHE = {}


def serving_bspline_data(nv, ks=8, ng=8):
    """B-spline workspace on the exponential grid of ``nv`` intervals (cached)."""
    from .bspline import prepare_bspline_data
    key = (nv, ks, ng)
    if key not in _BSP_CACHE:
        ns = nv + ks - 1
        t = make_exponential_knots(ns, ks, 1e-5, 60.0)
        _BSP_CACHE[key] = prepare_bspline_data(t, ns, ks, n_gauss=ng)
    return _BSP_CACHE[key]


def exchange_coefficients(ls, kmax):
    """Synthetic exchange weights c[t, b, k] = 1/(2k+1) where the triangle and
    parity rules of (l_t, l_b, k) hold, k <= kmax."""
    n = len(ls)
    c = np.zeros((n, n, kmax + 1))
    for t in range(n):
        for b in range(n):
            lt, lb = int(ls[t]), int(ls[b])
            for k in range(abs(lt - lb), min(lt + lb, kmax) + 1):
                if (lt + lb + k) % 2 == 0:
                    c[t, b, k] = 1.0 / (2 * k + 1)
    return c


def serving_stream(n_requests, rate=200.0, seed=0, fock_fraction=0.2, precision="split"):
    """[(arrival_s, request)]: seeded exponential inter-arrival times of mean
    1/rate; per request n_orb ~ logU[5, 60], N ~ logU[1000, 6000] (nv =
    round(N/8), 8 Gauss points), k_max ~ logU[2, 10] (rounded); a fraction
    ``fock_fraction`` are Fock-build requests (random decaying B-spline
    coefficients of order 8 on the same kind of grid), the rest all-R^k
    requests on synthetic orbitals."""
    from .serve import FockRequest, SlaterRequest
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        n_orb = int(round(_log_uniform(rng, 5, 60)))
        nv = int(round(_log_uniform(rng, 1000, 6000) / 8))
        kmax = int(round(_log_uniform(rng, 2, 10)))
        sub = int(rng.integers(0, 2 ** 31 - 1))
        if rng.uniform() < fock_fraction:
            bd = serving_bspline_data(nv)
            r2 = np.random.default_rng(sub)
            ns = bd["ns"]
            decay = np.exp(-np.arange(ns) / max(ns / 8.0, 1.0))
            P = r2.normal(size=(n_orb, ns)) * decay
            Q = 0.01 * r2.normal(size=(n_orb, ns)) * decay
            ls = [p[0] for p in draw_orbital_params(n_orb, sub)]
            req = FockRequest(bd, P, Q, np.full(n_orb, 2.0), exchange_coefficients(ls, kmax))
        else:
            gr, grw = synthetic_grid(nv=nv, ng=8)
            P, Q, ls = synthetic_orbitals(gr, grw, n_orb=n_orb, seed=sub)
            req = SlaterRequest(gr, grw, P, Q, ls, kmax, precision=precision)
        out.append((t, req))
    return out


# ---------------------------------------------------------------------------
# CI / CCSD B-spline product workload
# ---------------------------------------------------------------------------
def bspline_product_workload(ns=120, ks=8, r_min=1e-5, r_max=50.0, ng=8):
    """B-splines of order ``ks`` on make_exponential_knots(ns, ks, r_min, r_max)
    sampled on the Gauss grid, and the band pairs (i, i+s), s < ks, i ascending
    then s ascending.  Returns (gr, grw, B (ns, nv*ng), pa, pb)."""
    from .bspline import evaluate_bsplines_on_grid
    t = make_exponential_knots(ns, ks, r_min, r_max)
    gr, grw = setup_gauss_points(t, ns, ks, ng)
    nv = ns - ks + 1
    bsp = evaluate_bsplines_on_grid(t, ns, ks, gr)[:, :, :, 0]  # (nv, ng, ks)
    B = np.zeros((ns, nv * ng))
    for iv in range(nv):
        for i in range(ks):
            B[iv + i, iv * ng:(iv + 1) * ng] = bsp[iv, :, i]
    pa, pb = [], []
    for i in range(ns):
        for s in range(ks):
            if i + s < ns:
                pa.append(i)
                pb.append(i + s)
    return gr, grw, B, np.asarray(pa, np.int32), np.asarray(pb, np.int32)


# ---------------------------------------------------------------------------
# Batched uneven SCF: the 26 charge states of iron
# ---------------------------------------------------------------------------
# relativistic subshells in filling order: (label, kappa, capacity)
_FE_SHELLS = (("1s", -1, 2), ("2s", -1, 2), ("2p-", 1, 2), ("2p+", -2, 4), ("3s", -1, 2),
              ("3p-", 1, 2), ("3p+", -2, 4), ("3d-", 2, 4), ("3d+", -3, 6), ("4s", -1, 2))


def fe_configuration(charge):
    """Ground configuration of Fe^{charge+} as [(label, kappa, occupation)]:
    neutral and singly ionised iron keep 4s (3d^6 4s^2, 3d^6 4s^1); from
    Fe^2+ on the 3d shell holds what is left above the argon core; below it
    the shells fill in order (an average-of-configuration occupation per
    relativistic subshell, j-subshells filled lower j first)."""
    n_e = 26 - int(charge)
    if not 1 <= n_e <= 26:
        raise ValueError("charge must be in 0..25")
    four_s = {0: 2, 1: 1}.get(int(charge), 0)
    rest = n_e - four_s
    out = []
    for label, kappa, cap in _FE_SHELLS[:-1]:
        if rest <= 0:
            break
        occ = min(cap, rest)
        out.append((label, kappa, float(occ)))
        rest -= occ
    if four_s:
        out.append(("4s", -1, float(four_s)))
    return out


def fe_ion_workload(seed=0, ns_range=(40, 160), ks_choices=(7, 8), r_min=1e-4, r_max=30.0):
    """The 26 iron ions of the batched uneven Fock build: per ion a basis of
    ns ~ U{40..160} B-splines of order ks ~ U{7, 8} (kinetic balance, P in
    order ks-1) on make_exponential_knots(ns, ks, r_min, r_max), the ground
    configuration's subshells, and seeded smooth coefficient vectors
    (decaying Gaussian noise, unit norm).  Returns a list of dicts with keys
    charge, bsp_data, kappas, occupations, P, Q."""
    from .bspline import prepare_bspline_data
    rng = np.random.default_rng(seed)
    ions = []
    for charge in range(26):
        ns = int(rng.integers(ns_range[0], ns_range[1] + 1))
        ks = int(rng.choice(ks_choices))
        t = make_exponential_knots(ns, ks, r_min, r_max)
        bd = prepare_bspline_data(t, ns, ks, kinetic_balance=True)
        conf = fe_configuration(charge)
        n_orb = len(conf)
        decay = np.exp(-np.arange(bd["ns"]) / max(bd["ns"] / 6.0, 1.0))
        P = rng.normal(size=(n_orb, bd["ns"])) * decay
        Q = 0.05 * rng.normal(size=(n_orb, bd["ns"])) * decay
        P[:, bd["nsp"]:] = 0.0
        norm = np.sqrt((P * P).sum(axis=1) + (Q * Q).sum(axis=1))
        ions.append({"charge": charge, "bsp_data": bd, "kappas": [c[1] for c in conf],
                     "occupations": [c[2] for c in conf], "P": P / norm[:, None],
                     "Q": Q / norm[:, None]})
    return ions


def small_atom_requests(n_atoms, seed=0, precision="fp64"):
    """All-R^k requests of small atoms and ions (the launch-bound regime of
    per-atom processing): nv ~ U{40..160} cells of 8 Gauss nodes on
    r 1e-4..30, 1..10 orbitals with l <= 3, k_max = 2 max(l) (at most 6)."""
    from .serve import SlaterRequest
    rng = np.random.default_rng(seed)
    out = []
    for a in range(n_atoms):
        nv = int(rng.integers(40, 161))
        n_orb = int(rng.integers(1, 11))
        gr, grw = synthetic_grid(nv=nv, ng=8, r_min=1e-4, r_max=30.0)
        P, Q, _ = synthetic_orbitals(gr, grw, n_orb=n_orb, seed=seed * 100003 + a)
        ls = rng.integers(0, 4, size=n_orb).astype(np.int32)
        kmax = int(min(6, 2 * ls.max()))
        out.append(SlaterRequest(gr, grw, P, Q, ls, kmax, precision))
    return out
