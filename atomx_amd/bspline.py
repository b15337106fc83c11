"""B-spline radial geometry: knots, Gauss-Legendre nodes, B-spline tables.

These are the input tables of the Slater-integral engine (the ``bsp_data``
workspace).  They are built once per grid on the host in FP64, exactly as the
reference builds them, and then uploaded to the device by the engine.

Reference behaviour followed:
  * ``make_knots`` / ``make_exponential_knots``  numerics/bspline/knots.py:10-45
  * ``make_physics_knots``                       numerics/bspline/knots.py:48-166
  * ``gauss_legendre_points``                    numerics/bspline/basis.py:195-208
  * ``setup_gauss_points``                       numerics/bspline/basis.py:211-249
  * ``evaluate_bsplines_on_grid``                numerics/bspline/basis.py:252-341
  * ``prepare_bspline_data``                     numerics/bspline/basis.py:348-438
  * ``eval_orbital_on_grid`` (host convenience)  numerics/bspline/evaluate.py:71-97
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "make_knots",
    "make_exponential_knots",
    "make_physics_knots",
    "gauss_legendre_points",
    "setup_gauss_points",
    "evaluate_bsplines_on_grid",
    "prepare_bspline_data",
    "eval_orbital_on_grid",
    "bsplvb",
]


def _check_knot_request(ns, ks, r_min, r_max):
    """Shared argument check of the exponential knot builders: at least two
    cells and a positive, increasing radial range."""
    if ks < 1 or ns - ks + 1 < 2:
        raise ValueError(f"exponential knots need order ks >= 1 and ns >= ks + 1 "
                         f"splines (at least two cells); got ns={ns}, ks={ks}")
    if not (0 < r_min < r_max):
        raise ValueError(f"exponential knots need 0 < r_min < r_max; got r_min={r_min}, "
                         f"r_max={r_max}")


def make_knots(ns: int, ks: int, r_min: float, r_max: float, *, mode: str = "exponential"):
    """Exponential knot sequence of ns splines of order ks on [0, r_max] with
    the first breakpoint at r_min (knots.py:10-17; ``mode`` is accepted for
    the reference's signature, only the exponential grid exists)."""
    _check_knot_request(ns, ks, r_min, r_max)
    return make_exponential_knots(ns, ks, r_min, r_max)


def make_exponential_knots(ns: int, ks: int, r_min: float, r_max: float, *, power: float = 1.0):
    """Knots [0]*ks, breakpoints r_min .. r_max, [r_max]*ks (knots.py:20-45).

    The nv = ns - ks + 1 breakpoints are r_min * exp(log(r_max / r_min) * u^power)
    at nv equally spaced u in [0, 1].  For power 1 they are formed as the
    running product r_min * h * h * ... (h = (r_max / r_min)^(1/(nv - 1)),
    one multiplication per breakpoint, left to right), which is the rounding
    the reference's recurrence produces, so the knots agree bit for bit.
    """
    if not (np.isscalar(power) and np.isrealobj(power) and np.isfinite(power) and power > 0):
        raise ValueError(f"grid power must be a positive finite real number; got {power!r}")
    _check_knot_request(ns, ks, r_min, r_max)
    nv = ns - ks + 1
    if power == 1.0:
        h = np.exp(np.log(r_max / r_min) / (nv - 1))
        factors = np.full(nv - 1, h)
        inner = np.cumprod(np.concatenate(([float(r_min)], factors)))[:-1]
    else:
        u = np.linspace(0.0, 1.0, nv)
        inner = (r_min * np.exp(np.log(r_max / r_min) * u ** power))[:-1]
    return np.concatenate((np.zeros(ks), inner, np.full(ks, float(r_max))))


def make_physics_knots(ks: int, Z: int, ns: int = 113, r_max: float | None = None,
                       nuclear: str = "point", hi: float | None = None,
                       he: float | None = None, hmax: float | None = None, ml: int = 1):
    """DBSR_HF semi-exponential grid, point-nucleus form (knots.py:48-166).

    Returns ``(t, ns)``.  Only the point-nucleus initial step hs = hi/Z is
    provided here (the Fermi step belongs to the nuclear-model package).
    """
    if nuclear != "point":
        raise ValueError("only the point-nucleus physics grid is provided")
    if r_max is None:
        r_max = 50.0
    if hi is None:
        hi = 0.25
    if he is None:
        he = 0.0625 if Z <= 2 else 0.25
    if hmax is None:
        hmax = 1.0
    hs = hi / max(Z, 1)
    breakpoints = [hs * i for i in range(1, ml + 1)]
    ti = breakpoints[-1]
    ht = 1.0 + he
    while True:
        tj = ti * ht
        if (tj - ti) > hmax:
            break
        ti = tj
        if ti > r_max:
            break
        breakpoints.append(ti)
    if ti < r_max:
        while True:
            ti_next = ti + hmax
            if ti_next >= r_max:
                break
            ti = ti_next
            breakpoints.append(ti)
    tmax = breakpoints[-1] + hmax
    breakpoints.append(tmax)
    nv = len(breakpoints)
    ns_out = nv + ks - 1
    t = np.zeros(ns_out + ks)
    t[ks:ks + nv] = breakpoints
    t[ns_out:ns_out + ks] = tmax
    return t, ns_out


def gauss_legendre_points(n: int):
    """Gauss-Legendre nodes and weights on [-1, 1] (basis.py:195-208)."""
    return np.polynomial.legendre.leggauss(n)


def setup_gauss_points(t, ns, ks, n_gauss=None):
    """Nodes ``gr`` and weights ``grw`` (nv, n_gauss) per knot interval (basis.py:211-249)."""
    if n_gauss is None:
        n_gauss = ks
    nv = ns - ks + 1
    xi, wi = gauss_legendre_points(n_gauss)
    a = np.asarray(t[ks - 1:ks - 1 + nv], dtype=float)
    b = np.asarray(t[ks:ks + nv], dtype=float)
    half = 0.5 * (b - a)
    mid = 0.5 * (a + b)
    gr = mid[:, None] + half[:, None] * xi[None, :]
    grw = half[:, None] * wi[None, :]
    return gr, grw


def bsplvb(t, ks, x, left):
    """All ks nonzero B-splines of order ks at x, t[left] <= x < t[left+1] (Cox-de Boor)."""
    b = np.zeros(ks)
    b[0] = 1.0
    dr = np.zeros(ks)
    dl = np.zeros(ks)
    for j in range(1, ks):
        dr[j] = t[left + j] - x
        dl[j] = x - t[left + 1 - j]
        saved = 0.0
        for r in range(j):
            term = b[r] / (dr[r + 1] + dl[j - r])
            b[r] = saved + dr[r + 1] * term
            saved = dl[j - r] * term
        b[j] = saved
    return b


def _bsplvb_vec(t, ks, x, left):
    """Vectorised Cox-de Boor over arrays x/left; returns (npts, ks)."""
    npts = x.shape[0]
    b = np.zeros((npts, ks))
    b[:, 0] = 1.0
    dr = np.zeros((npts, ks))
    dl = np.zeros((npts, ks))
    for j in range(1, ks):
        dr[:, j] = t[left + j] - x
        dl[:, j] = x - t[left + 1 - j]
        saved = np.zeros(npts)
        for r in range(j):
            term = b[:, r] / (dr[:, r + 1] + dl[:, j - r])
            b[:, r] = saved + dr[:, r + 1] * term
            saved = dl[:, j - r] * term
        b[:, j] = saved
    return b


def evaluate_bsplines_on_grid(t, ns, ks, gr, nderiv=1):
    """B-spline values (and first derivatives) at all grid points (basis.py:252-341).

    Returns ``(nv, n_gauss, ks, nderiv)``.
    """
    t = np.asarray(t, dtype=float)
    nv, ng = gr.shape
    x = np.asarray(gr, dtype=float).ravel()
    left = np.repeat(ks - 1 + np.arange(nv), ng)
    out = np.zeros((nv * ng, ks, nderiv))
    out[:, :, 0] = _bsplvb_vec(t, ks, x, left)
    if nderiv >= 2:
        kk = ks - 1
        blow = _bsplvb_vec(t, kk, x, left) if kk >= 1 else np.zeros((nv * ng, 0))
        fac = float(kk)
        for i in range(ks):
            gi = left - ks + 1 + i
            b_lo = np.zeros(nv * ng)
            b_hi = np.zeros(nv * ng)
            if 0 <= i - 1 < kk:
                denom = t[gi + kk] - t[gi]
                ok = np.abs(denom) > 1e-30
                b_lo[ok] = blow[ok, i - 1] / denom[ok]
            if 0 <= i < kk:
                denom = t[gi + ks] - t[gi + 1]
                ok = np.abs(denom) > 1e-30
                b_hi[ok] = blow[ok, i] / denom[ok]
            out[:, i, 1] = fac * (b_lo - b_hi)
    return out.reshape(nv, ng, ks, nderiv)


def _build_knot_sequence(t_master, ks_master, ns_master, k_target):
    """Knots of order k_target on the breakpoints of t_master (basis.py:348-361):
    the interior knots are shared, each end knot repeated k_target times.
    Returns (knots, number of splines)."""
    t_master = np.asarray(t_master, dtype=float)
    interior = t_master[ks_master:ns_master]
    knots = np.concatenate((np.full(k_target, t_master[0]), interior,
                            np.full(k_target, t_master[ns_master])))
    return knots, interior.size + k_target


def prepare_bspline_data(t, ns, ks, n_gauss=None, kinetic_balance=False):
    """The ``bsp_data`` workspace dict (basis.py:364-438).

    Keys: nv, ks, ksp, ksq, nsp, nsq, gr, grw, bsp_p, bsp_q, bsp_p_d, bsp_q_d,
    t, tp, ns.  With kinetic balance ksp = ks-1 on its own knot sequence tp.
    """
    t = np.asarray(t, dtype=float)
    nv = ns - ks + 1
    gr, grw = setup_gauss_points(t, ns, ks, n_gauss)
    ksp, ksq = (ks - 1, ks) if kinetic_balance else (ks, ks)
    nsp = nv + ksp - 1
    nsq = nv + ksq - 1
    full_q = evaluate_bsplines_on_grid(t, nsq, ksq, gr, nderiv=2)
    bsp_q = np.ascontiguousarray(full_q[:, :, :ksq, 0])
    bsp_q_d = np.ascontiguousarray(full_q[:, :, :ksq, 1])
    tp = t
    if ksp == ksq:
        bsp_p = bsp_q.copy()
        bsp_p_d = bsp_q_d.copy()
    else:
        tp, nsp_check = _build_knot_sequence(t, ks, ns, ksp)
        assert nsp_check == nsp
        full_p = evaluate_bsplines_on_grid(tp, nsp, ksp, gr, nderiv=2)
        bsp_p = np.ascontiguousarray(full_p[:, :, :ksp, 0])
        bsp_p_d = np.ascontiguousarray(full_p[:, :, :ksp, 1])
    return {
        "nv": nv, "ks": ks, "ksp": ksp, "ksq": ksq, "nsp": nsp, "nsq": nsq,
        "gr": gr, "grw": grw, "bsp_p": bsp_p, "bsp_q": bsp_q,
        "bsp_q_d": bsp_q_d, "bsp_p_d": bsp_p_d, "t": t, "tp": tp,
        "ns": max(nsp, nsq),
    }


def eval_orbital_on_grid(coeffs, bsp, nv, ks, ns):
    """f(r) = sum_i c_i B_i(r) on the Gauss grid (evaluate.py:71-97), host FP64."""
    gi = np.arange(nv)[:, None] + np.arange(ks)[None, :]
    gi_safe = np.clip(gi, 0, ns - 1)
    c = np.where(gi < ns, np.asarray(coeffs)[gi_safe], 0.0)
    return np.einsum("is,ims->im", c, bsp)
