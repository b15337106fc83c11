"""Slater radial integrals R^k on the B-spline quadrature grid — drop-in for
numerics/bspline/slater.py.

    R^k(13, 24) = int int rho13(r1) [r_<^k / r_>^{k+1}] rho24(r2) dr1 dr2

``"poisson"`` (default): Y^k[rho24] by the O(N) scan, contracted with
grw*rho13 (slater.py:118-130).  ``"direct"``: the O(N^2) product quadrature
with the coincident-cell cusp split (slater.py:51-115).  Both run on the GPU.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np

from . import device as dv
from . import grid as _grid
from ._lib import call
from .yk import yk_rows

DEFAULT_METHOD = "poisson"
AVAILABLE_METHODS = ("poisson", "direct")

__all__ = ["DEFAULT_METHOD", "AVAILABLE_METHODS", "slater_rk", "slater_rk_poisson",
           "slater_rk_direct", "slater_rk_many"]


def _dot_rows(D, Y, p, q):
    out = dv.empty((len(p),))
    pd, qd = dv.i32(p), dv.i32(q)  # held until the launch is queued
    call("ax_rk_dot", dv.ptr(D), dv.ld(D), dv.ptr(Y), dv.ld(Y), D.shape[1], len(p),
         dv.ptr(pd), dv.ptr(qd), dv.ptr(out), dv.stream_handle())
    return out


def slater_rk_many(rho13, rho24, tables, ks):
    """R^k for many (rho13[t], rho24[t], ks[t]) on one grid; device tensors in,
    device tensor (n,) out."""
    Y = yk_rows(rho24, tables, ks)
    D = dv.empty(tuple(rho13.shape))
    call("ax_mul_rows", dv.ptr(D), dv.ptr(rho13), dv.ptr(tables.w), rho13.shape[0], rho13.shape[1],
         dv.stream_handle())
    idx = np.arange(rho13.shape[0], dtype=np.int32)
    return _dot_rows(D, Y, idx, idx)


def slater_rk_poisson(rho13, rho24, gr, grw, nv, n_gauss, k):
    """Fast O(N) Slater integral via the Poisson Y^k potential (slater.py:118-130)."""
    nv, n_gauss = int(nv), int(n_gauss)
    a = np.asarray(rho13).reshape(nv * n_gauss)
    b = np.asarray(rho24).reshape(nv * n_gauss)
    tables = _grid.get(np.ravel(gr), np.ravel(grw), nv, n_gauss, int(k))
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        a = a.astype(np.complex128)
        b = b.astype(np.complex128)
        # bilinear in (rho13, rho24): four real contractions
        A = dv.f64(np.stack([a.real, a.imag, a.real, a.imag]))
        B = dv.f64(np.stack([b.real, b.imag, b.imag, b.real]))
        v = dv.host(slater_rk_many(A, B, tables, [int(k)] * 4))
        return complex(v[0] - v[1], v[2] + v[3])
    A = dv.f64(a[None, :].astype(np.float64))
    B = dv.f64(b[None, :].astype(np.float64))
    return float(dv.host(slater_rk_many(A, B, tables, [int(k)]))[0])


@lru_cache(maxsize=16)
def _cell_triangle_rule(n_gauss, order):
    """Gauss nodes and density interpolation for a split square cell (slater.py:63-78).

    On its lower triangle r = a + h x and s = a + h x y, x, y in (0, 1); the
    Jacobian h^2 x cancels the integrable Coulomb corner at a = 0.
    """
    from numpy.polynomial.legendre import leggauss, legvander
    nodes, _ = leggauss(n_gauss)
    x, w = leggauss(order)
    x, w = (x + 1) / 2, w / 2
    inverse = np.linalg.solve(legvander(nodes, n_gauss - 1), np.eye(n_gauss))
    outer = legvander(2 * x - 1, n_gauss - 1) @ inverse
    inner = legvander(2 * x[:, None] * x[None, :] - 1, n_gauss - 1) @ inverse
    return nodes, x, w, outer, inner


def slater_rk_direct(rho13, rho24, gr, grw, k):
    """Independent double quadrature, split along each cell's r = s cusp (slater.py:81-115)."""
    gr, grw = np.asarray(gr), np.asarray(grw)
    if gr.ndim != 2 or gr.shape != grw.shape:
        raise ValueError("direct Slater integration requires a 2-D cell Gauss grid and weights")
    if k < 0 or int(k) != k:
        raise ValueError("Slater multipole rank must be a nonnegative integer")
    k = int(k)
    nv, n_gauss = gr.shape
    a = np.asarray(rho13, dtype=np.float64).reshape(gr.shape)
    b = np.asarray(rho24, dtype=np.float64).reshape(gr.shape)
    order = max(2 * n_gauss, (k + n_gauss + 1) // 2)
    nodes, x, w, outer, inner = _cell_triangle_rule(n_gauss, order)
    dev = [dv.f64(v) for v in (a, b, gr, grw, x, w, outer, inner)]
    ws = dv.empty((1024 + nv,))
    out = dv.empty((1,))
    call("ax_slater_rk_direct", *[dv.ptr(v) for v in dev[:4]], nv, n_gauss, k, order,
         *[dv.ptr(v) for v in dev[4:]], float(nodes[0]), dv.ptr(ws), dv.ptr(out),
         dv.stream_handle())
    return float(dv.host(out)[0])


def slater_rk(rho13, rho24, bsp_data, k, method=DEFAULT_METHOD):
    """Slater radial integral R^k(13, 24) with a selectable integrator (slater.py:133-170)."""
    gr = bsp_data["gr"]
    grw = bsp_data["grw"]
    nv = bsp_data["nv"]
    n_gauss = gr.shape[1]
    if method == "poisson":
        return slater_rk_poisson(rho13, rho24, gr, grw, nv, n_gauss, k)
    if method == "direct":
        return slater_rk_direct(rho13, rho24, gr, grw, k)
    raise ValueError(f"unknown Slater integrator {method!r}; choose from {AVAILABLE_METHODS}")
