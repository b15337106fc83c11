"""Shared inputs for the parity tests (grids, closed-form densities, fixtures)."""

import json
import os

import numpy as np

from atomx_amd import bspline as B

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Hydrogenic 1s (Z = 1) scalar references: R^0 = 5/8, R^1 = 3/8 (closed forms),
# R^2 by a 45-digit quadrature of the definition integral
# (reference tests/test_slater_kernel_crosscheck.py:166-170).
RK_1S = {0: 5.0 / 8.0, 1: 3.0 / 8.0, 2: 0.260532333438687426013570914996237633812003225}


def physics_workspace(Z=1, ks=9, r_max=50.0, kinetic_balance=False):
    t, ns = B.make_physics_knots(ks=ks, Z=Z, r_max=r_max)
    return B.prepare_bspline_data(t, ns, ks, kinetic_balance=kinetic_balance), t


def rho1s(r, Z=1.0):
    return 4.0 * Z ** 3 * r ** 2 * np.exp(-2.0 * Z * r)


def p1s(r, Z=1.0):
    return 2.0 * Z ** 1.5 * r * np.exp(-Z * r)


def project_1s(bd, t, Z=1.0):
    """L2 projection of the exact hydrogenic 1s onto the B-spline basis, with a
    48-point Gauss rule per interval for the right-hand side
    (reference tests/test_slater_kernel_crosscheck.py:256-281)."""
    ks, ns, nv = bd["ks"], bd["ns"], bd["nv"]
    bsp, grw = bd["bsp_p"], bd["grw"]
    overlap = np.zeros((ns, ns))
    rhs = np.zeros(ns)
    xq, wq = np.polynomial.legendre.leggauss(48)
    for iv in range(nv):
        idx = iv + np.arange(ks)
        block = bsp[iv]
        overlap[np.ix_(idx, idx)] += (block * grw[iv][:, None]).T @ block
        a, b = t[ks - 1 + iv], t[ks + iv]
        pts = 0.5 * (b - a) * xq + 0.5 * (a + b)
        wts = 0.5 * (b - a) * wq
        vals = np.array([B.bsplvb(t, ks, float(x), ks - 1 + iv) for x in pts])
        rhs[idx] += (wts * p1s(pts, Z)) @ vals
    return np.linalg.solve(overlap, rhs), overlap


def dbsr_state():
    """DBSR_HF Ne fixed spinors: (bsp_data, P, Q, occ, kappas, reference json)."""
    ref = json.load(open(os.path.join(GOLDEN, "dbsr_ne_reference.json")))
    with np.load(os.path.join(GOLDEN, "dbsr_ne_spinors.npz"), allow_pickle=False) as a:
        knots, pa, qa, labels = (a[k].copy() for k in ("knots", "pa", "qa", "labels"))
    bd = B.prepare_bspline_data(knots, ref["ns"], ref["ks"], kinetic_balance=True)
    order = ["1s1/2", "2s1/2", "2p1/2", "2p3/2"]
    kappa = {"1s1/2": -1, "2s1/2": -1, "2p1/2": 1, "2p3/2": -2}
    occ = {"1s1/2": 2.0, "2s1/2": 2.0, "2p1/2": 2.0, "2p3/2": 4.0}
    labels = [str(x) for x in labels]
    idx = [labels.index(lab) for lab in order]
    return (bd, pa[idx].copy(), qa[idx].copy(), [occ[l] for l in order],
            [kappa[l] for l in order], ref)
