"""Y^k-quadrature two-electron adapters — drop-in for numerics/bspline/slater_yk.py.

R^k/F^k/G^k from B-spline coefficients (orbitals evaluated on the Gauss grid on
the device, pair densities, Y^k scan, contraction) and the Y^k-form exchange
matrix, all on the GPU.
"""

from __future__ import annotations

import numpy as np

from . import device as dv
from . import grid as _grid
from ._lib import call
from .slater import slater_rk_many

__all__ = ["compute_slater_rk", "compute_slater_fk", "compute_slater_gk",
           "compute_exchange_matrix", "eval_orbitals_device"]


def eval_orbitals_device(coeffs, bsp, nv, ks, ns):
    """Samples of several B-spline expansions (evaluate.py:71-97) as a device
    tensor (n_orb, nv*ng); ``coeffs`` (n_orb, >= ns) or (ns,)."""
    c = np.atleast_2d(np.asarray(coeffs, dtype=np.float64))
    n_orb = c.shape[0]
    cc = np.zeros((n_orb, ns))
    m = min(c.shape[1], ns)
    cc[:, :m] = c[:, :m]
    bsp = np.ascontiguousarray(bsp, dtype=np.float64)
    ng = bsp.shape[1]
    out = dv.empty((n_orb, nv * ng))
    cd, bd = dv.f64(cc), dv.f64(bsp)
    call("ax_bspline_eval", dv.ptr(cd), ns, n_orb, dv.ptr(bd), nv, ng, ks, ns, dv.ptr(out),
         nv * ng, dv.stream_handle())
    return out


def _pq_samples(P_list, Q_list, bd):
    nv = bd["nv"]
    ksp, nsp = bd.get("ksp", bd["ks"]), bd.get("nsp", bd["ns"])
    ksq, nsq = bd.get("ksq", bd["ks"]), bd.get("nsq", bd["ns"])
    fp = eval_orbitals_device(np.stack([_pad(p, nsp) for p in P_list]), bd["bsp_p"], nv, ksp, nsp)
    fq = eval_orbitals_device(np.stack([_pad(q, nsq) for q in Q_list]),
                              bd.get("bsp_q", bd["bsp_p"]), nv, ksq, nsq)
    return fp, fq


def _pad(v, n):
    v = np.asarray(v, dtype=np.float64)
    out = np.zeros(n)
    m = min(v.size, n)
    out[:m] = v[:m]
    return out


def compute_slater_rk(Pa, Qa, Pb, Qb, Pc, Qc, Pd, Qd, bsp_data, k):
    """General two-electron Slater integral R^k(ab, cd) (slater_yk.py:159-196)."""
    bd = bsp_data
    gr, grw = bd["gr"], bd["grw"]
    nv, ng = bd["nv"], gr.shape[1]
    fp, fq = _pq_samples([Pa, Pb, Pc, Pd], [Qa, Qb, Qc, Qd], bd)
    n = nv * ng
    rho = dv.empty((2, n))
    pa, pb = dv.i32([0, 2]), dv.i32([1, 3])
    call("ax_pair_density", dv.ptr(fp), dv.ptr(fq), n, n, 2, dv.ptr(pa), dv.ptr(pb), None,
         dv.ptr(rho), n, dv.stream_handle())
    tables = _grid.get(gr, grw, nv, ng, int(k))
    v = slater_rk_many(rho[0:1].contiguous(), rho[1:2].contiguous(), tables, [int(k)])
    return float(dv.host(v)[0])


def compute_slater_fk(orb_a_P, orb_a_Q, orb_b_P, orb_b_Q, bsp_data, k):
    """Direct Slater integral F^k(a,b) = R^k(aa,bb) (slater_yk.py:199-202)."""
    return compute_slater_rk(orb_a_P, orb_a_Q, orb_a_P, orb_a_Q,
                             orb_b_P, orb_b_Q, orb_b_P, orb_b_Q, bsp_data, k)


def compute_slater_gk(orb_a_P, orb_a_Q, orb_b_P, orb_b_Q, bsp_data, k):
    """Exchange Slater integral G^k(a,b) = R^k(ab,ba) (slater_yk.py:205-208)."""
    return compute_slater_rk(orb_a_P, orb_a_Q, orb_b_P, orb_b_Q,
                             orb_b_P, orb_b_Q, orb_a_P, orb_a_Q, bsp_data, k)


def compute_exchange_matrix(orb_b_P, orb_b_Q, bsp_data, k):
    """Non-local exchange matrices K_PP, K_QQ (slater_yk.py:125-152)."""
    bd = bsp_data
    nv, ns = bd["nv"], bd["ns"]
    gr, grw = bd["gr"], bd["grw"]
    ng = gr.shape[1]
    r, w = dv.f64(np.ravel(gr)), dv.f64(np.ravel(grw))
    out = []
    for coeffs, bsp, ksx, nsx in ((orb_b_P, bd["bsp_p"], bd.get("ksp", bd["ks"]), bd.get("nsp", ns)),
                                  (orb_b_Q, bd.get("bsp_q", bd["bsp_p"]), bd.get("ksq", bd["ks"]),
                                   bd.get("nsq", ns))):
        fb = eval_orbitals_device(_pad(coeffs, nsx)[None, :], bsp, nv, ksx, nsx)
        bspd = dv.f64(bsp)
        K = dv.empty((nsx, nsx))
        call("ax_exchange_matrix_yk", dv.ptr(fb), dv.ptr(bspd), dv.ptr(r), dv.ptr(w), nv, ng, ksx,
             nsx, int(k), dv.ptr(K), dv.stream_handle())
        out.append(dv.host(K))
    return tuple(out)
