"""CPU FP64 oracle for the Slater-integral path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or, for the
baseline, as the timed CPU restatement).  The product package ``atomx_amd``
never imports it.

The compiled part (``slater_oracle.c``, built by ``oracle/Makefile`` into
``oracle/_build/liboracle.so``) restates the reference's numba kernels; this
module adds the numpy-level glue of the reference (bank cache, four-block
scalar R^k, direct/exchange matrices, the per-target Fock build and the
average-level two-electron energy).  Every function cites the reference file
and lines it follows (paths relative to src/atomx/).

Pinning (see tests/test_oracle_*.py): closed-form hydrogenic R^k, the
40-digit mpmath exp(-r) R^k, monomial cell integrals, partial-matrix
invariants, the DBSR_HF Ne two-electron energy on fixed spinors
(tests/golden/dbsr_ne_*), bank symmetry, case1 = case2^T, direct/exchange
blocks contracting to the scalar R^k, and semiseparable = full bank.
Functions without such a check say "parity unpinned" in their docstring.
"""

from __future__ import annotations

import ctypes
import math
import os
from fractions import Fraction
from functools import lru_cache

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile the C oracle with its Makefile (gcc, no reference sources)."""
    import subprocess
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_NATIVE_PATH = os.path.join(_HERE, "_build", "liboracle_native.so")
_native = None


def lib_native():
    """The same C source built -O2 -march=native with OpenMP for the CPU this
    process runs on (rebuilt on first use in each process, since a library
    built elsewhere may name another CPU): the timed CPU baseline.  Its
    values are not the checker's (multiply-adds may be fused)."""
    global _native
    if _native is None:
        import subprocess
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "native"], check=True)
        _native = ctypes.CDLL(_NATIVE_PATH)
    return _native


def host_cpu():
    """nproc, the CPUs this process may use, and lscpu's identifying lines."""
    import subprocess
    info = {"nproc": int(subprocess.run(["nproc"], capture_output=True, text=True).stdout)}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity_cpus"] = os.cpu_count()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keep = ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
            "NUMA node(s)", "CPU max MHz", "L3 cache", "Flags")
    for line in out.splitlines():
        key, _, val = line.partition(":")
        key = key.strip()
        if key in keep:
            val = val.strip()
            if key == "Flags":  # the vector extensions only
                val = " ".join(f for f in val.split() if f.startswith(("avx", "fma", "amx")))
            info["lscpu " + key] = val
    return info


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_slater_rk_poisson.restype = ctypes.c_double
        L.orc_disjoint_cell_integral.restype = ctypes.c_double
        L.orc_sum_amb_banded.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------------------
# Quadrature tables
# ---------------------------------------------------------------------------

def gauss_legendre(n):
    x = np.empty(n)
    w = np.empty(n)
    lib().orc_gauss_legendre(ctypes.c_int(n), _p(x), _p(w))
    return x, w


def gauss_legendre_01(n):
    """convol.py:252-291 (_gauss_legendre_01)."""
    x = np.empty(n)
    w = np.empty(n)
    lib().orc_gauss_legendre_01(ctypes.c_int(n), _p(x), _p(w))
    return x, w


def yk_partial_matrices(n_gauss):
    """numerics/yk.py:65-84."""
    mo = np.empty((n_gauss, n_gauss))
    mi = np.empty((n_gauss, n_gauss))
    lib().orc_yk_partial_matrices(ctypes.c_int(n_gauss), _p(mo), _p(mi))
    return mo, mi


# ---------------------------------------------------------------------------
# (Y) Poisson discretisation
# ---------------------------------------------------------------------------

def yk_flat(rho, r, w, n, k, nv, ng, mo, mi):
    """numerics/yk.py:87-154."""
    rho, r, w, mo, mi = (_c(a) for a in (rho, r, w, mo, mi))
    out = np.empty(n)
    lib().orc_yk_flat(_p(rho), _p(r), _p(w), ctypes.c_long(n), ctypes.c_int(k),
                      ctypes.c_int(nv), ctypes.c_int(ng), _p(mo), _p(mi), _p(out))
    return out


def compute_yk(rho, gr, grw, nv, n_gauss, k):
    """numerics/yk.py:157-185 (real densities)."""
    mo, mi = yk_partial_matrices(n_gauss)
    return yk_flat(np.ravel(rho), np.ravel(gr), np.ravel(grw), nv * n_gauss, k, nv,
                   n_gauss, mo, mi).reshape(nv, n_gauss)


def slater_rk_poisson(rho13, rho24, gr, grw, nv, n_gauss, k):
    """numerics/bspline/slater.py:118-130."""
    yk = compute_yk(np.reshape(rho24, (nv, n_gauss)), gr, grw, nv, n_gauss, k)
    return float(np.sum(grw * np.reshape(rho13, (nv, n_gauss)) * yk))


@lru_cache(maxsize=16)
def _cell_triangle_rule(n_gauss, order):
    """numerics/bspline/slater.py:63-78."""
    from numpy.polynomial.legendre import leggauss, legvander
    nodes, _ = leggauss(n_gauss)
    x, w = leggauss(order)
    x, w = (x + 1) / 2, w / 2
    inverse = np.linalg.solve(legvander(nodes, n_gauss - 1), np.eye(n_gauss))
    outer = legvander(2 * x - 1, n_gauss - 1) @ inverse
    inner = legvander(2 * x[:, None] * x[None, :] - 1, n_gauss - 1) @ inverse
    return nodes, x, w, outer, inner


def slater_rk_direct(rho13, rho24, gr, grw, k):
    """numerics/bspline/slater.py:81-115: O(N^2) double quadrature, cusp split."""
    gr, grw = np.asarray(gr), np.asarray(grw)
    rho13 = np.asarray(rho13, dtype=float).reshape(gr.shape)
    rho24 = np.asarray(rho24, dtype=float).reshape(gr.shape)
    n_gauss = gr.shape[1]
    a13, a24, r, w = (_c(a.ravel()) for a in (rho13, rho24, gr, grw))
    off = lib().orc_disjoint_cell_integral(_p(a13), _p(a24), _p(r), _p(w),
                                           ctypes.c_long(r.size), ctypes.c_int(k),
                                           ctypes.c_int(n_gauss))
    order = max(2 * n_gauss, (k + n_gauss + 1) // 2)
    nodes, x, w1, outer, inner = _cell_triangle_rule(n_gauss, order)
    h = grw.sum(axis=1)
    a = gr[:, 0] - h * (1 + nodes[0]) / 2
    rr = a[:, None, None] + h[:, None, None] * x[None, :, None]
    s = a[:, None, None] + h[:, None, None] * x[None, :, None] * x[None, None, :]
    kj = (s / rr) ** k / rr * h[:, None, None] ** 2 * x[None, :, None]
    f_outer, g_outer = rho13 @ outer.T, rho24 @ outer.T
    f_inner = np.einsum("cj,xyj->cxy", rho13, inner)
    g_inner = np.einsum("cj,xyj->cxy", rho24, inner)
    dens = f_outer[:, :, None] * g_inner + g_outer[:, :, None] * f_inner
    diag = np.sum(dens * kj * w1[None, :, None] * w1[None, None, :])
    return float(off + diag)


def rk_quartets(rho, r, w, nv, ng, qp, qq, qk, nthreads=0, native=False):
    """All requested R^k(p;q) = sum_x w rho_p Y^k[rho_q] (slater.py:118-130 per
    quartet, with Y^k computed once per (q, k)), OpenMP over quartets.
    native: run the -march=native build (the timed baseline) instead."""
    rho = _c(rho)
    r, w = _c(np.ravel(r)), _c(np.ravel(w))
    qp, qq, qk = (_c(a, np.int32) for a in (qp, qq, qk))
    mo, mi = yk_partial_matrices(ng)
    out = np.empty(qp.size)
    kmax = int(qk.max()) if qk.size else 0
    (lib_native() if native else lib()).orc_rk_quartets(_p(rho), ctypes.c_long(rho.shape[0]), _p(r), _p(w),
                          ctypes.c_int(nv), ctypes.c_int(ng), _p(mo), _p(mi),
                          qp.ctypes.data_as(_I), qq.ctypes.data_as(_I), qk.ctypes.data_as(_I),
                          ctypes.c_long(qp.size), ctypes.c_int(kmax), ctypes.c_int(nthreads),
                          _p(out))
    return out


def exchange_matrix_kernel(fb, bsp, gr, grw, nv, ng, ks, ns, k):
    """numerics/bspline/slater_yk.py:35-122."""
    fb, bsp, gr, grw = (_c(a) for a in (fb, bsp, gr, grw))
    K = np.empty((ns, ns))
    lib().orc_exchange_matrix_kernel(_p(fb), _p(bsp), _p(gr), _p(grw), ctypes.c_int(nv),
                                     ctypes.c_int(ng), ctypes.c_int(ks), ctypes.c_int(ns),
                                     ctypes.c_int(k), _p(K))
    return K


def assemble_integral(nv, ks1, ks2, mm, bsp1, bsp2, weights):
    """numerics/bspline/integrals.py:18-56: <B1_i | w | B2_j>, (mm, mm)."""
    bsp1, bsp2, weights = _c(bsp1), _c(bsp2), _c(weights)
    ng = bsp1.shape[1]
    mat = np.empty((mm, mm))
    lib().orc_assemble_integral(ctypes.c_int(nv), ctypes.c_int(ng), ctypes.c_int(ks1),
                                ctypes.c_int(ks2), ctypes.c_int(mm), _p(bsp1), _p(bsp2),
                                _p(weights), _p(mat))
    return mat


def direct_potential_yk(P_stack, Q_stack, occ, bsp_data, target=None):
    """methods/dirac_fock.py:929-965 (_build_direct_potential_yk): Y^0 of the
    occupation-weighted density (minus one electron of ``target``), assembled
    in the P and Q bases and symmetrised.  Returns (V_J_PP, V_J_QQ)."""
    bd = bsp_data
    nv, ns = bd["nv"], bd["ns"]
    gr, grw = bd["gr"], bd["grw"]
    ng = gr.shape[1]
    ksp, nsp = bd.get("ksp", bd["ks"]), bd.get("nsp", ns)
    ksq, nsq = bd.get("ksq", bd["ks"]), bd.get("nsq", ns)
    bsp_p, bsp_q = bd["bsp_p"], bd.get("bsp_q", bd["bsp_p"])
    rho = np.zeros((nv, ng))
    for P, Q, q in zip(P_stack, Q_stack, occ):
        fP = eval_orbital_on_grid(P, bsp_p, nv, ksp, nsp)
        fQ = eval_orbital_on_grid(Q, bsp_q, nv, ksq, nsq)
        rho += q * (fP ** 2 + fQ ** 2)
    if target is not None:
        fP = eval_orbital_on_grid(P_stack[target], bsp_p, nv, ksp, nsp)
        fQ = eval_orbital_on_grid(Q_stack[target], bsp_q, nv, ksq, nsq)
        rho -= fP ** 2 + fQ ** 2
    yk = compute_yk(rho, gr, grw, nv, ng, 0)
    wts = grw * yk
    vpp = assemble_integral(nv, ksp, ksp, ns, bsp_p, bsp_p, wts)
    vqq = assemble_integral(nv, ksq, ksq, ns, bsp_q, bsp_q, wts)
    return 0.5 * (vpp + vpp.T), 0.5 * (vqq + vqq.T)


def eval_orbital_on_grid(coeffs, bsp, nv, ks, ns):
    """numerics/bspline/evaluate.py:71-97."""
    gi = np.arange(nv)[:, None] + np.arange(ks)[None, :]
    gi_safe = np.clip(gi, 0, ns - 1)
    c = np.where(gi < ns, np.asarray(coeffs)[gi_safe], 0.0)
    return np.einsum("is,ims->im", c, bsp)


def compute_slater_rk(Pa, Qa, Pb, Qb, Pc, Qc, Pd, Qd, bsp_data, k):
    """numerics/bspline/slater_yk.py:159-196."""
    bd = bsp_data
    nv = bd["nv"]
    ksp, ksq = bd.get("ksp", bd["ks"]), bd.get("ksq", bd["ks"])
    nsp, nsq = bd.get("nsp", bd["ns"]), bd.get("nsq", bd["ns"])
    ev_p = lambda c: eval_orbital_on_grid(c, bd["bsp_p"], nv, ksp, nsp)
    ev_q = lambda c: eval_orbital_on_grid(c, bd.get("bsp_q", bd["bsp_p"]), nv, ksq, nsq)
    rho_ab = ev_p(Pa) * ev_p(Pb) + ev_q(Qa) * ev_q(Qb)
    rho_cd = ev_p(Pc) * ev_p(Pd) + ev_q(Qc) * ev_q(Qd)
    yk = compute_yk(rho_cd, bd["gr"], bd["grw"], nv, bd["gr"].shape[1], k)
    return float(np.sum(bd["grw"] * rho_ab * yk))


def compute_exchange_matrix(orb_b_P, orb_b_Q, bsp_data, k):
    """numerics/bspline/slater_yk.py:125-152."""
    bd = bsp_data
    nv, ns = bd["nv"], bd["ns"]
    ng = bd["gr"].shape[1]
    ksp, nsp = bd.get("ksp", bd["ks"]), bd.get("nsp", ns)
    ksq, nsq = bd.get("ksq", bd["ks"]), bd.get("nsq", ns)
    fp = eval_orbital_on_grid(orb_b_P, bd["bsp_p"], nv, ksp, nsp)
    fq = eval_orbital_on_grid(orb_b_Q, bd.get("bsp_q", bd["bsp_p"]), nv, ksq, nsq)
    kpp = exchange_matrix_kernel(fp, bd["bsp_p"], bd["gr"], bd["grw"], nv, ng, ksp, nsp, k)
    kqq = exchange_matrix_kernel(fq, bd.get("bsp_q", bd["bsp_p"]), bd["gr"], bd["grw"], nv, ng,
                                 ksq, nsq, k)
    return kpp, kqq


# ---------------------------------------------------------------------------
# (C) cell algorithm
# ---------------------------------------------------------------------------

def moments_sym(bsp, gr, grw, nv, ks, k):
    """convol.py:57-90."""
    bsp, gr, grw = _c(bsp), _c(gr), _c(grw)
    out = np.empty((ks * (ks + 1) // 2, nv))
    lib().orc_moments_sym(_p(bsp), _p(gr), _p(grw), ctypes.c_int(nv), ctypes.c_int(gr.shape[1]),
                          ctypes.c_int(ks), ctypes.c_int(k), _p(out))
    return out


def moments_nonsym(bsp1, bsp2, gr, grw, nv, ks1, ks2, k):
    """convol.py:93-128."""
    bsp1, bsp2, gr, grw = _c(bsp1), _c(bsp2), _c(gr), _c(grw)
    out = np.empty((ks1 * ks2, nv))
    lib().orc_moments_nonsym(_p(bsp1), _p(bsp2), _p(gr), _p(grw), ctypes.c_int(nv),
                             ctypes.c_int(gr.shape[1]), ctypes.c_int(ks1), ctypes.c_int(ks2),
                             ctypes.c_int(k), _p(out))
    return out


def triang_sym(bsp, gr, grw, t, nv, ks, k):
    """convol.py:135-249."""
    bsp, gr, grw, t = _c(bsp), _c(gr), _c(grw), _c(t)
    kk = ks * (ks + 1) // 2
    out = np.empty((kk, kk, nv))
    lib().orc_triang_sym(_p(bsp), _p(gr), _p(grw), _p(t), ctypes.c_int(nv),
                         ctypes.c_int(gr.shape[1]), ctypes.c_int(ks), ctypes.c_int(k), _p(out))
    return out


def triang_pqpq(bsp_p, bsp_q, gr, grw, tp, tq, nv, ksp, ksq, k):
    """convol.py:387-529."""
    bsp_p, bsp_q, gr, grw, tp, tq = (_c(a) for a in (bsp_p, bsp_q, gr, grw, tp, tq))
    out = np.empty((ksp * (ksp + 1) // 2, ksq * (ksq + 1) // 2, nv))
    lib().orc_triang_pqpq(_p(bsp_p), _p(bsp_q), _p(gr), _p(grw), _p(tp), _p(tq),
                          ctypes.c_int(nv), ctypes.c_int(gr.shape[1]), ctypes.c_int(ksp),
                          ctypes.c_int(ksq), ctypes.c_int(k), _p(out))
    return out


def assemble_rkb(lo_l, hi_r, lo_r, hi_l, diag, nv, ksl, ksr, ns):
    """convol.py:323-380 (sym) and :532-572 (pqpq)."""
    arrs = [_c(a) for a in (lo_l, hi_r, lo_r, hi_l, diag)]
    out = np.empty((ns, ns, ksl, ksr))
    lib().orc_assemble_rkb(*[_p(a) for a in arrs], ctypes.c_int(nv), ctypes.c_int(ksl),
                           ctypes.c_int(ksr), ctypes.c_int(ns), _p(out))
    return out


def density_sym(p1, p2, ns, ks):
    """convol.py:688-753."""
    p1, p2 = _c(p1), _c(p2)
    out = np.empty((ns, ks))
    lib().orc_density_sym(_p(p1), ctypes.c_int(p1.size), _p(p2), ctypes.c_int(p2.size),
                          ctypes.c_int(ns), ctypes.c_int(ks), _p(out))
    return out


def density_full(p1, p2, ns):
    """convol.py:756-771."""
    p1, p2 = _c(p1), _c(p2)
    out = np.empty((ns, ns))
    lib().orc_density_full(_p(p1), ctypes.c_int(p1.size), _p(p2), ctypes.c_int(p2.size),
                           ctypes.c_int(ns), _p(out))
    return out


def convol_case1(rkb, d, ns, ksl, ksr):
    """convol.py:788-811, :829-849 (and the _t form :971-997)."""
    rkb, d = _c(rkb), _c(d)
    out = np.empty((ns, ksl))
    lib().orc_convol_case1(_p(rkb), _p(d), ctypes.c_int(ns), ctypes.c_int(ksl),
                           ctypes.c_int(ksr), _p(out))
    return out


def convol_case2(rkb, d, ns, ksl, ksr, n_left=None):
    """convol.py:852-874, :1310-1345."""
    rkb, d = _c(rkb), _c(d)
    if n_left is None:
        n_left = d.shape[0]
    out = np.empty((ns, ksr))
    lib().orc_convol_case2(_p(rkb), _p(d), ctypes.c_int(n_left), ctypes.c_int(ns),
                           ctypes.c_int(ksl), ctypes.c_int(ksr), _p(out))
    return out


def convol_case4(rkb, d, ns, ks, ksr=None, symmetric_output=False):
    """convol.py:877-948 / :1000-1058 (transposed-bank order)."""
    rkb, d = _c(rkb), _c(d)
    if ksr is None:
        ksr = ks
    out = np.empty((ns, ns))
    lib().orc_convol_case4_t(_p(rkb), _p(d), ctypes.c_int(ns), ctypes.c_int(ks),
                             ctypes.c_int(ksr), ctypes.c_int(1 if symmetric_output else 0),
                             _p(out))
    return out


def convol_case4_gather(rkb, d, ns, ks, ksr=None):
    """convol.py:877-948 (gather order)."""
    rkb, d = _c(rkb), _c(d)
    if ksr is None:
        ksr = ks
    out = np.empty((ns, ns))
    lib().orc_convol_case4(_p(rkb), _p(d), ctypes.c_int(ns), ctypes.c_int(ks),
                           ctypes.c_int(ksr), _p(out))
    return out


def sum_amb_banded(a, b, ns, ks):
    """convol.py:1348-1356."""
    a, b = _c(a), _c(b)
    return float(lib().orc_sum_amb_banded(_p(a), ctypes.c_int(a.shape[1]), _p(b),
                                          ctypes.c_int(b.shape[1]), ctypes.c_int(ns),
                                          ctypes.c_int(ks)))


def banded_to_full_add(V, vb, ns, ks, coeff):
    """convol.py:814-825 (in place on V)."""
    assert V.flags.c_contiguous and V.dtype == np.float64
    vb = _c(vb)
    lib().orc_banded_to_full_add(_p(V), _p(vb), ctypes.c_int(ns), ctypes.c_int(ks),
                                 ctypes.c_double(coeff))


def global_pair_moments(local, nv, ks, ns):
    """convol.py:1061-1073."""
    local = _c(local)
    out = np.empty((ks, ns))
    lib().orc_global_pair_moments(_p(local), ctypes.c_int(nv), ctypes.c_int(ks),
                                  ctypes.c_int(ns), _p(out))
    return out


def assemble_near_rkb(rising_left, falling_right, rising_right, falling_left, diag, nv, ksl,
                      ksr, ns):
    """convol.py:1076-1122."""
    arrs = [_c(a) for a in (rising_left, falling_right, rising_right, falling_left, diag)]
    out = np.empty((ksl, ksr, ns, ksl + ksr - 1))
    lib().orc_assemble_near_rkb(*[_p(a) for a in arrs], ctypes.c_int(nv), ctypes.c_int(ksl),
                                ctypes.c_int(ksr), ctypes.c_int(ns), _p(out))
    return out


def convol_case1_semiseparable(near, rl, fl, rr, fr, density, ns, ksl, ksr):
    """convol.py:1125-1162."""
    arrs = [_c(a) for a in (near, rl, fl, rr, fr, density)]
    out = np.empty((ns, ksl))
    lib().orc_convol_case1_semiseparable(*[_p(a) for a in arrs], ctypes.c_int(ns),
                                         ctypes.c_int(ksl), ctypes.c_int(ksr), _p(out))
    return out


def convol_case4_semiseparable(near, rl, fl, rr, fr, density, ns, ksl, ksr, symmetric):
    """convol.py:1187-1222 (symmetric) and :1250-1303 (mixed)."""
    arrs = [_c(a) for a in (near, rl, fl, rr, fr, density)]
    out = np.empty((ns, ns))
    lib().orc_convol_case4_semiseparable(*[_p(a) for a in arrs], ctypes.c_int(ns),
                                         ctypes.c_int(ksl), ctypes.c_int(ksr),
                                         ctypes.c_int(1 if symmetric else 0), _p(out))
    return out


SEMISEPARABLE_DIRECT_NS = 96          # convol.py:1411
SEMISEPARABLE_EXCHANGE_NS = 320       # convol.py:1403
SEMISEPARABLE_MIXED_EXCHANGE_NS = 320  # convol.py:1406

_COMPONENTS = {"pppp": ("p", "p"), "qqqq": ("q", "q"), "pqpq": ("p", "q"), "qpqp": ("q", "p")}


class Banks:
    """In-memory bank memo for one grid (convol.py:1478-1681, ConvolCache).

    get_rkb(k, itype) -> (rkb, ns, ks_left);  get_coulomb_factors(k, itype)
    -> (near, rising_left, falling_left, rising_right, falling_right).
    """

    def __init__(self, bsp_data):
        self.bd = bsp_data
        self._cache = {}

    def _comp(self, c):
        bd = self.bd
        if c == "p":
            return bd["bsp_p"], bd.get("ksp", bd["ks"]), bd.get("tp", bd["t"])
        return bd.get("bsp_q", bd["bsp_p"]), bd.get("ksq", bd["ks"]), bd["t"]

    def get_rkb(self, k, itype="pppp"):
        """convol.py:1507-1570 (and _build_rkb_* :532-602, :1684-1689)."""
        key = (k, itype)
        if key in self._cache:
            return self._cache[key]
        if itype not in _COMPONENTS:
            raise ValueError(f"Unknown integral type: {itype}")
        bd = self.bd
        nv, gr, grw, ns = bd["nv"], bd["gr"], bd["grw"], bd["ns"]
        left, right = _COMPONENTS[itype]
        bl, ksl, tl = self._comp(left)
        br, ksr, tr = self._comp(right)
        if left == right:
            rkd1 = moments_sym(bl, gr, grw, nv, ksl, k)
            rkd2 = moments_sym(bl, gr, grw, nv, ksl, -(k + 1))
            diag = triang_sym(bl, gr, grw, tl, nv, ksl, k)
            rkb = assemble_rkb(rkd1, rkd2, rkd1, rkd2, diag, nv, ksl, ksr, ns)
        else:
            lo_l = moments_sym(bl, gr, grw, nv, ksl, k)
            hi_r = moments_sym(br, gr, grw, nv, ksr, -(k + 1))
            lo_r = moments_sym(br, gr, grw, nv, ksr, k)
            hi_l = moments_sym(bl, gr, grw, nv, ksl, -(k + 1))
            diag = triang_pqpq(bl, br, gr, grw, tl, tr, nv, ksl, ksr, k)
            rkb = assemble_rkb(lo_l, hi_r, lo_r, hi_l, diag, nv, ksl, ksr, ns)
        self._cache[key] = (rkb, ns, ksl)
        return self._cache[key]

    def _pair_moments(self, k, c):
        """convol.py:1593-1611."""
        key = (k, c, "pm")
        if key not in self._cache:
            bd = self.bd
            bsp, ks, _ = self._comp(c)
            lr = moments_sym(bsp, bd["gr"], bd["grw"], bd["nv"], ks, k)
            lf = moments_sym(bsp, bd["gr"], bd["grw"], bd["nv"], ks, -(k + 1))
            self._cache[key] = (lr, lf, global_pair_moments(lr, bd["nv"], ks, bd["ns"]),
                                global_pair_moments(lf, bd["nv"], ks, bd["ns"]))
        return self._cache[key]

    def get_coulomb_factors(self, k, itype="pppp"):
        """convol.py:1613-1658."""
        if itype not in _COMPONENTS:
            raise ValueError(f"Unknown integral type: {itype}")
        key = (k, itype, "semi")
        if key in self._cache:
            return self._cache[key]
        bd = self.bd
        left, right = _COMPONENTS[itype]
        lrl, lfl, rl, fl = self._pair_moments(k, left)
        lrr, lfr, rr, fr = self._pair_moments(k, right)
        bl, ksl, tl = self._comp(left)
        br, ksr, tr = self._comp(right)
        if left == right:
            diag = triang_sym(bl, bd["gr"], bd["grw"], tl, bd["nv"], ksl, k)
        else:
            diag = triang_pqpq(bl, br, bd["gr"], bd["grw"], tl, tr, bd["nv"], ksl, ksr, k)
        near = assemble_near_rkb(lrl, lfr, lrr, lfl, diag, bd["nv"], ksl, ksr, bd["ns"])
        self._cache[key] = (near, rl, fl, rr, fr)
        return self._cache[key]


def compute_slater_rk_cell(Pa, Qa, Pb, Qb, Pc, Qc, Pd, Qd, bsp_data, k, banks=None,
                           force=None):
    """convol.py:1756-1882: four-block scalar R^k; semiseparable for ns >= 96.

    ``force`` = "classic" | "semiseparable" overrides the threshold (the
    reference tests do this by monkeypatching SEMISEPARABLE_DIRECT_NS).
    """
    bd = bsp_data
    if banks is None:
        banks = Banks(bd)
    ksp, nsp = bd.get("ksp", bd["ks"]), bd.get("nsp", bd["ns"])
    ksq, nsq = bd.get("ksq", bd["ks"]), bd.get("nsq", bd["ns"])
    ns = bd["ns"]
    semi = ns >= SEMISEPARABLE_DIRECT_NS if force is None else force == "semiseparable"
    if semi:
        lp, lq = density_sym(Pa, Pb, ns, ksp), density_sym(Qa, Qb, ns, ksq)
        rp, rq = density_sym(Pc, Pd, ns, ksp), density_sym(Qc, Qd, ns, ksq)
        value = 0.0
        for itype, left, right in (("pppp", lp, rp), ("qqqq", lq, rq), ("pqpq", lp, rq),
                                   ("qpqp", lq, rp)):
            f = banks.get_coulomb_factors(k, itype)
            ksl, ksr = f[1].shape[0], f[3].shape[0]
            con = convol_case1_semiseparable(*f, right, ns, ksl, ksr)
            value += sum_amb_banded(con, left, ns, ksl)
        return value
    result = 0.0
    rkb, _, _ = banks.get_rkb(k, "pppp")
    d_ab = density_sym(Pa, Pb, nsp, ksp)
    conv = convol_case2(rkb[:nsp, :nsp] if rkb.shape[0] != nsp else rkb, d_ab, nsp, ksp, ksp)
    result += sum_amb_banded(conv, density_sym(Pc, Pd, nsp, ksp), nsp, ksp)
    rkb, _, _ = banks.get_rkb(k, "qqqq")
    d_ab = density_sym(Qa, Qb, nsq, ksq)
    conv = convol_case2(rkb[:nsq, :nsq] if rkb.shape[0] != nsq else rkb, d_ab, nsq, ksq, ksq)
    result += sum_amb_banded(conv, density_sym(Qc, Qd, nsq, ksq), nsq, ksq)
    rkb, _, _ = banks.get_rkb(k, "pqpq")
    conv = convol_case2(rkb, density_sym(Pa, Pb, nsp, ksp), ns, ksp, ksq)
    result += sum_amb_banded(conv, density_sym(Qc, Qd, nsq, ksq), min(ns, nsq), ksq)
    rkb, _, _ = banks.get_rkb(k, "qpqp")
    conv = convol_case2(rkb, density_sym(Qa, Qb, nsq, ksq), ns, ksq, ksp)
    result += sum_amb_banded(conv, density_sym(Pc, Pd, nsp, ksp), min(ns, nsp), ksp)
    return result


def compute_direct_potential_cell(orb_P, orb_Q, bsp_data, k, banks=None):
    """convol.py:1694-1717."""
    bd = bsp_data
    ns = bd["ns"]
    orders = {"p": bd.get("ksp", bd["ks"]), "q": bd.get("ksq", bd["ks"])}
    if banks is None:
        banks = Banks(bd)
    dens = {c: density_sym(v, v, ns, orders[c]) for c, v in (("p", orb_P), ("q", orb_Q))}
    blocks = []
    for a in ("p", "q"):
        banded = np.zeros((ns, orders[a]))
        for b in ("p", "q"):
            rkb, _, _ = banks.get_rkb(k, a + b + a + b)
            banded += convol_case1(rkb, dens[b], ns, orders[a], orders[b])
        full = np.zeros((ns, ns))
        banded_to_full_add(full, banded, ns, orders[a], 1.0)
        blocks.append(full)
    return tuple(blocks)


def compute_exchange_matrix_cell(Pa, Qa, Pb, Qb, bsp_data, k, banks=None):
    """convol.py:1722-1751."""
    bd = bsp_data
    ns = bd["ns"]
    orders = {"p": bd.get("ksp", bd["ks"]), "q": bd.get("ksq", bd["ks"])}
    left = {"p": Pa, "q": Qa}
    right = {"p": Pb, "q": Qb}
    if banks is None:
        banks = Banks(bd)
    diagonal = Pa is Pb and Qa is Qb
    blocks = {}
    for a, b in (("p", "p"), ("q", "q"), ("p", "q"), ("q", "p")):
        if diagonal and (a, b) == ("q", "p"):
            blocks[a, b] = blocks["p", "q"].T.copy()
            continue
        rkb, _, _ = banks.get_rkb(k, a + b + a + b)
        dens = density_full(left[a], right[b], ns)
        blocks[a, b] = convol_case4(rkb, dens, ns, orders[a], orders[b],
                                    symmetric_output=diagonal and a == b)
    return tuple(blocks[x] for x in (("p", "p"), ("q", "q"), ("p", "q"), ("q", "p")))


# ---------------------------------------------------------------------------
# Angular coefficients (integer rules -> FP64 weights)
# ---------------------------------------------------------------------------

def _fact(n):
    return math.factorial(n)


def wigner_3j_exact(tj1, tj2, tj3, tm1, tm2, tm3):
    """Exact 3j symbol (twice-j integer arguments) by Racah's sum, in rationals.

    Returns (sign*sqrt(num/den)) as float.  Independent of the product's
    implementation (which uses log-factorials).
    """
    if tm1 + tm2 + tm3 != 0:
        return 0.0
    if abs(tm1) > tj1 or abs(tm2) > tj2 or abs(tm3) > tj3:
        return 0.0
    if (tj1 + tm1) % 2 or (tj2 + tm2) % 2 or (tj3 + tm3) % 2:
        return 0.0
    if tj3 < abs(tj1 - tj2) or tj3 > tj1 + tj2 or (tj1 + tj2 + tj3) % 2:
        return 0.0
    j1p, j2p, j3p = tj1, tj2, tj3
    a = (j1p + j2p - j3p) // 2
    b = (j1p - j2p + j3p) // 2
    c = (-j1p + j2p + j3p) // 2
    d = (j1p + j2p + j3p) // 2 + 1
    tri = Fraction(_fact(a) * _fact(b) * _fact(c), _fact(d))
    pre = tri * _fact((tj1 + tm1) // 2) * _fact((tj1 - tm1) // 2) * _fact((tj2 + tm2) // 2) \
        * _fact((tj2 - tm2) // 2) * _fact((tj3 + tm3) // 2) * _fact((tj3 - tm3) // 2)
    s = Fraction(0)
    for t in range(0, a + b + c + 1):
        args = [t, a - t, (tj1 - tm1) // 2 - t, (tj2 + tm2) // 2 - t,
                (tj3 - tj2 + tm1) // 2 + t, (tj3 - tj1 - tm2) // 2 + t]
        if min(args) < 0:
            continue
        den = 1
        for x in args:
            den *= _fact(x)
        s += Fraction((-1) ** t, den)
    phase = -1 if ((tj1 - tj2 - tm3) // 2) % 2 else 1
    val = phase * s
    # 3j = phase * sqrt(pre) * s ; compute in float at the end
    return float(phase) * math.copysign(1.0, float(s)) * math.sqrt(float(pre * s * s))


@lru_cache(maxsize=None)
def wigner_6j_exact(tj1, tj2, tj3, tj4, tj5, tj6):
    """Exact 6j symbol {j1 j2 j3; j4 j5 j6} (twice-j integer arguments) by
    Racah's single sum in rationals: the four triangle coefficients' squares
    and the alternating sum are exact, one square root at the end.
    Independent of the product's implementation (log-factorials)."""
    triads = ((tj1, tj2, tj3), (tj1, tj5, tj6), (tj4, tj2, tj6), (tj4, tj5, tj3))
    for a, b, c in triads:
        if (a + b + c) % 2 or not abs(a - b) <= c <= a + b:
            return 0.0
    tri2 = Fraction(1)
    for a, b, c in triads:
        tri2 *= Fraction(_fact((a + b - c) // 2) * _fact((a - b + c) // 2)
                         * _fact((b + c - a) // 2), _fact((a + b + c) // 2 + 1))
    lo = [(a + b + c) // 2 for a, b, c in triads]
    hi = [(tj1 + tj2 + tj4 + tj5) // 2, (tj2 + tj3 + tj5 + tj6) // 2,
          (tj3 + tj1 + tj6 + tj4) // 2]
    s = Fraction(0)
    for t in range(max(lo), min(hi) + 1):
        den = 1
        for x in lo:
            den *= _fact(t - x)
        for x in hi:
            den *= _fact(x - t)
        s += Fraction((-1) ** t * _fact(t + 1), den)
    if s == 0:
        return 0.0
    return math.copysign(1.0, float(s)) * math.sqrt(float(tri2 * s * s))


def l_of_kappa(kappa):
    return kappa if kappa > 0 else -kappa - 1


def twoj_of_kappa(kappa):
    return 2 * abs(kappa) - 1


def exchange_lambda(kappa_a, kappa_b, k):
    """angular/coefficients.py:153-183: (2ja+1)(2jb+1) (ja k jb; -1/2 0 1/2)^2."""
    ta, tb = twoj_of_kappa(kappa_a), twoj_of_kappa(kappa_b)
    w = wigner_3j_exact(ta, 2 * k, tb, -1, 0, 1)
    if abs(w) < 1e-15:
        return 0.0
    return (ta + 1) * (tb + 1) * w * w


def exchange_k_range(kappa_a, kappa_b):
    """angular/coefficients.py:186-210."""
    la, lb = l_of_kappa(kappa_a), l_of_kappa(kappa_b)
    ta, tb = twoj_of_kappa(kappa_a), twoj_of_kappa(kappa_b)
    return [k for k in range(abs(ta - tb) // 2, (ta + tb) // 2 + 1) if (k + la + lb) % 2 == 0]


def get_ww(occ, i, j):
    """methods/dirac_fock.py:714-733 (occupation-product model)."""
    return occ[i] * (occ[i] - 1.0) if i == j else occ[i] * occ[j]


def direct_coef(occ, kappas, i, j, k):
    """methods/dirac_fock.py:1321-1355."""
    tji, tjj = twoj_of_kappa(kappas[i]), twoj_of_kappa(kappas[j])
    if k % 2 != 0 or k > (tji + tjj) // 2:
        return 0.0
    if i == j:
        c = get_ww(occ, i, j) / 2.0
        if k == 0:
            return c
        lam = exchange_lambda(kappas[i], kappas[i], k)
        if abs(lam) < 1e-15:
            return 0.0
        return -c * lam / (tji * (tji + 1.0))
    if k != 0:
        return 0.0
    return get_ww(occ, i, j)


def exchange_coef(occ, kappas, i, j, k):
    """methods/dirac_fock.py:1357-1387."""
    if i == j:
        return 0.0
    li, lj = l_of_kappa(kappas[i]), l_of_kappa(kappas[j])
    tji, tjj = twoj_of_kappa(kappas[i]), twoj_of_kappa(kappas[j])
    if (k + li + lj) % 2 or k < abs(tji - tjj) // 2 or k > (tji + tjj) // 2:
        return 0.0
    ww = get_ww(occ, i, j)
    if abs(ww) < 1e-15:
        return 0.0
    lam = exchange_lambda(kappas[i], kappas[j], k)
    if abs(lam) < 1e-15:
        return 0.0
    return -ww * lam / ((tji + 1.0) * (tjj + 1.0))


def angular_tables(occ, kappas):
    """methods/dirac_fock.py:1070-1126: dense direct/exchange (n, n, kmax+1)."""
    n = len(kappas)
    kmax = max(abs(a) + abs(b) for a in kappas for b in kappas)
    direct = np.zeros((n, n, kmax + 1))
    exch = np.zeros((n, n, kmax + 1))
    k_for_target = []
    for i in range(n):
        ks = set()
        for j in range(n):
            for k in range(kmax + 1):
                dc = direct_coef(occ, kappas, i, j, k)
                ec = exchange_coef(occ, kappas, i, j, k)
                if abs(dc) > 1e-15 or abs(ec) > 1e-15:
                    direct[i, j, k] = dc
                    exch[i, j, k] = ec
                    ks.add(k)
        k_for_target.append(tuple(sorted(ks)))
    return direct, exch, k_for_target


def build_hf_matrix_dbsr(target_i, P_stack, Q_stack, occ, kappas, bsp_data, banks=None,
                         skip_pq_exchange=False, tables=None):
    """methods/dirac_fock.py:1128-1319 with the threshold dispatch of :1249-1301.

    P_stack/Q_stack: (n_orb, ns) coefficient stacks.  Returns PP, QQ, PQ, QP.
    """
    bd = bsp_data
    ksp, ksq, ns = bd.get("ksp", bd["ks"]), bd.get("ksq", bd["ks"]), bd["ns"]
    if banks is None:
        banks = Banks(bd)
    direct, exch, k_for_target = tables if tables is not None else angular_tables(occ, kappas)
    n_orb = len(kappas)
    hpp, hqq, hpq, hqp = (np.zeros((ns, ns)) for _ in range(4))
    for ii, jj in ((1, 1), (1, 2), (2, 1), (2, 2)):
        block = {(1, 1): "pppp", (1, 2): "pqpq", (2, 1): "qpqp", (2, 2): "qqqq"}[ii, jj]
        ks_bra = ksp if ii == 1 else ksq
        ks_ket = ksp if jj == 1 else ksq
        cjj = Q_stack if jj == 2 else P_stack
        cii = Q_stack if ii == 2 else P_stack
        for k in k_for_target[target_i]:
            a = direct[target_i, :, k]
            b = exch[target_i, :, k]
            if np.any(a):
                dd = np.zeros((ns, ks_ket))
                for j in range(n_orb):
                    c = a[j]
                    if c == 0.0:
                        continue
                    if j == target_i:
                        c = c + c
                    dd += c * density_sym(cjj[j], cjj[j], ns, ks_ket)
                if ns >= SEMISEPARABLE_DIRECT_NS:
                    v = convol_case1_semiseparable(*banks.get_coulomb_factors(k, block), dd, ns,
                                                   ks_bra, ks_ket)
                else:
                    v = convol_case1(banks.get_rkb(k, block)[0], dd, ns, ks_bra, ks_ket)
                banded_to_full_add(hpp if ii == 1 else hqq, v, ns, ks_bra, 1.0)
            if skip_pq_exchange and ii == 1 and jj == 2:
                continue
            if np.any(b):
                xx = np.zeros((ns, ns))
                for j in range(n_orb):
                    if b[j] == 0.0:
                        continue
                    xx += np.outer(b[j] * cii[j], cjj[j])
                if ii == jj and ns >= SEMISEPARABLE_EXCHANGE_NS:
                    near, rising, falling, _, _ = banks.get_coulomb_factors(k, block)
                    kk = convol_case4_semiseparable(near, rising, falling, rising, falling, xx,
                                                    ns, ks_bra, ks_bra, True)
                elif ii != jj and ns >= SEMISEPARABLE_MIXED_EXCHANGE_NS:
                    kk = convol_case4_semiseparable(*banks.get_coulomb_factors(k, block), xx,
                                                    ns, ks_bra, ks_ket, False)
                else:
                    kk = convol_case4(banks.get_rkb(k, block)[0], xx, ns, ks_bra, ks_ket,
                                      symmetric_output=(ii == jj))
                {(1, 1): hpp, (1, 2): hpq, (2, 1): hqp, (2, 2): hqq}[ii, jj][...] += kk
    hpp = 0.5 * (hpp + hpp.T)
    hqq = 0.5 * (hqq + hqq.T)
    if skip_pq_exchange:
        hpq = hqp.T.copy()
    else:
        hpq = 0.5 * (hpq + hqp.T)
        hqp = hpq.T.copy()
    norm = float(occ[target_i])
    if norm > 1e-15:
        hpp /= norm
        hqq /= norm
        hpq /= norm
        hqp /= norm
    return hpp, hqq, hpq, hqp


def two_electron_energy(P_stack, Q_stack, occ, kappas, bsp_data, banks=None):
    """methods/dirac_fock.py:1618-1661 (E_J, E_K and E_2e = E_J - E_K)."""
    if banks is None:
        banks = Banks(bsp_data)
    fk = lambda a, b, k: compute_slater_rk_cell(P_stack[a], Q_stack[a], P_stack[a], Q_stack[a],
                                                P_stack[b], Q_stack[b], P_stack[b], Q_stack[b],
                                                bsp_data, k, banks)
    gk = lambda a, b, k: compute_slater_rk_cell(P_stack[a], Q_stack[a], P_stack[b], Q_stack[b],
                                                P_stack[b], Q_stack[b], P_stack[a], Q_stack[a],
                                                bsp_data, k, banks)
    EJ = EK = 0.0
    n = len(kappas)
    for i in range(n):
        ka = kappas[i]
        tja, la = twoj_of_kappa(ka), l_of_kappa(ka)
        for j in range(i + 1):
            kb = kappas[j]
            tjb = twoj_of_kappa(kb)
            if i == j:
                c = get_ww(occ, i, i) / 2.0
                if abs(c) < 1e-15:
                    continue
                EJ += c * fk(i, i, 0)
                for k in range(1, tja + 1):
                    if (k + 2 * la) % 2:
                        continue
                    lam = exchange_lambda(ka, ka, k)
                    if lam < 1e-15:
                        continue
                    EK += c * lam / (tja * (tja + 1.0)) * fk(i, i, k)
            else:
                ww = get_ww(occ, i, j)
                if abs(ww) < 1e-15:
                    continue
                EJ += ww * fk(i, j, 0)
                for k in exchange_k_range(ka, kb):
                    lam = exchange_lambda(ka, kb, k)
                    if lam < 1e-15:
                        continue
                    EK += ww * lam / ((tja + 1.0) * (tjb + 1.0)) * gk(i, j, k)
    return EJ - EK, EJ, EK
