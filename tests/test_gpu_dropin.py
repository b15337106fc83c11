"""GPU: the reference's private cell kernels by name and argument order
(numerics/bspline/convol.py:609-1373, as dirac_fock.py:1169-1181 and
physics/csf/matrix.py:24-34 import them) and the _build_hf_matrix_dbsr
drop-in (methods/dirac_fock.py:1128-1319) fed the reference's orbitals dict.

Each private kernel is compared with the oracle's restatement of the same
function at 1e-13 of the largest magnitude (in-place kernels modify the
caller's array exactly once); the Fock blocks at 1e-12 of each block's
largest entry against oracle.build_hf_matrix_dbsr."""

from types import SimpleNamespace

import numpy as np
import pytest

import oracle as O
from atomx_amd import bspline as B

pytestmark = pytest.mark.gpu


def _close(a, b, rel=1e-13):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= rel * max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def ws():
    t = B.make_knots(44, 7, 1e-3, 30.0)
    return B.prepare_bspline_data(t, 44, 7, kinetic_balance=True)


def test_private_density_kernels(ws):
    from atomx_amd import convol as C
    ns = ws["ns"]
    rng = np.random.default_rng(3)
    p1, p2 = rng.normal(size=ns - 1), rng.normal(size=ns)
    for ks in (7, 8):
        assert _close(C._density_sym(p1, p2, ns, ks), O.density_sym(p1, p2, ns, ks))
    assert _close(C._density_full(p1, p2, ns), O.density_full(p1, p2, ns))
    dd0 = rng.normal(size=(ns, 7))
    dd = dd0.copy()
    C._accumulate_density_sym(dd, p1, p2, ns, 7, 0.75)
    assert _close(dd, dd0 + 0.75 * O.density_sym(p1, p2, ns, 7))
    xx0 = rng.normal(size=(ns, ns))
    xx = xx0.copy()
    C._accumulate_density_full(xx, p1, p2, ns, -1.5)
    assert _close(xx, xx0 - 1.5 * O.density_full(p1, p2, ns))
    stack = rng.normal(size=(5, ns))
    coef = np.array([0.5, 0.0, -1.25, 2.0, 0.3])
    dd = dd0.copy()
    C._accumulate_direct_density_batch(dd, stack, coef, 2, ns, 7)
    want = dd0.copy()
    for j in range(5):
        c = coef[j] * (2.0 if j == 2 else 1.0)
        want += c * O.density_sym(stack[j], stack[j], ns, 7)
    assert _close(dd, want)
    stack2 = rng.normal(size=(5, ns))
    xx = xx0.copy()
    C._accumulate_exchange_density_batch(xx, stack, stack2, coef, ns)
    assert _close(xx, xx0 + sum(c * np.outer(a, b) for c, a, b in zip(coef, stack, stack2)))


def test_private_contractions(ws):
    from atomx_amd import convol as C
    bd = ws
    ns = bd["ns"]
    cache = C.ConvolCache(bd)
    ob = O.Banks(bd)
    order = {"p": bd["ksp"], "q": bd["ksq"]}
    rng = np.random.default_rng(4)
    x = rng.normal(size=(3, ns)) * np.exp(-np.arange(ns) / 9.0)
    for k in (0, 2):
        for itype in ("pppp", "qqqq", "pqpq", "qpqp"):
            ksl, ksr = order[itype[0]], order[itype[1]]
            rkb = ob.get_rkb(k, itype)[0]
            f = ob.get_coulomb_factors(k, itype)
            rkbT_res = cache.get_rkb_t(k, itype)[0]          # resident, read-only
            rkbT_host = np.array(rkbT_res)                    # plain numpy copy
            fres = cache.get_coulomb_factors(k, itype)
            d = O.density_sym(x[0], x[1], ns, ksr)
            ref1 = O.convol_case1(rkb, d, ns, ksl, ksr)
            assert _close(C._convol_case1_t(rkbT_res, d, ns, ksl, ksr), ref1), itype
            assert _close(C._convol_case1_t(rkbT_host, d, ns, ksl, ksr), ref1), itype
            assert _close(C._convol_case1_mixed(rkb, d, ns, ksl, ksr), ref1), itype
            assert _close(C._convol_case1_semiseparable_t(*fres, d, ns, ksl, ksr),
                          O.convol_case1_semiseparable(*f, d, ns, ksl, ksr)), itype
            dl = O.density_sym(x[0], x[2], ns, ksl)[: ns - 1]
            assert _close(C._convol_case2_mixed(rkb, dl, ns, ksl, ksr),
                          O.convol_case2(rkb, dl, ns, ksl, ksr, n_left=ns - 1)), itype
            xx = np.outer(x[0], x[2])
            ref4 = O.convol_case4(rkb, xx, ns, ksl, ksr)
            assert _close(C._convol_case4_t(rkbT_res, xx, ns, ksl, ksr), ref4), itype
            assert _close(C._convol_case4(rkb, xx, ns, ksl, ksr), ref4), itype
            if itype in ("pppp", "qqqq"):
                xs = np.outer(x[2], x[2])
                assert _close(C._convol_case4_t(rkbT_res, xs, ns, ksl, symmetric_output=True),
                              O.convol_case4(rkb, xs, ns, ksl, ksl, symmetric_output=True))
                assert _close(C._convol_case4_semiseparable_t(fres[0], fres[1], fres[2], xs, ns,
                                                              ksl),
                              O.convol_case4_semiseparable(f[0], f[1], f[2], f[1], f[2], xs,
                                                           ns, ksl, ksl, True))
                assert _close(C._convol_case1(rkb, d, ns, ksl), ref1)
            else:
                assert _close(C._convol_case4_semiseparable_mixed_t(*fres, xx, ns, ksl, ksr),
                              O.convol_case4_semiseparable(*f, xx, ns, ksl, ksr, False))
    with pytest.raises(ValueError):
        rkbT_res[0, 0, 0, 0] = 1.0  # banks are handed out read-only


def test_private_traces_and_band_add(ws):
    from atomx_amd import convol as C
    ns = ws["ns"]
    rng = np.random.default_rng(5)
    a, b = rng.normal(size=(ns, 7)), rng.normal(size=(ns, 7))
    assert C._sum_amb_banded(a, b, ns, 7) == pytest.approx(O.sum_amb_banded(a, b, ns, 7),
                                                           rel=1e-13)
    A, Bm = rng.normal(size=(ns, ns)), rng.normal(size=(ns, ns))
    assert C._sum_amb_full(A, Bm, ns) == pytest.approx(float(np.sum(A * Bm)), rel=1e-12)
    V0 = rng.normal(size=(ns, ns))
    V, W = V0.copy(), V0.copy()
    C._banded_to_full_add(V, a, ns, 7, 0.5)
    O.banded_to_full_add(W, a, ns, 7, 0.5)
    assert _close(V, W)


@pytest.mark.parametrize("skip", [False, True])
def test_build_hf_matrix_dbsr_dropin_with_orbitals_dict(skip):
    from atomx_amd.dirac_fock import build_hf_matrix_dbsr
    from atomx_amd.workloads import fe_ion_workload
    ion = fe_ion_workload(seed=2, ns_range=(60, 110))[4]
    bd, kap, occ = ion["bsp_data"], ion["kappas"], ion["occupations"]
    direct, exchange, k_for_target = O.angular_tables(occ, kap)
    shells = [SimpleNamespace(kappa=k, occupation=q) for k, q in zip(kap, occ)]
    solver = SimpleNamespace(bsp_data=bd, subshells=shells, _angular_direct=direct,
                             _angular_exchange=exchange, _angular_k_for_target=k_for_target,
                             _mean_field_divisor=lambda i: float(occ[i]))
    orbitals = {}
    for j, kv in enumerate(kap):  # kappa -> [(P, Q, occ), ...] in subshell order
        orbitals.setdefault(kv, []).append((ion["P"][j], ion["Q"][j], occ[j]))
    banks = O.Banks(bd)
    tables = O.angular_tables(occ, kap)
    for rep in range(2):  # the second call reuses the cached build on new orbitals
        scale = 1.0 + 0.1 * rep
        orb = {kv: [(P * scale, Q, q) for P, Q, q in lst] for kv, lst in orbitals.items()}
        for t in range(len(kap)):
            got = build_hf_matrix_dbsr(solver, t, orb, skip_pq_exchange=skip)
            ref = O.build_hf_matrix_dbsr(t, ion["P"] * scale, ion["Q"], occ, kap, bd, banks,
                                         skip_pq_exchange=skip, tables=tables)
            for name, g, r in zip(("PP", "QQ", "PQ", "QP"), got, ref):
                assert _close(g, r, 1e-12), (rep, t, name)
