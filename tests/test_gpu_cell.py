"""GPU parity of the (C) DBSR cell path against the pinned CPU oracle.

Tolerances: banks, factors and contractions agree with the FP64 oracle to
1e-13 of their largest magnitude (different summation order only, several
kernels reproduce the reference order exactly); scalar R^k to 1e-12
relative; the DBSR_HF Ne two-electron energy to 2e-10 Ha (the reference's
gate) and to 1e-11 Ha of the oracle; Fock blocks to 1e-12 of their largest
element.
"""

import math

import numpy as np
import pytest

import oracle as O
from atomx_amd import bspline as B
from helpers import RK_1S, dbsr_state, physics_workspace, project_1s

pytestmark = pytest.mark.gpu


def _close(a, b, rel):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= rel * max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def dbsr():
    return dbsr_state()


@pytest.fixture(scope="module")
def kb_small():
    t = B.make_knots(40, 7, 1e-3, 30.0)
    return B.prepare_bspline_data(t, 40, 7, kinetic_balance=True)


@pytest.mark.parametrize("k", [0, 1, 3])
def test_banks_and_factors_match_oracle(kb_small, k):
    from atomx_amd.convol import ConvolCache
    bd = kb_small
    cache = ConvolCache(bd)
    ob = O.Banks(bd)
    for itype in ("pppp", "qqqq", "pqpq", "qpqp"):
        got, ns, ksl = cache.get_rkb(k, itype)
        ref, ns2, ksl2 = ob.get_rkb(k, itype)
        assert (ns, ksl) == (ns2, ksl2)
        assert _close(got, ref, 1e-13), itype
        for g, r in zip(cache.get_coulomb_factors(k, itype), ob.get_coulomb_factors(k, itype)):
            assert g.shape == r.shape
            assert _close(g, r, 1e-13), itype


def test_contractions_match_oracle(kb_small):
    from atomx_amd import device as dv
    from atomx_amd.convol import ConvolCache, density_sym_device
    bd = kb_small
    cache = ConvolCache(bd)
    ob = O.Banks(bd)
    ns = bd["ns"]
    rng = np.random.default_rng(8)
    x = rng.normal(size=(3, ns)) * np.exp(-np.arange(ns) / 9.0)
    order = {"p": bd["ksp"], "q": bd["ksq"]}
    for k in (0, 2):
        for itype in ("pppp", "qqqq", "pqpq", "qpqp"):
            l, r = itype[0], itype[1]
            rkb = ob.get_rkb(k, itype)[0]
            f = ob.get_coulomb_factors(k, itype)
            X = dv.f64(x)
            dr = density_sym_device(X, [(0, 1), (2, 2)], ns, order[r])
            got = dv.host(cache.case1(k, itype, dr, force="bank"))
            gsemi = dv.host(cache.case1(k, itype, dr, force="semiseparable"))
            for b, (p, q) in enumerate(((0, 1), (2, 2))):
                d = O.density_sym(x[p], x[q], ns, order[r])
                ref = O.convol_case1(rkb, d, ns, order[l], order[r])
                assert _close(got[b], ref, 1e-13)
                rs = O.convol_case1_semiseparable(*f, d, ns, order[l], order[r])
                assert _close(gsemi[b], rs, 1e-13)
            xx = np.stack([np.outer(x[0], x[1]), np.outer(x[2], x[2])])
            sym = l == r
            kx = dv.host(cache.case4(k, itype, dv.f64(xx), symmetric=False, force="bank"))
            ks = dv.host(cache.case4(k, itype, dv.f64(xx), symmetric=False, force="semiseparable"))
            for b in range(2):
                ref = O.convol_case4(rkb, xx[b], ns, order[l], order[r])
                assert _close(kx[b], ref, 1e-13)
                refs = O.convol_case4_semiseparable(*f, xx[b], ns, order[l], order[r], False)
                assert _close(ks[b], refs, 1e-13)
            if sym:
                xs = dv.f64(np.outer(x[2], x[2])[None])
                g1 = dv.host(cache.case4(k, itype, xs, symmetric=True, force="bank"))[0]
                g2 = dv.host(cache.case4(k, itype, xs, symmetric=True, force="semiseparable"))[0]
                r1 = O.convol_case4(rkb, np.outer(x[2], x[2]), ns, order[l], order[r], symmetric_output=True)
                r2 = O.convol_case4_semiseparable(f[0], f[1], f[2], f[1], f[2], np.outer(x[2], x[2]),
                                                  ns, order[l], order[r], True)
                assert _close(g1, r1, 1e-13) and _close(g2, r2, 1e-13)


@pytest.mark.parametrize("path", ["classic", "semiseparable"])
def test_scalar_cell_rk_matches_oracle(dbsr, path):
    from atomx_amd.convol import ConvolCache, _cell_rk_many
    from atomx_amd import device as dv
    bd, P, Q, occ, kap, ref = dbsr
    cache = ConvolCache(bd)
    ob = O.Banks(bd)
    quartets = [(P[0], Q[0], P[1], Q[1], P[2], Q[2], P[3], Q[3]),
                (P[3], Q[3], P[3], Q[3], P[1], Q[1], P[1], Q[1]),
                (P[1], Q[1], P[3], Q[3], P[3], Q[3], P[1], Q[1])]
    for k in (0, 1, 2):
        got = dv.host(_cell_rk_many(cache, k, quartets, force=path))
        for qt, g in zip(quartets, got):
            r = O.compute_slater_rk_cell(*qt, bd, k, ob, force=path)
            assert g == pytest.approx(r, rel=1e-12, abs=1e-15)


def test_projected_1s_closed_forms_gpu():
    from atomx_amd.convol import compute_slater_rk_cell, ConvolCache
    bd, t = physics_workspace(Z=1, ks=9)
    c, _ = project_1s(bd, t)
    q0 = np.zeros(bd["ns"])
    cache = ConvolCache(bd)
    for k in (0, 1, 2):
        v = compute_slater_rk_cell(c, q0, c, q0, c, q0, c, q0, bd, k, cache)
        assert v == pytest.approx(RK_1S[k], rel=1e-12)


def test_dbsr_two_electron_energy_gpu(dbsr):
    from atomx_amd.dirac_fock import FockBuilder
    bd, P, Q, occ, kap, ref = dbsr
    fb = FockBuilder(bd, kap, occ)
    e2, ej, ek = fb.energy(P, Q)
    assert e2 == pytest.approx(ref["components"]["E_2e"], abs=2e-10, rel=0)
    oe2, oej, oek = O.two_electron_energy(P, Q, occ, kap, bd)
    assert e2 == pytest.approx(oe2, abs=1e-11, rel=0)
    assert ej == pytest.approx(oej, abs=1e-11, rel=0)


@pytest.mark.parametrize("skip_pq", [True, False])
def test_dbsr_fock_blocks_match_oracle(dbsr, skip_pq):
    from atomx_amd.dirac_fock import FockBuilder
    bd, P, Q, occ, kap, ref = dbsr
    fb = FockBuilder(bd, kap, occ)
    H = fb.build_all(P, Q, skip_pq_exchange=skip_pq)
    banks = O.Banks(bd)
    tables = O.angular_tables(occ, kap)
    for t in range(len(kap)):
        r = O.build_hf_matrix_dbsr(t, P, Q, occ, kap, bd, banks, skip_pq_exchange=skip_pq,
                                   tables=tables)
        for g, rr in zip(H[t], r):
            assert _close(g, rr, 1e-12), t


def test_direct_and_exchange_blocks_gpu(kb_small):
    from atomx_amd.convol import (ConvolCache, compute_direct_potential_cell,
                                  compute_exchange_matrix_cell, compute_slater_rk_cell)
    bd = kb_small
    ns = bd["ns"]
    rng = np.random.default_rng(713)
    a, b, c = [tuple(rng.normal(size=ns) for _ in range(2)) for _ in range(3)]
    cache = ConvolCache(bd)
    ob = O.Banks(bd)
    for rank in (0, 1, 2):
        jpp, jqq = compute_direct_potential_cell(*c, bd, rank, cache)
        rpp, rqq = O.compute_direct_potential_cell(*c, bd, rank, ob)
        assert _close(jpp, rpp, 1e-13) and _close(jqq, rqq, 1e-13)
        got = compute_exchange_matrix_cell(*c, *c, bd, rank, cache)
        refx = O.compute_exchange_matrix_cell(*c, *c, bd, rank, ob)
        for g, r in zip(got, refx):
            assert _close(g, r, 1e-13)
        # the blocks contract to the scalar cell R^k (reference test_convol.py:409-432)
        direct = a[0] @ jpp @ b[0] + a[1] @ jqq @ b[1]
        ej = compute_slater_rk_cell(*a, *b, *c, *c, bd, rank, cache)
        assert direct == pytest.approx(ej, rel=3e-13, abs=1e-11)


def test_batch_and_bilinear_match_scalar(kb_small):
    from atomx_amd.convol import (ConvolCache, build_slater_rk_bilinear,
                                  compute_slater_rk_cell_batch, eval_slater_rk_bilinear)
    bd = kb_small
    ns = bd["ns"]
    rng = np.random.default_rng(21)
    orbs = {n: (rng.normal(size=ns) * np.exp(-np.arange(ns) / 8.0),
                0.05 * rng.normal(size=ns) * np.exp(-np.arange(ns) / 8.0)) for n in "abcd"}
    cache = ConvolCache(bd)
    ob = O.Banks(bd)
    keys = [("a", "b", "c", "d", 0), ("a", "a", "b", "b", 2), ("c", "d", "d", "c", 1)]
    got = compute_slater_rk_cell_batch(orbs, keys, bd, cache)
    for (a, b, c, d, k), g in zip(keys, got):
        r = O.compute_slater_rk_cell(*orbs[a], *orbs[b], *orbs[c], *orbs[d], bd, k, ob,
                                     force="classic")
        assert g == pytest.approx(r, rel=1e-12)
    M = build_slater_rk_bilinear(*orbs["a"], *orbs["b"], bd, 1, cache)
    v = eval_slater_rk_bilinear(M, *orbs["c"], *orbs["d"])
    r = O.compute_slater_rk_cell(*orbs["a"], *orbs["c"], *orbs["b"], *orbs["d"], bd, 1, ob,
                                 force="classic")
    assert v == pytest.approx(r, rel=1e-12)


@pytest.mark.parametrize("ks,balanced,p_power,q_power,ranks", [
    (4, False, 1, 1, (0, 1, 6)),
    (6, True, 4, 5, (6,)),
])
def test_signed_monomial_components_gpu(ks, balanced, p_power, q_power, ranks):
    """Closed-form cell integrals through the device banks and factors."""
    from atomx_amd import device as dv
    from atomx_amd.convol import ConvolCache, density_sym_device, sum_amb_device
    radius = 0.7
    knots = np.r_[np.zeros(ks), np.full(ks, radius)]
    data = B.prepare_bspline_data(knots, ks, ks, n_gauss=16, kinetic_balance=balanced)
    orders = data["ksp"], data["ksq"]
    coefficients = []
    for order, power in zip(orders, (p_power, q_power)):
        vector = np.zeros(data["ns"])
        for i in range(power, order):
            vector[i] = radius ** power * math.comb(i, power) / math.comb(order - 1, power)
        coefficients.append(vector)
    waves = [(s * coefficients[0], s * coefficients[1]) for s in (1.0, -2.0, 3.0, 4.0)]
    ns = data["ns"]
    cache = ConvolCache(data)
    for rank in ranks:
        for left, right in ((0, 0), (1, 1), (0, 1), (1, 0)):
            comp = ("p", "q")[left] + ("p", "q")[right]
            comp += comp
            tgt = density_sym_device(dv.f64(np.stack([waves[0][left], waves[1][left]])),
                                     [(0, 1)], ns, orders[left])
            src = density_sym_device(dv.f64(np.stack([waves[2][right], waves[3][right]])),
                                     [(0, 1)], ns, orders[right])
            a = 2 * (p_power, q_power)[left]
            b = 2 * (p_power, q_power)[right]
            expected = -24 * radius ** (a + b + 1) / (a + b + 1) * (1 / (a + rank + 1) + 1 / (b + rank + 1))
            for force in ("bank", "semiseparable"):
                pot = cache.case1(rank, comp, src, force=force)
                v = float(dv.host(sum_amb_device(tgt, pot, ns, orders[left]))[0])
                assert v == pytest.approx(expected, rel=2e-13, abs=0.0)
