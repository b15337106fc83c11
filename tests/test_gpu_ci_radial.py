"""GPU: the CI radial evaluator, pair potentials, memoised F^k/G^k/R^k and
the Breit component-resolved R^L against the oracle (scalar cell R^k,
Poisson-form Slater integral) and closed forms (Dirac 1s energy)."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _workspace(ns, ks, kb, r_max=20.0):
    from atomx_amd.bspline import make_knots, prepare_bspline_data
    return prepare_bspline_data(make_knots(ns, ks, 0.002, r_max), ns, ks, kinetic_balance=kb)


def _orbitals(bd, seed):
    rng = np.random.default_rng(seed)
    ns = bd["ns"]
    keys = [(1, -1), (2, -1), (2, 1), (2, -2), (3, 2)]
    return {k: (rng.normal(size=ns), rng.normal(size=ns)) for k in keys}


@pytest.mark.parametrize("ns,ks,kb", [(24, 5, True), (24, 6, False), (100, 6, True)])
def test_evaluator_rk_keys_match_cell_rk(ns, ks, kb):
    from atomx_amd.ci_radial import RadialEvaluator
    bd = _workspace(ns, ks, kb)
    orbs = _orbitals(bd, 11)
    ev = RadialEvaluator(orbs, bd, 10, bd["ns"])
    banks = O.Banks(bd)
    names = sorted(orbs)
    rng = np.random.default_rng(4)
    keys = []
    for _ in range(40):
        a, c, b, d = (names[i] for i in rng.integers(0, len(names), 4))
        keys.append(("r", int(rng.integers(0, 5)), a, c, b, d))
    print(f"ns={ns} ks={ks} kb={kb}: {len(keys)} keys", flush=True)
    got = ev.evaluate_many(keys)
    for key, v in zip(keys, got):
        _, k, a, c, b, d = key
        want = O.compute_slater_rk_cell(*orbs[a], *orbs[c], *orbs[b], *orbs[d], bd, k, banks)
        assert v == pytest.approx(want, rel=1e-12, abs=1e-12 * max(1.0, abs(want))), key
    # memoised single-key calls return the batched values
    assert ev(keys[3]) == got[3]
    assert ev.hits >= 1


def test_evaluator_one_body_key_is_the_dirac_eigenvalue():
    """For an eigenvector of the one-body Dirac problem, <a|h|a>/<a|a> is its
    eigenvalue; the 1s root of Z = 10 (point nucleus) is the Dirac energy
    c^2 (sqrt(1 - (Z/c)^2) - 1) to the basis accuracy."""
    from atomx_amd import dirac as D
    from atomx_amd.ci_radial import MissingOrbitalError, RadialEvaluator
    Z = 10
    bd = _workspace(40, 7, False, r_max=30.0)
    ns = bd["ns"]
    H, S = D.dirac_matrices(-1, Z, bd)
    e, v = D.solve_dirac(H, S, ns, bd["nsp"], bd["nsq"])
    orbs = {(1, -1): (v[:ns, 0], v[ns:, 0]), (2, -1): (v[:ns, 1], v[ns:, 1])}
    ev = RadialEvaluator(orbs, bd, Z, ns)
    assert ev(("h", 1, -1, 1)) == pytest.approx(e[0], rel=1e-10)
    assert ev(("h", 2, -1, 2)) == pytest.approx(e[1], rel=1e-10)
    assert abs(ev(("h", 1, -1, 2))) < 1e-9 * abs(e[0])
    exact = D.C_SQ * (np.sqrt(1.0 - (Z / D.C) ** 2) - 1.0)
    assert e[0] == pytest.approx(exact, rel=1e-9)
    with pytest.raises(MissingOrbitalError):
        ev(("h", 3, -1, 3))
    with pytest.raises(ValueError, match="capacity"):
        RadialEvaluator(orbs, bd, Z, ns, pair_potential_cache_size=0)


def test_pair_potential_and_memo_functions():
    from atomx_amd import ci_radial as CR
    from atomx_amd.convol import ConvolCache
    bd = _workspace(24, 5, True)
    orbs = _orbitals(bd, 3)
    banks = O.Banks(bd)
    ns, ksp, ksq = bd["ns"], bd["ksp"], bd["ksq"]
    (Pa, Qa), (Pb, Qb) = orbs[(1, -1)], orbs[(2, 1)]
    cache = ConvolCache(bd)
    src = (O.density_sym(Pb, Pb, ns, ksp), O.density_sym(Qb, Qb, ns, ksq))
    vp, vq = CR.bspline_pair_potential(cache, ns, (ksp, ksq), 2, src)
    tgt = (O.density_sym(Pa, Pa, ns, ksp), O.density_sym(Qa, Qa, ns, ksq))
    val = O.sum_amb_banded(vp, tgt[0], ns, ksp) + O.sum_amb_banded(vq, tgt[1], ns, ksq)
    want = O.compute_slater_rk_cell(Pa, Qa, Pa, Qa, Pb, Qb, Pb, Qb, bd, 2, banks)
    assert val == pytest.approx(want, rel=1e-12)
    CR.clear_memo()
    f = CR.fk(Pa, Qa, Pb, Qb, bd, 2)
    assert f == pytest.approx(want, rel=1e-12)
    assert CR.fk(Pb, Qb, Pa, Qa, bd, 2) == f          # symmetric key
    g = CR.gk(Pa, Qa, Pb, Qb, bd, 1)
    assert g == pytest.approx(O.compute_slater_rk_cell(Pa, Qa, Pb, Qb, Pb, Qb, Pa, Qa, bd, 1,
                                                        banks), rel=1e-12)
    r = CR.rk(Pa, Qa, Pb, Qb, Pa, Qa, Pb, Qb, bd, 3)
    assert r == pytest.approx(O.compute_slater_rk_cell(Pa, Qa, Pb, Qb, Pa, Qa, Pb, Qb, bd, 3,
                                                        banks), rel=1e-12)


def test_breit_component_integrals_match_poisson_form():
    from atomx_amd.ci_radial import breit_rk, breit_rk_many
    bd = _workspace(24, 5, True)
    orbs = _orbitals(bd, 8)
    gr, grw = bd["gr"], bd["grw"]
    nv, ng = gr.shape
    tables = {"P": (bd["bsp_p"], bd["ksp"], bd["nsp"]), "Q": (bd["bsp_q"], bd["ksq"], bd["nsq"])}

    def on_grid(f, c):
        tab, ks, ns = tables[c]
        return O.eval_orbital_on_grid(f[:ns], tab, nv, ks, ns)

    vals = [v for pair in orbs.values() for v in pair]
    rng = np.random.default_rng(2)
    items = []
    for _ in range(12):
        f = [vals[i] for i in rng.integers(0, len(vals), 4)]
        comps = tuple(rng.choice(["P", "Q"], 4))
        items.append((*f, comps, int(rng.integers(0, 4))))
    got = breit_rk_many(items, bd)
    for it, v in zip(items, got):
        f1, f2, f3, f4, comps, L = it
        rho1 = on_grid(f1, comps[0]) * on_grid(f3, comps[2])
        rho2 = on_grid(f2, comps[1]) * on_grid(f4, comps[3])
        want = 0.5 * (O.slater_rk_poisson(rho1, rho2, gr, grw, nv, ng, L)
                      + O.slater_rk_poisson(rho2, rho1, gr, grw, nv, ng, L))
        assert v == pytest.approx(want, rel=1e-12, abs=1e-14), it[4:]
    f1, f2, f3, f4, comps, L = items[0]
    swapped = (comps[1], comps[0], comps[3], comps[2])
    assert breit_rk(f1, f2, f3, f4, comps, bd, L) == breit_rk(f2, f1, f4, f3, swapped, bd, L)
    with pytest.raises(ValueError, match="component"):
        breit_rk(f1, f2, f3, f4, ("P", "X", "P", "P"), bd, 0)
