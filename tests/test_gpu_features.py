"""GPU: the auxiliary-node action and the pair contraction against the
oracle (loop restatement of _apply, operator-form contraction, scalar cell
R^k) — the checks of the reference's tests/test_cell_coulomb_features.py and
tests/test_cell_pair_contraction.py, through the C ABI kernels."""

import itertools

import numpy as np
import pytest

import oracle as O
from oracle import features as OF

pytestmark = pytest.mark.gpu

TYPES = ["native", "polynomial"]


def _features(kind, bd, **kw):
    from atomx_amd.features import CellCoulombFeatures, PolynomialCellCoulombFeatures
    return (CellCoulombFeatures if kind == "native" else PolynomialCellCoulombFeatures)(bd, **kw)


def _bd(ns, ks, r_max, kb):
    from atomx_amd.bspline import make_knots, prepare_bspline_data
    return prepare_bspline_data(make_knots(ns, ks, 0.02, r_max), ns, ks, kinetic_balance=kb)


def _host_sample(f, P, Q):
    out = []
    for c, coef in enumerate((P, Q)):
        tab = f.tables[c]
        nv, _, order = tab.shape
        out.append(O.eval_orbital_on_grid(coef, tab, nv, order, f.coefficient_sizes[c]).ravel())
    return np.concatenate(out)


@pytest.mark.parametrize("kind", TYPES)
@pytest.mark.parametrize("kb", [False, True])
def test_sample_and_action_match_oracle(kind, kb):
    ns, ks = 16, 5
    bd = _bd(ns, ks, 4.0, kb)
    f = _features(kind, bd)
    rng = np.random.default_rng(7)
    P, Q = rng.normal(size=ns), rng.normal(size=ns)
    s = f.sample(P, Q)
    ref = _host_sample(f, P, Q)
    assert np.max(np.abs(s - ref)) <= 1e-14 * np.max(np.abs(ref))
    rho = rng.normal(size=f.radii.size)
    for rank in (0, 1, 2, 5):
        print(f"{kind} kb={kb} rank={rank}", flush=True)
        got = f.apply_density(rho, rank)
        want = OF.apply_features(f, rho, rank)
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want)), rank
    # many columns at once, complex c-product = real and imaginary parts
    cols = rng.normal(size=(f.radii.size, 5)) + 1j * rng.normal(size=(f.radii.size, 5))
    got = f.apply_density(cols, 3)
    for j in range(5):
        want = OF.apply_features(f, cols[:, j], 3)
        assert np.max(np.abs(got[:, j] - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("kind", TYPES)
@pytest.mark.parametrize("kb", [False, True])
@pytest.mark.parametrize("rank", [0, 1, 2, 3, 7])
def test_integral_matches_cell_rk(kind, kb, rank):
    ns, ks = 16, 5
    bd = _bd(ns, ks, 4.0, kb)
    f = _features(kind, bd)
    rng = np.random.default_rng(781)
    orbitals = [(rng.normal(size=ns), rng.normal(size=ns)) for _ in range(4)]
    expected = O.compute_slater_rk_cell(*(v for o in orbitals for v in o), bd, rank)
    assert f.integral(orbitals, rank) == pytest.approx(expected, abs=1e-11, rel=1e-13)
    a, b, c, d = (f.sample(*o) for o in orbitals)
    assert np.dot(a * b, f.apply_density(c * d, rank)) == pytest.approx(
        np.dot(c * d, f.apply_density(a * b, rank)), abs=1e-11)


@pytest.mark.parametrize("kind", TYPES)
def test_monomial_double_integral_and_complex_linearity(kind):
    ns = ks = 4
    radius = 0.7
    from atomx_amd.bspline import prepare_bspline_data
    bd = prepare_bspline_data(np.r_[np.zeros(ks), np.full(ks, radius)], ns, ks, n_gauss=5)
    f = _features(kind, bd)
    p, q = np.linspace(0, radius, ns), np.zeros(ns)
    for rank in (0, 1, 3):
        expected = radius ** 5 / 5 * (2 / (3 + rank))
        assert f.integral([(p, q)] * 4, rank) == pytest.approx(expected, rel=2e-13)
    g = _features(kind, _bd(12, 4, 3.0, True), rank_cache_size=1)
    rng = np.random.default_rng(213)
    a, b = rng.normal(size=len(g.radii)), rng.normal(size=len(g.radii))
    act = g.apply_density(a + 1j * b, 2)
    np.testing.assert_allclose(act, g.apply_density(a, 2) + 1j * g.apply_density(b, 2),
                               atol=1e-12)
    before = g.apply_density(a, 2)
    np.testing.assert_array_equal(before, g.apply_density(a, 2))


@pytest.mark.parametrize("kb", [False, True])
@pytest.mark.parametrize("rank", [0, 3, 7])
def test_pair_contraction_matches_quartet_sum(kb, rank):
    from atomx_amd.features import CellCoulombFeatures, CellPairContraction
    ns, ks = 14, 5
    bd = _bd(ns, ks, 4.0, kb)
    f = CellCoulombFeatures(bd)
    banks = O.Banks(bd)
    rng = np.random.default_rng(381)
    waves = [[(rng.normal(size=ns), rng.normal(size=ns)) for _ in range(n)]
             for n in (3, 2, 2, 4)]
    samples = [np.column_stack([f.sample(*a) for a in g]) for g in waves]
    rho = rng.normal(size=(2, 2))
    expected = np.zeros((3, 4))
    for m, r, s, n in itertools.product(range(3), range(2), range(2), range(4)):
        q = (waves[0][m], waves[1][r], waves[2][s], waves[3][n])
        expected[m, n] += rho[r, s] * O.compute_slater_rk_cell(
            *(v for a in q for v in a), bd, rank, banks)
    a = CellPairContraction(f, *samples, rank, edge_batch_size=17)
    b = CellPairContraction(f, *samples, rank, edge_batch_size=4096)
    np.testing.assert_allclose(a(rho), expected, atol=2e-11, rtol=3e-13)
    np.testing.assert_allclose(b(rho), expected, atol=2e-11, rtol=3e-13)
    tiny = np.array([[0.0, 1e-14], [-2e-14, 0.0]])
    np.testing.assert_allclose(a(rho + 1j * tiny), a(rho) + 1j * a(tiny), atol=2e-11,
                               rtol=3e-13)
    old = a(rho)
    samples[0][:] *= -2
    np.testing.assert_array_equal(a(rho), old)


def test_pair_contraction_polynomial_and_complex_operands():
    from atomx_amd.bspline import prepare_bspline_data
    from atomx_amd.features import (CellCoulombFeatures, CellPairContraction,
                                    PolynomialCellCoulombFeatures)
    ns = ks = 4
    bd = prepare_bspline_data(np.r_[np.zeros(ks), np.full(ks, 0.7)], ns, ks, n_gauss=5)
    f = CellCoulombFeatures(bd)
    z = f.sample(np.linspace(0, 0.7, ns), np.zeros(ns))[:, None]
    a = CellPairContraction(f, z * (1 + 1j), z * 2, z * 3, z * (1 - 2j), 1)
    expected = 0.7 ** 5 / 5 * (2 / 4) * (1 + 1j) * 2 * 3 * (1 - 2j)
    assert a(np.ones((1, 1)))[0, 0] == pytest.approx(expected, rel=3e-13)
    with pytest.raises(MemoryError, match="complete cell-pair"):
        CellPairContraction(f, z, z, z, z, 0, max_work_bytes=1)
    with pytest.raises(ValueError, match="integer"):
        CellPairContraction(f, z, z, z, z, 0, edge_batch_size=True)
    with pytest.raises(ValueError, match="density matrix"):
        a(np.zeros((1, 2)))
    with pytest.raises(ValueError, match="column"):
        CellPairContraction(f, z[:-1], z, z, z, 0)
    # the polynomial coordinates against the oracle's operator form
    g = PolynomialCellCoulombFeatures(_bd(12, 4, 3.0, True))
    rng = np.random.default_rng(9)
    ops = [rng.normal(size=(g.radii.size, k)) for k in (2, 3, 2, 2)]
    ops[1] = ops[1] + 1j * rng.normal(size=ops[1].shape)
    rho = rng.normal(size=(3, 2))
    got = CellPairContraction(g, *ops, 4)(rho)
    want = OF.pair_contraction(g, *ops, 4, rho)
    np.testing.assert_allclose(got, want, atol=1e-12 * np.max(np.abs(want)), rtol=1e-12)
