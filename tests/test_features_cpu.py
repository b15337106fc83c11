"""CPU pins of the auxiliary-node geometry (host set-up of
atomx_amd.features) through the oracle's loop restatement of the action,
against the oracle's scalar cell R^k — the checks of the reference's
tests/test_cell_coulomb_features.py:17-56,93-110 — plus the oracle's
operator-form pair contraction against the brute-force quartet sum
(tests/test_cell_pair_contraction.py:15-40)."""

import itertools

import numpy as np
import pytest

import oracle as O
from oracle import features as OF
from atomx_amd.bspline import make_knots, prepare_bspline_data
from atomx_amd.features import (CellCoulombFeatures, PolynomialCellCoulombFeatures,
                                cardinal_values)

TYPES = [CellCoulombFeatures, PolynomialCellCoulombFeatures]


def host_sample(f, P, Q):
    """Samples at the auxiliary nodes by the oracle's B-spline evaluation."""
    out = []
    for c, coef in enumerate((P, Q)):
        tab = f.tables[c]
        nv, _, order = tab.shape
        out.append(O.eval_orbital_on_grid(coef, tab, nv, order, f.coefficient_sizes[c]).ravel())
    return np.concatenate(out)


def host_integral(f, orbitals, rank):
    a, b, c, d = (host_sample(f, *o) for o in orbitals)
    return float(np.dot(a * b, OF.apply_features(f, c * d, rank)))


@pytest.mark.parametrize("kinetic_balance", [False, True])
@pytest.mark.parametrize("rank", [0, 1, 2, 3, 7])
@pytest.mark.parametrize("ftype", TYPES)
def test_signed_quartets_match_cell_integral(kinetic_balance, rank, ftype):
    ns, ks = 16, 5
    bd = prepare_bspline_data(make_knots(ns, ks, 0.02, 4.0), ns, ks,
                              kinetic_balance=kinetic_balance)
    f = ftype(bd)
    banks = O.Banks(bd)
    rng = np.random.default_rng(781)
    orbitals = [(rng.normal(size=ns), rng.normal(size=ns)) for _ in range(4)]
    expected = O.compute_slater_rk_cell(*(v for o in orbitals for v in o), bd, rank, banks)
    got = host_integral(f, orbitals, rank)
    assert got == pytest.approx(expected, abs=1e-11, rel=1e-13)
    parts = []
    for left, right in itertools.product(range(2), repeat=2):
        sel = [(p if (left if i < 2 else right) == 0 else np.zeros(ns),
                q if (left if i < 2 else right) == 1 else np.zeros(ns))
               for i, (p, q) in enumerate(orbitals)]
        ref = O.compute_slater_rk_cell(*(v for o in sel for v in o), bd, rank, banks)
        val = host_integral(f, sel, rank)
        assert val == pytest.approx(ref, abs=1e-11, rel=1e-13)
        parts.append(val)
    assert sum(parts) == pytest.approx(got, abs=1e-11, rel=1e-13)
    a, b, c, d = (host_sample(f, *o) for o in orbitals)
    assert np.dot(a * b, OF.apply_features(f, c * d, rank)) == pytest.approx(
        np.dot(c * d, OF.apply_features(f, a * b, rank)), abs=1e-11)


@pytest.mark.parametrize("ftype", TYPES)
def test_origin_cell_monomial_double_integral(ftype):
    """P(r) = r on one cell from the origin:
    R^k = R^5/5 * 2/(3+k) (the double integral of r^2 r'^2 r_<^k/r_>^(k+1))."""
    ns = ks = 4
    radius = 0.7
    t = np.r_[np.zeros(ks), np.full(ks, radius)]
    bd = prepare_bspline_data(t, ns, ks, n_gauss=5)
    f = ftype(bd)
    p, q = np.linspace(0, radius, ns), np.zeros(ns)
    for rank in (0, 1, 3):
        expected = radius ** 5 / 5 * (2 / (3 + rank))
        assert host_integral(f, [(p, q)] * 4, rank) == pytest.approx(expected, rel=2e-13)


def test_cardinal_polynomials_reproduce_polynomials():
    nodes = (np.polynomial.legendre.leggauss(7)[0] + 1) / 2
    x = np.linspace(0, 1, 23)
    L = cardinal_values(x, nodes)
    for deg in range(7):
        assert np.allclose(L @ nodes ** deg, x ** deg, atol=1e-13, rtol=0)
    assert np.array_equal(cardinal_values(nodes[2:3], nodes)[0], np.eye(7)[2])


def test_geometry_guards_and_read_only_arrays():
    ns, ks = 12, 4
    bd = prepare_bspline_data(make_knots(ns, ks, 0.02, 3.0), ns, ks)
    with pytest.raises(MemoryError, match="complete auxiliary"):
        CellCoulombFeatures(bd, memory_bytes=1)
    with pytest.raises(ValueError, match="integer"):
        CellCoulombFeatures(bd, rank_cache_size=True)
    f = CellCoulombFeatures(bd, rank_cache_size=1)
    with pytest.raises(ValueError, match="multipole"):
        f.rank_arrays(-1)
    with pytest.raises(MemoryError, match="rank-aware"):
        f.rank_arrays(100_000_000)
    with pytest.raises(ValueError, match="coefficient"):
        f._check_coefficients(np.ones(ns - 1))
    for array in (f.radii, f.outer_indices, f.edge_left, *f.rank_arrays(0)):
        with pytest.raises(ValueError):
            array.setflags(write=True)
    f.rank_arrays(2)
    assert len(f._rank_cache) == 1
    bad = dict(bd)
    bad["bsp_p"] = bd["bsp_p"].copy()
    bad["bsp_p"][0, 0, 0] += 0.01
    with pytest.raises(ValueError, match="basis tables"):
        CellCoulombFeatures(bad)


def test_polynomial_payload_and_sizes():
    ns, ks = 12, 4
    bd = prepare_bspline_data(make_knots(ns, ks, 0.02, 3.0), ns, ks, kinetic_balance=True)
    native = CellCoulombFeatures(bd)
    poly = PolynomialCellCoulombFeatures(bd)
    assert poly.identity["auxiliary_nodes"] < native.identity["auxiliary_nodes"]
    need = poly.identity["payload_estimate_bytes"]
    assert need > native.identity["payload_estimate_bytes"]
    with pytest.raises(MemoryError, match="complete polynomial"):
        PolynomialCellCoulombFeatures(bd, memory_bytes=need - 1)
    assert PolynomialCellCoulombFeatures(bd, memory_bytes=need).identity == poly.identity
    nv, m = bd["nv"], 2 * ks - 1
    assert poly.radii.size == 2 * nv * m
    assert poly.edge_left.size == nv * (2 * m) * (2 * m + 1) // 2


def test_oracle_pair_contraction_is_the_quartet_sum():
    ns, ks = 10, 4
    bd = prepare_bspline_data(make_knots(ns, ks, 0.02, 4.0), ns, ks, kinetic_balance=True)
    f = CellCoulombFeatures(bd)
    banks = O.Banks(bd)
    rng = np.random.default_rng(381)
    waves = [[(rng.normal(size=ns), rng.normal(size=ns)) for _ in range(n)]
             for n in (2, 2, 1, 3)]
    cols = [np.column_stack([host_sample(f, *a) for a in g]) for g in waves]
    rho = rng.normal(size=(2, 1))
    for rank in (0, 3):
        want = np.zeros((2, 3))
        for m, r, s, n in itertools.product(range(2), range(2), range(1), range(3)):
            quartet = (waves[0][m], waves[1][r], waves[2][s], waves[3][n])
            want[m, n] += rho[r, s] * O.compute_slater_rk_cell(
                *(v for a in quartet for v in a), bd, rank, banks)
        got = OF.pair_contraction(f, *cols, rank, rho)
        np.testing.assert_allclose(got, want, atol=2e-11, rtol=3e-13)
