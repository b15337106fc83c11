"""Pins the CPU oracle of the (Y) Poisson discretisation before it is trusted.

Checks closed forms and invariants that a dropped term, a wrong r-power, a
swapped r_</r_> or a mis-indexed partial matrix would fail (reference tests
tests/test_slater_integrator.py and tests/test_slater_kernel_crosscheck.py).
"""

import math

import numpy as np
import pytest

import oracle as O
from atomx_amd import bspline as B
from helpers import RK_1S, physics_workspace, rho1s


@pytest.fixture(scope="module")
def h_grid():
    bd, t = physics_workspace(Z=1, ks=9)
    return bd


def test_gauss_legendre_matches_numpy():
    for n in (2, 5, 7, 8, 9, 16, 33):
        x, w = O.gauss_legendre(n)
        xn, wn = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xn)) < 2e-16
        # numpy's weights carry ~1e-13 relative error at n = 33
        assert np.max(np.abs(w - wn) / wn) < 3e-13
        assert abs(w.sum() - 2.0) < 4e-15


def test_gauss_legendre_01_integrates_polynomials():
    for n in (3, 9, 12):
        x, w = O.gauss_legendre_01(n)
        for p in range(2 * n):
            assert np.dot(w, x ** p) == pytest.approx(1.0 / (p + 1), rel=1e-14)


@pytest.mark.parametrize("n_gauss", [5, 8, 9, 17])
def test_partial_matrices_invariants(n_gauss):
    """mo + mi = 1, mo@w = 1+x, mi@w = 1-x, exact cubic partials (yk.py:42-47)."""
    xi, wi = np.polynomial.legendre.leggauss(n_gauss)
    mo, mi = O.yk_partial_matrices(n_gauss)
    # absolute gates a few ulps above the rounding of the checks themselves
    # (the reference gates these at atol 1e-15 with numpy's default rtol 1e-5)
    assert np.allclose(mo + mi, 1.0, atol=1e-15, rtol=0)
    assert np.allclose(mo @ wi, 1.0 + xi, atol=4e-15, rtol=0)
    assert np.allclose(mi @ wi, 1.0 - xi, atol=4e-15, rtol=0)
    g = wi * (xi ** 3 + 2.0 * xi)
    assert np.allclose(g @ mo.T, xi ** 4 / 4.0 + xi ** 2 - 1.25, atol=8e-15, rtol=0)
    assert np.allclose(g @ mi.T, 1.25 - xi ** 4 / 4.0 - xi ** 2, atol=8e-15, rtol=0)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_hydrogenic_1s_poisson_closed_form(h_grid, k):
    """R^k(1s,1s) on the physics grid: 5/8, 3/8, R^2 (measured 3.6e-16/3.0e-16/4.9e-12)."""
    gr, grw, nv = h_grid["gr"], h_grid["grw"], h_grid["nv"]
    rho = rho1s(gr)
    got = O.slater_rk_poisson(rho, rho, gr, grw, nv, gr.shape[1], k)
    assert got == pytest.approx(RK_1S[k], rel=1e-11)


@pytest.mark.parametrize("Z", [1, 10, 26])
def test_r0_hydrogenic_both_methods(Z):
    """R^0(1s,1s) = 5Z/8 by the Poisson and the direct integrators."""
    bd, _ = physics_workspace(Z=Z, ks=9)
    gr, grw, nv = bd["gr"], bd["grw"], bd["nv"]
    rho = rho1s(gr, Z)
    assert O.slater_rk_poisson(rho, rho, gr, grw, nv, gr.shape[1], 0) == pytest.approx(
        5.0 * Z / 8.0, rel=1e-10)
    assert O.slater_rk_direct(rho, rho, gr, grw, 0) == pytest.approx(5.0 * Z / 8.0, rel=1e-8)


@pytest.fixture(scope="module")
def fine_grid():
    """Exponential knots from 1e-5 (the exp(-r) density does not vanish at 0,
    so Y^k of it is endpoint-singular and needs a small first cell)."""
    t = B.make_knots(60, 9, 1e-5, 40.0)
    return B.prepare_bspline_data(t, 60, 9)


def exp_density_rk_reference(k):
    """R^k(rho, rho), rho = exp(-r): Y^k(r) = [gamma(k+1, r) + r^(2k+1) Gamma(-k, r)]
    / r^(k+1) integrated against exp(-r) with 40-digit mpmath (k=0: ln 4)."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40

    def yk_ref(rv):
        r = mp.mpf(rv)
        return (mp.gammainc(k + 1, 0, r) + r ** (2 * k + 1) * mp.gammainc(-k, r, mp.inf)) / r ** (k + 1)

    return float(mp.quad(lambda r: mp.exp(-r) * yk_ref(r), [0, 1, 3, 10, mp.inf]))


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_exponential_density_against_mpmath(fine_grid, k):
    """R^k of exp(-r) against mpmath (measured <= 2e-15 for both integrators)."""
    ref = exp_density_rk_reference(k)
    if k == 0:
        assert ref == pytest.approx(math.log(4.0), rel=1e-15)
    gr, grw, nv = fine_grid["gr"], fine_grid["grw"], fine_grid["nv"]
    rho = np.exp(-gr)
    assert O.slater_rk_poisson(rho, rho, gr, grw, nv, gr.shape[1], k) == pytest.approx(ref, rel=1e-12)
    assert O.slater_rk_direct(rho, rho, gr, grw, k) == pytest.approx(ref, rel=1e-12)


@pytest.mark.parametrize("n_gauss,p,q", [(2, 0, 0), (5, 2, 3), (9, 4, 6)])
@pytest.mark.parametrize("rank", [0, 3, 12])
def test_direct_origin_cell_monomial(n_gauss, p, q, rank):
    """Analytic double integral of r^p, r^q on one cell at the origin (rel 2e-13)."""
    x, w = np.polynomial.legendre.leggauss(n_gauss)
    radius = 0.7
    r, weights = (radius * (x + 1) / 2)[None, :], (radius * w / 2)[None, :]
    expected = radius ** (p + q + 1) / (p + q + 1) * (1 / (p + rank + 1) + 1 / (q + rank + 1))
    assert O.slater_rk_direct(r ** p, r ** q, r, weights, rank) == pytest.approx(expected, rel=2e-13)
    assert O.slater_rk_direct(r ** q, r ** p, r, weights, rank) == pytest.approx(expected, rel=2e-13)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_yk_pointwise_against_incomplete_gamma(h_grid, k):
    """Y^k(r) of the hydrogenic density outside the first knot cell (<= 1e-12),
    against the incomplete-gamma closed form."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    gr, grw, nv = h_grid["gr"], h_grid["grw"], h_grid["nv"]
    y = O.compute_yk(rho1s(gr), gr, grw, nv, gr.shape[1], k)
    rr = gr.ravel()
    sel = (rr > gr[1, 0]) & (rr < 20.0)
    sel_idx = np.flatnonzero(sel)[::7]

    def ref(rv):
        r = mp.mpf(float(rv))
        lo = 4 * mp.gammainc(k + 3, 0, 2 * r) / mp.mpf(2) ** (k + 3)
        hi = 4 * mp.gammainc(2 - k, 2 * r, mp.inf) / mp.mpf(2) ** (2 - k)
        return float(r ** (-(k + 1)) * lo + r ** k * hi)

    refs = np.array([ref(rr[i]) for i in sel_idx])
    got = y.ravel()[sel_idx]
    assert np.max(np.abs(got - refs) / np.abs(refs)) < 1e-12


def test_poisson_matches_direct_on_smooth_pair_densities(fine_grid):
    """Independent integration paths agree (rel 2e-12) on products of
    hydrogenic-like orbitals (reference tests/test_slater_integrator.py:60-66)."""
    gr, grw, nv = fine_grid["gr"], fine_grid["grw"], fine_grid["nv"]
    r = gr
    a = r * np.exp(-2.0 * r) * (1 - 0.5 * r)
    b = r ** 2 * np.exp(-0.9 * r)
    c = r * np.exp(-1.3 * r)
    d = r ** 3 * np.exp(-0.7 * r) * (1 - 0.2 * r)
    for k in range(5):
        fast = O.slater_rk_poisson(a * b, c * d, gr, grw, nv, gr.shape[1], k)
        slow = O.slater_rk_direct(a * b, c * d, gr, grw, k)
        assert fast == pytest.approx(slow, rel=2e-12, abs=1e-14)


def test_exchange_kernel_against_brute_force():
    """_exchange_matrix_kernel restated as an O(N^2) double sum with the
    node-wise cumulative kernel (both r_< orderings, diagonal node counted in
    both cumulative sums), on a tiny grid."""
    t = B.make_exponential_knots(10, 4, 0.05, 8.0)
    bd = B.prepare_bspline_data(t, 10, 4, n_gauss=5)
    gr, grw, nv, ks, ns = bd["gr"], bd["grw"], bd["nv"], bd["ks"], bd["ns"]
    rng = np.random.default_rng(5)
    fb = rng.normal(size=gr.shape)
    for k in (0, 1, 3):
        K = O.exchange_matrix_kernel(fb, bd["bsp_p"], gr, grw, nv, gr.shape[1], ks, ns, k)
        r = gr.ravel()
        w = grw.ravel()
        f = fb.ravel()
        Bfull = np.zeros((ns, r.size))
        for iv in range(nv):
            for s in range(ks):
                Bfull[iv + s, iv * gr.shape[1]:(iv + 1) * gr.shape[1]] = bd["bsp_p"][iv, :, s]
        ri, rj = r[:, None], r[None, :]
        lower = (np.arange(r.size)[None, :] <= np.arange(r.size)[:, None])
        upper = (np.arange(r.size)[None, :] >= np.arange(r.size)[:, None])
        kern = np.where(lower, rj ** k / ri ** (k + 1), 0.0) + np.where(upper, ri ** k / rj ** (k + 1), 0.0)
        left = Bfull * (w * f)[None, :]
        right = Bfull * (w * f)[None, :]
        ref = left @ kern @ right.T
        assert np.allclose(K, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())


def test_rk_quartets_matches_per_integral_poisson(h_grid):
    gr, grw, nv = h_grid["gr"], h_grid["grw"], h_grid["nv"]
    ng = gr.shape[1]
    r = gr.ravel()
    rho = np.stack([r ** 2 * np.exp(-a * r) for a in (1.0, 2.0, 3.5)])
    qp = np.array([0, 1, 2, 0], dtype=np.int32)
    qq = np.array([0, 2, 1, 2], dtype=np.int32)
    qk = np.array([0, 1, 2, 3], dtype=np.int32)
    out = O.rk_quartets(rho, gr, grw, nv, ng, qp, qq, qk)
    for t in range(4):
        ref = O.slater_rk_poisson(rho[qp[t]], rho[qq[t]], gr, grw, nv, ng, int(qk[t]))
        assert out[t] == pytest.approx(ref, rel=1e-13)
