"""GPU: ax_combine_rows, out[t] = sum_j coef[t, j] dens[j] over rows of
length L, against a float64 host product on shapes that leave partial
groups of eight targets and of four source rows, with zero (and negative
zero) coefficients skipped, and with no source rows at all."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nt,n_src,L", [(13, 7, 1000), (8, 4, 256), (1, 1, 3), (21, 0, 50),
                                        (17, 130, 4097)])
def test_combine_rows_matches_host(nt, n_src, L):
    from atomx_amd import device as dv
    from atomx_amd._lib import call
    rng = np.random.default_rng(nt * 1000 + n_src)
    dens = rng.standard_normal((n_src, L))
    coef = rng.standard_normal((nt, n_src))
    coef[rng.random((nt, n_src)) < 0.3] = 0.0
    if coef.size:
        coef.flat[0] = -0.0
    d, c = dv.f64(dens), dv.f64(coef)
    out = dv.empty((nt, L))
    call("ax_combine_rows", dv.ptr(d), L, n_src, dv.ptr(c), nt, dv.ptr(out),
         dv.stream_handle())
    got = dv.host(out)
    ref = coef @ dens if n_src else np.zeros((nt, L))
    scale = np.abs(coef) @ np.abs(dens) if n_src else np.ones((nt, L))
    # both sums round at most n_src times: |got - ref| <= 2 n u sum |c||d|
    tol = 2 * max(n_src, 1) * 2.0 ** -53
    assert np.all(np.abs(got - ref) <= tol * np.maximum(scale, 1e-300))
    if n_src == 0:
        assert np.array_equal(got, np.zeros((nt, L)))
