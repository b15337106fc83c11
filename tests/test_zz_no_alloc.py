"""GPU: ax_rk_gemm_noalloc (the contraction with the library's static device
scratch instead of a caller workspace) against ax_rk_gemm on the same
operands.  Both run the same kernels with the same K split, so results must
be bitwise equal; two calls in a row, the second on a smaller block, show
that a call does not depend on what the previous one left in the shared
scratch.  A block whose workspace exceeds the scratch is refused with
MemoryError before any launch."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gemm(name, D, Y, rows, m, n, prec):
    from atomx_amd import device as dv
    from atomx_amd._lib import PRECISIONS, call, load
    Dd, Yd, rd = dv.f64(D), dv.f64(Y), dv.i32(rows)
    out = dv.empty((m * (m + 1) // 2,))
    if name == "ax_rk_gemm":
        nbytes = load().ax_rk_gemm_workspace_bytes(m, n, PRECISIONS[prec])
        ws = dv.empty((max(nbytes, 8) // 8 + 1,))
        call(name, dv.ptr(Dd), n, dv.ptr(rd), dv.ptr(Yd), n, m, n, PRECISIONS[prec],
             dv.ptr(ws), dv.ptr(out), dv.stream_handle())
    else:
        call(name, dv.ptr(Dd), n, dv.ptr(rd), dv.ptr(Yd), n, m, n, PRECISIONS[prec],
             dv.ptr(out), dv.stream_handle())
    return dv.host(out)


@pytest.mark.parametrize("prec", ["fp64", "split"])
def test_noalloc_matches_workspace_entry_point(prec):
    rng = np.random.default_rng(17)
    for m, n in ((200, 1500), (37, 100)):   # the second call is the smaller one
        print(f"ax_rk_gemm_noalloc {prec} m={m} n={n}", flush=True)
        D = rng.normal(size=(m + 3, n))
        Y = rng.normal(size=(m, n))
        rows = rng.permutation(m + 3)[:m].astype(np.int32)
        got = _gemm("ax_rk_gemm_noalloc", D, Y, rows, m, n, prec)
        ref = _gemm("ax_rk_gemm", D, Y, rows, m, n, prec)
        assert np.array_equal(got, ref), (m, n, prec)
        i, j = np.triu_indices(m)
        exact = (D[rows] @ Y.T)[i, j]
        absref = (np.abs(D[rows]) @ np.abs(Y).T)[i, j]
        tol = 1e-14 if prec == "fp64" else 1e-11
        assert np.max(np.abs(got - exact) / absref) < tol


def test_noalloc_refuses_blocks_beyond_its_scratch():
    from atomx_amd import device as dv
    from atomx_amd._lib import PRECISIONS, call, load
    m, n = 2000, 6000
    need = load().ax_rk_gemm_workspace_bytes(m, n, PRECISIONS["split"])
    assert need > (32 << 20)
    D = dv.empty((m, n))
    rows = dv.i32(np.arange(m, dtype=np.int32))
    out = dv.empty((m * (m + 1) // 2,))
    with pytest.raises(MemoryError, match="scratch"):
        call("ax_rk_gemm_noalloc", dv.ptr(D), n, dv.ptr(rows), dv.ptr(D), n, m, n,
             PRECISIONS["split"], dv.ptr(out), dv.stream_handle())
