"""Y^k potentials on the dense Gauss grid — drop-in for numerics/yk.py.

    Y^k(r) = r^-(k+1) int_0^r rho r'^k dr' + r^k int_r^inf rho r'^-(k+1) dr'

Same discretisation as the reference (numerics/yk.py:87-154): completed cells
contribute full Gauss sums, the straddled cell the interpolatory partial
integrals of ``yk_partial_matrices``.  The scan runs on the GPU in FP64
(``ax_yk_batch``); complex c-product densities are handled as their real and
imaginary rows (Y^k is linear in rho).
"""

from __future__ import annotations

import numpy as np

from . import device as dv
from . import grid as _grid
from ._lib import call

__all__ = ["yk_partial_matrices", "yk_flat", "compute_yk", "yk_rows"]


def yk_partial_matrices(n_gauss):
    """Return ``(mo, mi)`` for the interpolatory partial-interval rule (yk.py:65-84)."""
    return _grid.partial_matrices(int(n_gauss))


def yk_rows(rho_dev, tables, ks, weighted=False, out=None, stream=None, mo=None, mi=None):
    """Y^k of many densities on the device.

    rho_dev : (n_rows, n) FP64 device tensor (w*rho rows when ``weighted``).
    ks      : one rank per row (sequence) or a single int for all rows.
    mo, mi  : optional (ng, ng) FP64 device tensors replacing the grid's own
              partial-interval matrices for this call.
    Returns an (n_rows, n) device tensor.
    """
    n_rows, n = rho_dev.shape
    if n != tables.n:
        raise ValueError(f"density rows have {n} samples, grid has {tables.n}")
    ks = np.broadcast_to(np.asarray(ks, dtype=np.int32), (n_rows,)).copy()
    if ks.size and (ks.min() < 0):
        raise ValueError("multipole rank must be a nonnegative integer")
    tables.ensure_k(int(ks.max()) if ks.size else 0)
    if out is None:
        out = dv.empty((n_rows, n))
    rows = dv.i32(np.arange(n_rows, dtype=np.int32))
    kk = dv.i32(ks)
    call("ax_yk_batch", dv.ptr(rho_dev), dv.ld(rho_dev), None if weighted else dv.ptr(tables.w),
         dv.ptr(tables.rk), dv.ptr(tables.rkm), n, tables.nv, tables.ng, tables.kmax,
         dv.ptr(tables.mo if mo is None else mo), dv.ptr(tables.mi if mi is None else mi),
         n_rows, dv.ptr(rows), dv.ptr(kk), dv.ptr(rows),
         dv.ptr(out), dv.ld(out), dv.stream_handle(stream))
    return out


def _caller_matrix(m, own, ng, name):
    """Device copy of a caller's partial-interval matrix, or None when it is
    absent or equal to the grid's own (whose device copy is reused)."""
    if m is None:
        return None
    m = np.asarray(m, dtype=np.float64)
    if m.shape != (ng, ng):
        raise ValueError(f"{name} must be ({ng}, {ng}), got {m.shape}")
    if np.array_equal(m, own):
        return None
    return dv.f64(m)


def yk_flat(rho_flat, r_flat, w_flat, n, k, nv, ng, mo=None, mi=None):
    """Y^k on a flattened Gauss grid (yk.py:87-154), computed on the GPU.

    ``mo``/``mi`` are the partial-interval matrices of the straddled cell, as
    in the reference; when omitted (or equal to ``yk_partial_matrices(ng)``)
    the engine's resident copies are used, otherwise the caller's matrices
    are uploaded for this call.
    """
    if k < 0 or int(k) != k:
        raise ValueError("multipole rank must be a nonnegative integer")
    rho = np.asarray(rho_flat).ravel()
    if rho.size != n or int(n) != int(nv) * int(ng):
        raise ValueError(f"density has {rho.size} samples, expected n = nv*ng = {n}")
    tables = _grid.get(r_flat, w_flat, nv, ng, int(k))
    own_mo, own_mi = yk_partial_matrices(int(ng))
    mo_d = _caller_matrix(mo, own_mo, int(ng), "mo")
    mi_d = _caller_matrix(mi, own_mi, int(ng), "mi")
    if np.iscomplexobj(rho):
        parts = np.stack([rho.real, rho.imag]).astype(np.float64)
        y = dv.host(yk_rows(dv.f64(parts), tables, int(k), mo=mo_d, mi=mi_d))
        return y[0] + 1j * y[1]
    rho = np.asarray(rho, dtype=np.result_type(rho, np.float64))
    y = yk_rows(dv.f64(rho[None, :]), tables, int(k), mo=mo_d, mi=mi_d)
    return dv.host(y)[0]


def compute_yk(rho, gr, grw, nv, n_gauss, k):
    """Y^k on the (nv, n_gauss) quadrature grid (yk.py:157-185)."""
    rho = np.asarray(rho)
    rho = np.asarray(rho, dtype=np.result_type(rho, np.float64))
    y = yk_flat(rho.ravel(), np.ravel(gr), np.ravel(grw), int(nv) * int(n_gauss), int(k),
                int(nv), int(n_gauss))
    return y.reshape(int(nv), int(n_gauss))
