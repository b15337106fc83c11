"""Auxiliary-node form of the cell Coulomb quadrature and the pair-density
matrix action built on it.

``CellCoulombFeatures`` / ``PolynomialCellCoulombFeatures`` restate
numerics/bspline/coulomb_features.py:64-524 and ``CellPairContraction``
restates numerics/bspline/pair_contraction.py:18-148, with the same names,
arguments, attributes, identities and exceptions.

The cell algorithm's R^k (outer Gauss nodes, triangular sub-quadrature of
order ks + k//2 inside a cell, cross-cell moments) is written as a sparse
symmetric operator K_k on auxiliary nodes: every spinor component is
sampled at the outer nodes, at the sub-nodes of each admitted order and at
2 ks - 1 interpolation nodes per cell, and R^k(ab;cd) = rho_ab . K_k
rho_cd with rho = the c-product of samples.  K_k = cross-cell part
(rising/falling moments, prefix/suffix over cells) + same-cell edges
between an outer node and the nodes below it in the cell.  For k >= 2 the
sub-nodes of order ks + k//2 are reached by exact polynomial interpolation
from the interpolation nodes (a pair density is a polynomial of degree
2 ks - 2 in each cell).

Geometry (node radii, B-spline tables at the nodes, edge lists, per-rank
weights) is host FP64 set-up, once per grid and rank, like the B-spline
tables.  Sampling, the action on many densities at once, integrals and the
pair contraction run on the GPU (``ax_feat_*``, ``ax_cpc_*``).  Host
arrays in give host arrays out; device tensors in give device tensors out.
"""

from __future__ import annotations

import ctypes
import hashlib
from collections import OrderedDict
from types import MappingProxyType

import numpy as np

from . import device as dv
from ._lib import call, load
from .bspline import evaluate_bsplines_on_grid

__all__ = ["CellCoulombFeatures", "PolynomialCellCoulombFeatures", "CellPairContraction",
           "cardinal_values"]

_MiB = 1024 ** 2
_dp = ctypes.POINTER(ctypes.c_double)


def _frozen(a):
    """Read-only copy that cannot be made writeable again."""
    a = np.ascontiguousarray(a)
    return np.frombuffer(a.tobytes(), dtype=a.dtype).reshape(a.shape)


def _positive_int(label, value):
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)) or value < 1:
        raise ValueError(f"positive integer {label} required")


def _rank(rank):
    if isinstance(rank, bool) or not isinstance(rank, (int, np.integer)) or rank < 0:
        raise ValueError("nonnegative integer multipole required")
    return int(rank)


def _gl01(n):
    """The cell algorithm's n-point Gauss-Legendre rule on [0, 1]."""
    x, w = np.empty(n), np.empty(n)
    call("ax_gauss_legendre_01", int(n), x.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    return x, w


def cardinal_values(points, nodes):
    """Lagrange cardinal polynomials of ``nodes`` evaluated at ``points``
    (shape points.shape + (len(nodes),)), in the second barycentric form
    (Berrut & Trefethen, SIAM Review 46 (2004) 501); a point on a node gives
    that node's unit vector."""
    nodes = np.asarray(nodes, dtype=float)
    gap = nodes[:, None] - nodes[None, :]
    np.fill_diagonal(gap, 1.0)
    lam = 1.0 / np.prod(gap, axis=1)
    d = np.asarray(points, dtype=float)[..., None] - nodes
    hit = d == 0
    q = lam / np.where(hit, 1.0, d)
    q = q / q.sum(axis=-1, keepdims=True)
    return np.where(hit.any(axis=-1, keepdims=True), hit.astype(float), q)


def _s():
    return dv.stream_handle()


class CellCoulombFeatures:
    """Exact sparse radial action on the native B-spline quadrature samples
    (coulomb_features.py:64-337).

    ``dot(sample(a) * sample(b), apply_density(sample(c) * sample(d), k))``
    is the cell algorithm's R^k(ab;cd) with all four component terms;
    complex values use c-products.  ``apply_density`` returns a weighted
    action (not a pointwise potential)."""

    def __init__(self, workspace, *, memory_bytes=128 * _MiB, rank_cache_size=8):
        _positive_int("memory_bytes", memory_bytes)
        _positive_int("rank_cache_size", rank_cache_size)
        gr = np.asarray(workspace["gr"], dtype=float)
        grw = np.asarray(workspace["grw"], dtype=float)
        nv, ns, ks = int(workspace["nv"]), int(workspace["ns"]), int(workspace["ks"])
        orders = (int(workspace.get("ksp", ks)), int(workspace.get("ksq", ks)))
        sizes = (int(workspace.get("nsp", ns)), int(workspace.get("nsq", ns)))
        knots = (np.asarray(workspace.get("tp", workspace["t"]), dtype=float),
                 np.asarray(workspace["t"], dtype=float))
        if (gr.ndim != 2 or gr.shape[0] != nv or grw.shape != gr.shape
                or not np.all(np.isfinite(gr)) or not np.all(np.isfinite(grw))
                or np.any(gr <= 0) or np.any(grw <= 0) or min(orders) < 1):
            raise ValueError("finite positive native cell quadrature required")
        breaks = [t[k - 1:k + nv] for t, k in zip(knots, orders)]
        if (any(b.size != nv + 1 or np.any(np.diff(b) <= 0) for b in breaks)
                or not np.array_equal(breaks[0], breaks[1])
                or np.any(gr <= breaks[0][:-1, None]) or np.any(gr >= breaks[0][1:, None])):
            raise ValueError("P/Q must have identical ordered cell boundaries")
        ng = gr.shape[1]
        top = max(orders)
        m = 2 * top - 1                                   # interpolation nodes per cell
        subs = [tuple(sorted({o, top})) for o in orders]  # sub-rule orders per component
        widths = tuple(ng * (1 + sum(s)) + m for s in subs)
        n_nodes = nv * sum(widths)
        legacy = nv * ng * (sum(orders) + 2 * top)
        n_edges = legacy + 4 * nv * ng * m
        # the reference's payload bound (coulomb_features.py:128-135)
        estimate = (8 * nv * sum(w * k for w, k in zip(widths, orders)) + 8 * n_nodes
                    + 32 * nv * ng + 32 * n_edges
                    + rank_cache_size * 8 * (2 * nv * ng + n_edges) + 8 * (nv + 1 + m))
        if estimate > memory_bytes:
            raise MemoryError(
                f"complete auxiliary Coulomb payload needs {estimate} bytes; cap={memory_bytes}")

        lo, width_cell = breaks[0][:-1], np.diff(breaks[0])
        span = gr - lo[:, None]
        dens_x = (np.polynomial.legendre.leggauss(m)[0] + 1.0) / 2.0
        outer = np.empty((2, nv, ng), dtype=np.int64)
        sub_nodes, dens_nodes, radii, tables = [{}, {}], [], [], []
        base = 0
        for c in range(2):
            R = np.empty((nv, widths[c]))
            first = base + np.arange(nv)[:, None] * widths[c]      # (nv, 1) first node per cell
            R[:, :ng] = gr
            outer[c] = first + np.arange(ng)
            col = ng
            for q in subs[c]:
                x, w = _gl01(q)
                rad = lo[:, None, None] + span[:, :, None] * x     # (nv, ng, q)
                R[:, col:col + ng * q] = rad.reshape(nv, ng * q)
                idx = first[:, :, None] + col + np.arange(ng * q).reshape(ng, q)
                sub_nodes[c][q] = (idx, span[:, :, None] * w, rad)
                col += ng * q
            R[:, col:] = lo[:, None] + width_cell[:, None] * dens_x
            dens_nodes.append(first + col + np.arange(m))
            tab = evaluate_bsplines_on_grid(knots[c], sizes[c], orders[c], R, nderiv=1)[..., 0]
            native = np.asarray(workspace["bsp_p"] if c == 0
                                else workspace.get("bsp_q", workspace["bsp_p"]))
            if native.shape != (nv, ng, orders[c]) or np.max(abs(tab[:, :ng] - native)) > 1e-11:
                raise ValueError("native basis tables do not match declared spline geometry")
            tab[:, :ng] = native  # the outer nodes use the workspace's own tables
            tables.append(_frozen(tab))
            radii.append(R.ravel())
            base += nv * widths[c]

        left, right, e_base, e_ratio = [], [], [], []
        for a in range(2):          # outer node's component
            for b in range(2):      # lower node's component
                q = orders[a] if a == b else top
                idx, sw, rad = sub_nodes[b][q]
                left.append(np.broadcast_to(outer[a][:, :, None], idx.shape).ravel())
                right.append(idx.ravel())
                e_base.append((grw[:, :, None] * sw / gr[:, :, None]).ravel())
                e_ratio.append((rad / gr[:, :, None]).ravel())
        for a in range(2):
            for b in range(2):
                shape = (nv, ng, m)
                left.append(np.broadcast_to(outer[a][:, :, None], shape).ravel())
                right.append(np.broadcast_to(dens_nodes[b][:, None, :], shape).ravel())
                e_base.append(np.zeros(nv * ng * m))
                e_ratio.append(np.zeros(nv * ng * m))

        self._tables = tuple(tables)
        self._orders, self._sizes, self._ns = orders, sizes, ns
        self._nv, self._ng = nv, ng
        self._gr, self._weights = _frozen(gr), _frozen(grw)
        self._breaks = _frozen(breaks[0])
        self._dens_x = _frozen(dens_x)
        self._legacy_edges = legacy
        self._memory_bytes = int(memory_bytes)
        self.radii, self.outer_indices = _frozen(np.concatenate(radii)), _frozen(outer)
        self.edge_left, self.edge_right = (_frozen(np.concatenate(left)),
                                           _frozen(np.concatenate(right)))
        self.edge_base, self.edge_ratio = (_frozen(np.concatenate(e_base)),
                                           _frozen(np.concatenate(e_ratio)))
        self._rank_cache, self._rank_capacity = OrderedDict(), int(rank_cache_size)
        self._node_bases = (0, nv * widths[0])
        self._widths = widths
        self._dev = None
        self._dev_rank = OrderedDict()
        h = hashlib.sha256()
        for value in (gr, grw, *knots, *tables):
            h.update(repr((value.shape, value.dtype.str)).encode())
            h.update(np.ascontiguousarray(value).tobytes())
        self.identity = MappingProxyType(dict(
            representation="native cell Coulomb auxiliary nodes",
            geometry_sha256=h.hexdigest(),
            coefficient_sizes=sizes,
            component_orders=orders,
            auxiliary_nodes=n_nodes,
            same_cell_edges=n_edges,
            quadrature_nodes=(nv, ng),
            payload_estimate_bytes=estimate,
            rank_cache_capacity=rank_cache_size,
            inner_quadrature="native order + floor(rank/2) for nonnegative ranks",
            contraction="weighted bilinear c-product; no new quadrature or screening",
        ))

    # -- host geometry per rank ------------------------------------------------
    def rank_arrays(self, rank):
        """(rising, falling, edge weights) of one multipole, read-only and
        cached (LRU of rank_cache_size ranks)."""
        rank = _rank(rank)
        if rank not in self._rank_cache:
            with np.errstate(over="raise", invalid="raise", divide="raise"):
                edges = self._edge_weights(rank)
                rising = self._weights * self._gr ** rank
                falling = self._weights * self._gr ** (-rank - 1)
            self._rank_cache[rank] = tuple(_frozen(x) for x in (rising, falling, edges))
            while len(self._rank_cache) > self._rank_capacity:
                self._rank_cache.popitem(last=False)
        else:
            self._rank_cache.move_to_end(rank)
        return self._rank_cache[rank]

    def _edge_weights(self, rank):
        """Same-cell edge weights of one rank: the native sub-rule edges for
        k < 2; for k >= 2 the rule of order ks + k//2 expressed on the
        interpolation nodes (coulomb_features.py:297-330)."""
        if rank < 2:
            return self.edge_base * self.edge_ratio ** rank
        nv, ng = self._gr.shape
        m = self._dens_x.size
        n_sub = max(self._orders) + rank // 2
        scratch = 8 * nv * ng * n_sub * (5 * m + 8) + 8 * self.edge_base.size
        if self.identity["payload_estimate_bytes"] + scratch > self._memory_bytes:
            raise MemoryError("complete rank-aware Coulomb workspace exceeds memory_bytes")
        weights = np.zeros(self.edge_base.size)
        pos = self._legacy_edges
        lo = self._breaks[:-1]
        span = self._gr - lo[:, None]
        width = np.diff(self._breaks)
        top = max(self._orders)
        for a in range(2):
            for b in range(2):
                q = (self._orders[a] if a == b else top) + rank // 2
                x, w = _gl01(q)
                t = span[:, :, None] * x / width[:, None, None]        # local coordinate
                card = cardinal_values(t, self._dens_x)                # (nv, ng, q, m)
                sub = lo[:, None, None] + span[:, :, None] * x
                radial = span[:, :, None] * w * (sub / self._gr[:, :, None]) ** rank
                block = np.einsum("cgq,cgqj->cgj", radial, card)
                block *= (self._weights / self._gr)[:, :, None]
                weights[pos:pos + nv * ng * m] = block.ravel()
                pos += nv * ng * m
        return weights

    # -- device tables ------------------------------------------------------------
    @property
    def _ngo(self):
        return self._ng

    def _device(self):
        if self._dev is None:
            dv.require_gpu()
            load()
            n = self.radii.size
            outer = np.asarray(self.outer_indices, dtype=np.int64)
            slot = np.full(n, -1, dtype=np.int32)
            nv, ngo = self._nv, self._ngo
            cell_slot = (np.arange(nv)[:, None] * ngo + np.arange(ngo)).astype(np.int32)
            for c in range(2):
                slot[outer[c].ravel()] = cell_slot.ravel()
            self._dev = {
                "outer": dv.i32(outer.ravel()),
                "slot": dv.i32(slot),
                "tables": [dv.f64(t) for t in self._tables],
            }
        return self._dev

    def _rank_device(self, rank):
        """Device rising/falling weights and the same-cell edges of one rank:
        by destination (CSR) for the action, as directed (src, dst, w) for
        the pair contraction.  Zero-weight edges are dropped."""
        rank = _rank(rank)
        d = self._dev_rank.get(rank)
        if d is None:
            rising, falling, w = self.rank_arrays(rank)
            keep = w != 0
            lft = np.asarray(self.edge_left)[keep]
            rgt = np.asarray(self.edge_right)[keep]
            ww = np.asarray(w)[keep]
            dst = np.stack((lft, rgt), axis=1).ravel()
            src = np.stack((rgt, lft), axis=1).ravel()
            wd = np.repeat(ww, 2)
            order = np.argsort(dst, kind="stable")   # edge order kept within a row
            n = self.radii.size
            row_ptr = np.zeros(n + 1, dtype=np.int64)
            np.cumsum(np.bincount(dst, minlength=n), out=row_ptr[1:])
            d = {
                "rising": dv.f64(rising), "falling": dv.f64(falling),
                "row_ptr": dv.to_dev(row_ptr), "csr_src": dv.i32(src[order]),
                "csr_w": dv.f64(wd[order]),
                "src": dv.i32(src), "dst": dv.i32(dst), "w": dv.f64(wd),
                "n_directed": int(src.size),
            }
            self._dev_rank[rank] = d
            while len(self._dev_rank) > self._rank_capacity:
                self._dev_rank.popitem(last=False)
        else:
            self._dev_rank.move_to_end(rank)
        return d

    # -- sampling -------------------------------------------------------------------
    def _check_coefficients(self, values):
        values = np.asarray(values)
        if values.shape != (self._ns,) or not np.all(np.isfinite(values)):
            raise ValueError("finite padded native coefficient vector required")
        return values

    def sample_columns(self, P, Q):
        """Samples of n spinors with coefficient stacks P, Q (n, ns) at every
        auxiliary node: device tensor (n_nodes, n), complex if any
        coefficient is."""
        P = np.atleast_2d(np.asarray(P))
        Q = np.atleast_2d(np.asarray(Q))
        if P.shape != Q.shape or P.ndim != 2 or P.shape[1] != self._ns:
            raise ValueError("finite padded native coefficient vector required")
        if not (np.all(np.isfinite(P)) and np.all(np.isfinite(Q))):
            raise ValueError("finite padded native coefficient vector required")
        cplx = np.iscomplexobj(P) or np.iscomplexobj(Q)
        dtype = np.complex128 if cplx else np.float64
        dev = self._device()
        torch = dv.torch()
        n_col = P.shape[0]
        out = torch.empty((self.radii.size, n_col),
                          dtype=torch.complex128 if cplx else torch.float64, device=dv.device())
        for c, coef in enumerate((P, Q)):
            cd = dv.to_dev(np.ascontiguousarray(coef, dtype=dtype))
            tab = dev["tables"][c]
            nv, width, order = tab.shape
            call("ax_feat_sample", dv.ptr(tab), nv, width, order, self._sizes[c], dv.ptr(cd),
                 self._ns, n_col, int(cplx), dv.ptr(out), n_col, self._node_bases[c], _s())
        return out

    def sample(self, P, Q):
        """Embed one native Dirac spinor in component-tagged exact nodes."""
        P = self._check_coefficients(P)
        Q = self._check_coefficients(Q)
        return dv.host(self.sample_columns(P[None], Q[None]))[:, 0]

    # -- the action -------------------------------------------------------------------
    def apply_columns(self, rho, rank, out=None):
        """K_rank applied to every column of a device tensor (n_nodes, n)
        (real or complex); returns a new device tensor (or fills ``out``)."""
        torch = dv.torch()
        d = self._rank_device(rank)
        dev = self._device()
        n = self.radii.size
        if rho.dim() == 1:
            rho = rho[:, None]
        if rho.shape[0] != n:
            raise ValueError("finite complete auxiliary-node density required")
        rho = rho.contiguous()
        real = torch.view_as_real(rho).reshape(n, -1) if rho.is_complex() else rho
        n_col = real.shape[1]
        if out is None:
            out = torch.empty_like(rho)
        real_out = torch.view_as_real(out).reshape(n, -1) if out.is_complex() else out
        work = dv.empty((int(load().ax_feat_apply_workspace(self._nv, n_col)) // 8,))
        call("ax_feat_apply", dv.ptr(real), n_col, n_col, n, dv.ptr(dev["outer"]),
             dv.ptr(dev["slot"]), dv.ptr(d["rising"]), dv.ptr(d["falling"]), self._nv,
             self._ngo, dv.ptr(d["row_ptr"]), dv.ptr(d["csr_src"]), dv.ptr(d["csr_w"]),
             dv.ptr(work), dv.ptr(real_out), n_col, _s())
        return out

    def apply_density(self, density, rank):
        """Return weighted K_k density in O(nodes + same-cell edges) work;
        a (n_nodes, n) array applies K_k to every column."""
        rank = _rank(rank)
        n = self.radii.size
        if dv._is_tensor(density):
            if density.shape[0] != n or not bool(dv.torch().isfinite(density).all()):
                raise ValueError("finite complete auxiliary-node density required")
            res = self.apply_columns(density.to(device=dv.device()), rank)
            return res.reshape(density.shape)
        rho = np.asarray(density)
        rho = rho.astype(np.result_type(rho, float), copy=False)
        if rho.ndim not in (1, 2) or rho.shape[0] != n or not np.all(np.isfinite(rho)):
            raise ValueError("finite complete auxiliary-node density required")
        self.rank_arrays(rank)  # rank checks and memory cap before any device work
        res = dv.host(self.apply_columns(dv.to_dev(np.ascontiguousarray(rho)), rank))
        return res.reshape(rho.shape)

    def integrals(self, samples, quartets, rank):
        """R^k for many quartets of sampled columns: ``samples`` (n_nodes, n)
        device tensor from ``sample_columns``, ``quartets`` (q, 4) column
        indices (a, b, c, d); returns a device tensor (q,)."""
        torch = dv.torch()
        quartets = np.asarray(quartets, dtype=np.int32).reshape(-1, 4)
        nq = quartets.shape[0]
        cplx = samples.is_complex()
        n = self.radii.size
        S = samples.contiguous()
        ia = dv.i32(np.concatenate((quartets[:, 0], quartets[:, 2])))
        ib = dv.i32(np.concatenate((quartets[:, 1], quartets[:, 3])))
        dens = torch.empty((n, 2 * nq), dtype=S.dtype, device=S.device)
        call("ax_feat_products", dv.ptr(S), S.shape[1], n, 2 * nq, dv.ptr(ia), dv.ptr(ib),
             int(cplx), dv.ptr(dens), 2 * nq, _s())
        right = dens[:, nq:].contiguous()
        act = self.apply_columns(right, rank)
        out = torch.empty((nq,), dtype=S.dtype, device=S.device)
        ix = dv.i32(np.arange(nq))
        call("ax_feat_dots", dv.ptr(dens), 2 * nq, dv.ptr(act), nq, n, nq, dv.ptr(ix),
             dv.ptr(ix), int(cplx), dv.ptr(out), _s())
        return out

    def integral(self, orbitals, rank):
        """R^k(ab,cd) for four (P,Q) coefficient pairs, without conjugation."""
        if len(orbitals) != 4:
            raise ValueError("four native coefficient pairs required")
        rank = _rank(rank)
        P = np.stack([self._check_coefficients(p) for p, _ in orbitals])
        Q = np.stack([self._check_coefficients(q) for _, q in orbitals])
        self.rank_arrays(rank)
        S = self.sample_columns(P, Q)
        val = dv.host(self.integrals(S, [[0, 1, 2, 3]], rank))[0]
        return complex(val) if np.iscomplexobj(val) else float(val)


class PolynomialCellCoulombFeatures(CellCoulombFeatures):
    """Exact polynomial-coordinate transform of the native cell quadrature
    (coulomb_features.py:359-524): each component of a spinor is sampled at
    2 ks - 1 interpolation nodes per cell and the native operator becomes
    A^T K A with A the interpolation map, one dense symmetric
    (2m x 2m) same-cell block per cell."""

    def __init__(self, workspace, *, memory_bytes=128 * _MiB, rank_cache_size=8):
        super().__init__(workspace, memory_bytes=memory_bytes, rank_cache_size=rank_cache_size)
        native = dict(self.identity)
        nv, ng = self._nv, self._ng
        m = 2 * max(self._orders) - 1
        n_nodes = 2 * nv * m
        per_cell = (2 * m) * (2 * m + 1) // 2
        n_edges = nv * per_cell
        n_native = self.radii.size
        # the reference's payload bound (coulomb_features.py:380-395)
        estimate = (native["payload_estimate_bytes"] + 32 * n_native * m
                    + 8 * nv * m * sum(self._orders) + 8 * n_nodes + 16 * nv * m
                    + 16 * n_edges + 32 * self.edge_left.size
                    + rank_cache_size * 8 * (2 * nv * m + n_edges)
                    + 24 * nv * (2 * m) ** 2 + 16 * nv * ng * max(self._orders) * m)
        if estimate > memory_bytes:
            raise MemoryError(
                f"complete polynomial Coulomb payload needs {estimate} bytes; cap={memory_bytes}")
        knots = (np.asarray(workspace.get("tp", workspace["t"]), dtype=float),
                 np.asarray(workspace["t"], dtype=float))
        lo, width = self._breaks[:-1], np.diff(self._breaks)
        grid = lo[:, None] + width[:, None] * self._dens_x
        maps, tables = [], []
        start = 0
        for c in range(2):
            w_c = self._tables[c].shape[1]
            old = self.radii[start:start + nv * w_c].reshape(nv, w_c)
            maps.append(cardinal_values((old - lo[:, None]) / width[:, None], self._dens_x)
                        .reshape(nv * w_c, m))
            tables.append(_frozen(evaluate_bsplines_on_grid(
                knots[c], self._sizes[c], self._orders[c], grid, nderiv=1)[..., 0]))
            start += nv * w_c
        self._interpolation = _frozen(np.concatenate(maps))
        self._native_outer = self.outer_indices
        boundary = nv * self._tables[0].shape[1]
        groups = []
        for a in range(2):
            for b in range(2):
                idx = np.flatnonzero(((self.edge_left < boundary) == (a == 0))
                                     & ((self.edge_right < boundary) == (b == 0)))
                cell = (self.edge_left[idx] - (boundary if a else 0)) // \
                    self._tables[a].shape[1]
                groups.append(_frozen(idx[np.argsort(cell, kind="stable")].reshape(nv, -1)))
        self._native_edge_groups = tuple(groups)
        self._native_edge_left, self._native_edge_right = self.edge_left, self.edge_right
        self._tables = tuple(tables)
        self.radii = _frozen(np.concatenate((grid.ravel(), grid.ravel())))
        self.outer_indices = _frozen(np.arange(n_nodes).reshape(2, nv, m))
        tl, tr = np.triu_indices(2 * m)
        self._local_left, self._local_right = _frozen(tl), _frozen(tr)
        self._diagonal_edge = _frozen(tl == tr)
        per_cell_nodes = self.outer_indices.transpose(1, 0, 2).reshape(nv, 2 * m)
        self.edge_left = _frozen(per_cell_nodes[:, tl].ravel())
        self.edge_right = _frozen(per_cell_nodes[:, tr].ravel())
        self._node_bases = (0, nv * m)
        self._m = m
        h = hashlib.sha256(native["geometry_sha256"].encode())
        for value in (*tables, self.radii, self._interpolation, self.edge_left,
                      self.edge_right):
            h.update(repr((value.shape, value.dtype.str)).encode())
            h.update(np.ascontiguousarray(value).tobytes())
        self.identity = MappingProxyType(dict(
            native,
            representation="native cell Coulomb polynomial density coordinates",
            native_geometry_sha256=native["geometry_sha256"],
            geometry_sha256=h.hexdigest(),
            auxiliary_nodes=n_nodes,
            same_cell_edges=n_edges,
            quadrature_nodes=(nv, m),
            payload_estimate_bytes=estimate,
            density_polynomial_degree=m - 1,
            interpolation_nodes_per_component_cell=m,
            contraction="A.T K_native A weighted bilinear c-product; original quadrature retained",
        ))

    @property
    def _ngo(self):
        return self._m

    def rank_arrays(self, rank):
        rank = _rank(rank)
        if rank not in self._rank_cache:
            m, nv = self._m, self._nv
            lag = self._interpolation[self._native_outer[0]]          # (nv, ng, m)
            with np.errstate(over="raise", invalid="raise", divide="raise"):
                native_w = CellCoulombFeatures._edge_weights(self, rank)
                rising = np.einsum("cgj,cg->cj", lag, self._weights * self._gr ** rank)
                falling = np.einsum("cgj,cg->cj", lag, self._weights * self._gr ** (-rank - 1))
            local = np.zeros((nv, 2 * m, 2 * m))
            for (a, b), idx in zip(((0, 0), (0, 1), (1, 0), (1, 1)), self._native_edge_groups):
                L = self._interpolation[self._native_edge_left[idx]]   # (nv, e, m)
                R = self._interpolation[self._native_edge_right[idx]]
                term = np.swapaxes(L, 1, 2) @ (native_w[idx][..., None] * R)
                local[:, a * m:(a + 1) * m, b * m:(b + 1) * m] += term
                local[:, b * m:(b + 1) * m, a * m:(a + 1) * m] += np.swapaxes(term, 1, 2)
            packed = local[:, self._local_left, self._local_right].copy()
            packed[:, self._diagonal_edge] *= 0.5   # a diagonal edge acts in both directions
            self._rank_cache[rank] = tuple(_frozen(x) for x in (rising, falling, packed.ravel()))
            while len(self._rank_cache) > self._rank_capacity:
                self._rank_cache.popitem(last=False)
        else:
            self._rank_cache.move_to_end(rank)
        return self._rank_cache[rank]


class CellPairContraction:
    """M[m,n] = sum[r,s] R^k(mr,sn) * rho[r,s], with c-products
    (pair_contraction.py:18-148).

    Operands are complete auxiliary-node column samples ordered target-left,
    source-left, source-right, target-right; they are copied to the device
    at construction (later changes by the caller do not matter).  The
    cross-cell part uses per-cell moments of the left and right operand
    pairs with prefix/suffix sums over cells; the same-cell part runs over
    every directed edge.  Results are FP64 (complex FP64 for complex
    operands)."""

    def __init__(self, features, target_left, source_left, source_right, target_right, rank, *,
                 edge_batch_size=1024, max_work_bytes=128 * _MiB):
        _positive_int("edge_batch_size", edge_batch_size)
        _positive_int("max_work_bytes", max_work_bytes)
        arrays = tuple(np.asarray(a) for a in (target_left, source_left, source_right,
                                               target_right))
        arrays = tuple(a.astype(np.result_type(a, float), copy=False) for a in arrays)
        if any(a.ndim != 2 or a.shape[0] != len(features.radii) or a.shape[1] < 1
               or not np.all(np.isfinite(a)) for a in arrays):
            raise ValueError("finite complete auxiliary-node column operands required")
        rank = _rank(rank)
        rising, falling, edges = features.rank_arrays(rank)
        m, r, s, n = (a.shape[1] for a in arrays)
        nv, ng = features.identity["quadrature_nodes"]
        batch = min(int(edge_batch_size), len(edges))
        # the reference's work bound (pair_contraction.py:68-79)
        bound = sum(a.nbytes for a in arrays) + 16 * (
            nv * (4 * m * r + 6 * n * s + 2 * m * s) + 4 * m * n + r * s
            + batch * (2 * m + 2 * r + 4 * s + 2 * n + 4) + 2 * nv * ng * (m + r + s + n))
        if bound > max_work_bytes:
            raise MemoryError(
                f"complete cell-pair contraction needs {bound} bytes; cap={max_work_bytes}")
        torch = dv.torch()
        self._complex = any(np.iscomplexobj(a) for a in arrays)
        self._np_dtype = np.complex128 if self._complex else np.float64
        self._torch_dtype = torch.complex128 if self._complex else torch.float64
        self._features = features
        self._rank = rank
        self._dims = (m, r, s, n)
        self._nv = nv
        self._batch = batch
        fd = features._device()
        self._d = features._rank_device(rank)
        A, B, C, D = (dv.to_dev(np.ascontiguousarray(a, dtype=self._np_dtype)) for a in arrays)
        self._ops = (A, B, C, D)
        ngo = features._ngo
        cplx = int(self._complex)
        mk = lambda *shape: torch.empty(shape, dtype=self._torch_dtype, device=dv.device())  # noqa: E731
        # left moments [cell][r][m] (rising, falling) stacked for one product
        self._left = mk(2, nv, r, m)
        call("ax_cpc_moments", dv.ptr(A), m, m, dv.ptr(B), r, r, dv.ptr(fd["outer"]),
             dv.ptr(self._d["rising"]), dv.ptr(self._d["falling"]), nv, ngo, cplx,
             dv.ptr(self._left[0]), dv.ptr(self._left[1]), _s())
        rr, rf = mk(nv, s, n), mk(nv, s, n)
        call("ax_cpc_moments", dv.ptr(D), n, n, dv.ptr(C), s, s, dv.ptr(fd["outer"]),
             dv.ptr(self._d["rising"]), dv.ptr(self._d["falling"]), nv, ngo, cplx,
             dv.ptr(rr), dv.ptr(rf), _s())
        self._prefix, self._suffix = mk(nv, s, n), mk(nv, s, n)
        call("ax_cpc_scan", dv.ptr(rr), dv.ptr(rf), nv, s * n, cplx, dv.ptr(self._prefix),
             dv.ptr(self._suffix), _s())
        self._density_shape = (r, s)
        self._output_shape = (m, n)
        n_edges = self._d["n_directed"]
        self._split = int(min(256, max(1, n_edges // 512)))
        ws = int(load().ax_cpc_workspace_bytes(nv, r, s, m, n, len(features.radii), n_edges,
                                                self._split, cplx))
        self._work = dv.empty((ws // 8 + 1,))
        h = hashlib.sha256()
        for a in arrays:
            h.update(repr((a.shape, a.dtype.str)).encode())
            h.update(np.ascontiguousarray(a).tobytes())
        self.identity = MappingProxyType(dict(
            representation="exact cell pair-density matrix contraction",
            feature_geometry_sha256=features.identity["geometry_sha256"],
            operand_sha256=h.hexdigest(),
            rank=int(rank),
            density_shape=(r, s),
            output_shape=(m, n),
            edge_batch_size=batch,
            complete_work_bound_bytes=bound,
            max_work_bytes=int(max_work_bytes),
            operator="weighted native Coulomb C-product; no amplitude threshold",
        ))

    def apply_device(self, rho):
        """The action on a device tensor (r, s) of the operands' dtype;
        returns a device tensor (m, n)."""
        torch = dv.torch()
        m, r, s, n = self._dims
        A, B, C, D = self._ops
        rho = rho.to(device=dv.device(), dtype=self._torch_dtype).contiguous()
        out = torch.empty((m, n), dtype=self._torch_dtype, device=dv.device())
        d = self._d
        call("ax_cpc_apply", dv.ptr(rho), r, s, dv.ptr(A), m, m, dv.ptr(B), r, dv.ptr(C), s,
             dv.ptr(D), n, n, len(self._features.radii), self._nv, dv.ptr(self._left),
             dv.ptr(self._suffix), dv.ptr(self._prefix), dv.ptr(d["src"]), dv.ptr(d["dst"]),
             dv.ptr(d["w"]), d["n_directed"], self._split, int(self._complex),
             dv.ptr(self._work), dv.ptr(out), _s())
        return out

    def __call__(self, density):
        rho = np.asarray(density)
        rho = rho.astype(np.result_type(rho, float), copy=False)
        if rho.shape != self._density_shape or not np.all(np.isfinite(rho)):
            raise ValueError("finite complete source density matrix required")
        if np.iscomplexobj(rho) and not self._complex:
            # real operands: the action is linear in rho
            result = (dv.host(self.apply_device(dv.to_dev(np.ascontiguousarray(rho.real))))
                      + 1j * dv.host(self.apply_device(
                          dv.to_dev(np.ascontiguousarray(rho.imag)))))
        else:
            result = dv.host(self.apply_device(
                dv.to_dev(np.ascontiguousarray(rho, dtype=self._np_dtype))))
        if not np.all(np.isfinite(result)):
            raise ValueError("nonfinite native cell-pair matrix action")
        return result
