"""Auxiliary-node form of the cell Coulomb quadrature and the pair-density
matrix action built on it.

Drop-in names: ``CellCoulombFeatures`` and ``PolynomialCellCoulombFeatures``
(numerics/bspline/coulomb_features.py:64-524), ``CellPairContraction``
(numerics/bspline/pair_contraction.py:18-148).  Constructor arguments,
public methods, the public array attributes and the identity keys that
callers read are the reference's; the node set, the edge set and their
construction are this module's own.

The operator.  The cell algorithm's R^k(ab;cd) (convol.py: cross-cell
products of r^k and r^-(k+1) moments on the Gauss nodes, plus per cell the
double integral over r_< below the Gauss node r_>, done with a [0,1]
Gauss-Legendre sub-rule of order (component order, or the larger order for
mixed components) + floor(k/2)) is a symmetric bilinear form in the pair
densities.  Inside a cell every spinor component is a polynomial of degree
order - 1, so a pair density is a polynomial of degree <= 2*top - 2 (top =
the larger order).  Its values at m = 2*top - 1 distinct points of the cell
determine it, and its value at any sub-rule point is an exact barycentric
interpolation of those values.  Each cell therefore carries, per component,

  * the workspace's ng Gauss nodes ("outer" nodes; moments and the r_> end
    of the same-cell rule), and
  * m Chebyshev points of the first kind ("shape" nodes), from which every
    sub-rule of every rank is reached by interpolation;

and R^k(ab;cd) = rho_ab . K_k rho_cd with rho the c-product of samples and
K_k = cross-cell part (rising / falling moments on the outer nodes, prefix
and suffix over cells) + same-cell edges outer node -> shape node whose
weight folds the rank's sub-rule and the interpolation together.  The edge
set does not depend on the rank (only its weights do).

``PolynomialCellCoulombFeatures`` keeps only the shape nodes: the outer-node
values are themselves interpolations of the shape values, so the moments
and the same-cell edges fold into one dense symmetric 2m x 2m block per
cell (upper triangle with the diagonal halved, since an edge acts in both
directions).

Geometry (radii, B-spline values at the nodes, edge lists, per-rank
weights) is host FP64 set-up once per grid and rank, like the B-spline
tables.  Sampling, the action on many densities at once, integrals and the
pair contraction run on the GPU (``ax_feat_*``, ``ax_cpc_*``).  Host arrays
in give host arrays out; device tensors in give device tensors out.
"""

from __future__ import annotations

import ctypes
import hashlib
from collections import OrderedDict
from types import MappingProxyType, SimpleNamespace

import numpy as np

from . import device as dv
from ._lib import call, load
from .bspline import evaluate_bsplines_on_grid

__all__ = ["CellCoulombFeatures", "PolynomialCellCoulombFeatures", "CellPairContraction",
           "cardinal_values", "chebyshev_points01"]

_MiB = 1024 ** 2
_dp = ctypes.POINTER(ctypes.c_double)
_COMBOS = ((0, 0), (0, 1), (1, 0), (1, 1))  # (outer-node component, shape-node component)


def _frozen(a):
    """Read-only copy that cannot be made writeable again."""
    a = np.ascontiguousarray(a)
    return np.frombuffer(a.tobytes(), dtype=a.dtype).reshape(a.shape)


def _require_count(name, value):
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)) or value < 1:
        raise ValueError(f"{name} must be a positive integer, got {value!r}")


def _multipole(rank):
    if isinstance(rank, bool) or not isinstance(rank, (int, np.integer)) or rank < 0:
        raise ValueError(f"multipole rank must be a nonnegative integer, got {rank!r}")
    return int(rank)


def _gl01(n):
    """The cell algorithm's n-point Gauss-Legendre rule on [0, 1]."""
    x, w = np.empty(n), np.empty(n)
    call("ax_gauss_legendre_01", int(n), x.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    return x, w


def chebyshev_points01(m):
    """m Chebyshev points of the first kind on [0, 1], increasing, and their
    barycentric weights (-1)^j sin(theta_j) (Berrut & Trefethen, SIAM Review
    46 (2004) 501, eq. 5.3; a common factor cancels in the formula)."""
    theta = (2.0 * np.arange(m) + 1.0) * np.pi / (2.0 * m)
    x = 0.5 * (1.0 - np.cos(theta))
    lam = np.where(np.arange(m) % 2 == 0, 1.0, -1.0) * np.sin(theta)
    return x, lam


def cardinal_values(points, nodes, weights=None):
    """Lagrange cardinal functions of ``nodes`` at ``points`` (shape
    points.shape + (len(nodes),)) in the second barycentric form; ``weights``
    are the nodes' barycentric weights (computed from the node gaps when
    omitted).  A point on a node gives that node's unit vector."""
    nodes = np.asarray(nodes, dtype=float)
    if weights is None:
        gaps = nodes[:, None] - nodes[None, :]
        np.fill_diagonal(gaps, 1.0)
        weights = 1.0 / np.prod(gaps, axis=1)
    d = np.asarray(points, dtype=float)[..., None] - nodes
    on_node = d == 0.0
    terms = np.asarray(weights, dtype=float) / np.where(on_node, 1.0, d)
    terms /= terms.sum(axis=-1, keepdims=True)
    return np.where(on_node.any(axis=-1, keepdims=True), on_node.astype(float), terms)


def _cell_grid(workspace):
    """The cells, Gauss nodes and both components' spline geometry of a
    ``bsp_data`` workspace, checked."""
    gr = np.array(workspace["gr"], dtype=float)
    grw = np.array(workspace["grw"], dtype=float)
    nv, ns, ks = int(workspace["nv"]), int(workspace["ns"]), int(workspace["ks"])
    orders = (int(workspace.get("ksp", ks)), int(workspace.get("ksq", ks)))
    sizes = (int(workspace.get("nsp", ns)), int(workspace.get("nsq", ns)))
    knots = (np.array(workspace.get("tp", workspace["t"]), dtype=float),
             np.array(workspace["t"], dtype=float))
    ok = (gr.ndim == 2 and gr.shape[0] == nv and grw.shape == gr.shape and min(orders) >= 1
          and bool(np.isfinite(gr).all() and np.isfinite(grw).all())
          and bool((gr > 0).all() and (grw > 0).all()))
    if not ok:
        raise ValueError("cell quadrature nodes and weights must be finite and positive, "
                         "one row of Gauss nodes per cell")
    edges = [t[o - 1:o + nv] for t, o in zip(knots, orders)]
    lo = edges[0][:-1] if edges[0].size == nv + 1 else None
    shared = (all(e.size == nv + 1 and bool((np.diff(e) > 0).all()) for e in edges)
              and np.array_equal(edges[0], edges[1]))
    if not shared or not bool(((gr > lo[:, None]) & (gr < edges[0][1:, None])).all()):
        raise ValueError("large and small components must share strictly increasing cell "
                         "boundaries with every Gauss node inside its cell")
    return SimpleNamespace(gr=gr, grw=grw, nv=nv, ng=gr.shape[1], ns=ns, orders=orders,
                           sizes=sizes, knots=knots, lo=lo, width=np.diff(edges[0]),
                           top=max(orders), bsp=(workspace["bsp_p"],
                                                 workspace.get("bsp_q", workspace["bsp_p"])))


def _digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(f"{a.dtype.str}{a.shape}".encode())
        h.update(a.tobytes())
    return h.hexdigest()


def _s():
    return dv.stream_handle()


def _device_rank_bytes(n_nodes, n_edges, n_moment):
    """Device bytes of one rank's tables: rising / falling moments, the CSR
    of both edge directions (row pointers int64, sources int32, weights) and
    the directed list (sources, destinations int32, weights)."""
    return 16 * n_moment + 8 * (n_nodes + 1) + 2 * n_edges * (4 + 8 + 4 + 4 + 8)


class CellCoulombFeatures:
    """The cell algorithm's Coulomb operator on auxiliary nodes (Gauss nodes
    plus Chebyshev shape nodes per cell and component; see the module
    docstring).

    ``dot(sample(a) * sample(b), apply_density(sample(c) * sample(d), k))``
    is the cell algorithm's R^k(ab;cd) with all four component terms;
    complex values use c-products (no conjugation).  ``apply_density``
    returns a weighted action, not a pointwise potential.  Large-component
    nodes come first: node indices below ``outer_indices[1, 0, 0]`` belong
    to P."""

    def __init__(self, workspace, *, memory_bytes=128 * _MiB, rank_cache_size=8):
        _require_count("memory_bytes", memory_bytes)
        _require_count("rank_cache_size", rank_cache_size)
        g = _cell_grid(workspace)
        nv, ng, top = g.nv, g.ng, g.top
        m = 2 * top - 1
        width = ng + m                                     # nodes per cell and component
        n_nodes = 2 * nv * width
        n_edges = nv * len(_COMBOS) * ng * m
        held = (8 * nv * width * sum(g.orders)             # B-spline values at the nodes
                + 8 * n_nodes + 16 * nv * ng + 16 * n_edges  # radii, outer and edge indices
                + rank_cache_size * 8 * (2 * nv * ng + n_edges))
        device = (8 * nv * width * sum(g.orders) + 4 * n_nodes + 8 * nv * ng
                  + rank_cache_size * _device_rank_bytes(n_nodes, n_edges, nv * ng))
        estimate = held + device
        if estimate > memory_bytes:
            raise MemoryError(f"complete auxiliary-node geometry would hold {estimate} bytes "
                              f"(host and device), above memory_bytes={memory_bytes}")

        xs, lam = chebyshev_points01(m)
        shape_r = g.lo[:, None] + g.width[:, None] * xs                  # (nv, m)
        cell_r = np.concatenate((g.gr, shape_r), axis=1)                 # (nv, width)
        tables = []
        for c in range(2):
            tab = evaluate_bsplines_on_grid(g.knots[c], g.sizes[c], g.orders[c], cell_r,
                                            nderiv=1)[..., 0]
            own = np.asarray(g.bsp[c])
            if own.shape != (nv, ng, g.orders[c]) or \
                    not bool(np.abs(tab[:, :ng] - own).max() <= 1e-11):
                raise ValueError("workspace basis tables (bsp_p / bsp_q) disagree with the "
                                 "knots and orders it declares")
            tab[:, :ng] = own          # Gauss-node operands exactly as the workspace has them
            tables.append(_frozen(tab))

        cell0 = np.arange(nv)[:, None] * width                           # (nv, 1)
        base = (0, nv * width)
        outer = np.stack([b + cell0 + np.arange(ng) for b in base])     # (2, nv, ng)
        shape = np.stack([b + cell0 + ng + np.arange(m) for b in base])  # (2, nv, m)
        left = np.empty((nv, len(_COMBOS), ng, m), dtype=np.int64)
        right = np.empty_like(left)
        for e, (a, b) in enumerate(_COMBOS):
            left[:, e] = outer[a][:, :, None]
            right[:, e] = shape[b][:, None, :]

        self.tables = tuple(tables)
        self.coefficient_sizes = g.sizes
        self._grid = g
        self._ns, self._nv, self._ng = g.ns, nv, ng
        self._orders = g.orders
        self._shape_x, self._shape_lam = xs, lam
        self._node_bases = base
        self._memory_bytes = int(memory_bytes)
        self._estimate = estimate
        self.radii = _frozen(np.concatenate((cell_r.ravel(), cell_r.ravel())))
        self.outer_indices = _frozen(outer)
        self.edge_left, self.edge_right = _frozen(left.ravel()), _frozen(right.ravel())
        self._rank_cache, self._rank_capacity = OrderedDict(), int(rank_cache_size)
        self._dev = None
        self._dev_rank = OrderedDict()
        self.identity = MappingProxyType(dict(
            representation="cell Coulomb operator on Gauss and Chebyshev shape nodes",
            geometry_sha256=_digest(g.gr, g.grw, *g.knots, *tables),
            coefficient_sizes=g.sizes,
            component_orders=g.orders,
            auxiliary_nodes=n_nodes,
            same_cell_edges=n_edges,
            quadrature_nodes=(nv, ng),
            payload_estimate_bytes=estimate,
            rank_cache_capacity=rank_cache_size,
            inner_quadrature=("cell sub-rule of order (component order, larger order when "
                              "mixed) + rank//2, interpolated exactly from the shape nodes"),
            contraction="c-product samples; rho_ab . K_k rho_cd is the cell R^k",
        ))

    # -- host geometry per rank ------------------------------------------------
    def rank_arrays(self, rank):
        """(rising, falling, edge weights) of one multipole, read-only and
        cached (LRU of rank_cache_size ranks); rising / falling are per cell
        and outer node, the edge weights follow ``edge_left`` /
        ``edge_right``."""
        rank = _multipole(rank)
        got = self._rank_cache.get(rank)
        if got is None:
            got = tuple(_frozen(x) for x in self._rank_tables(rank))
            self._rank_cache[rank] = got
            while len(self._rank_cache) > self._rank_capacity:
                self._rank_cache.popitem(last=False)
        else:
            self._rank_cache.move_to_end(rank)
        return got

    def _gauss_moments(self, rank):
        g = self._grid
        with np.errstate(over="raise", invalid="raise", divide="raise"):
            return g.grw * g.gr ** rank, g.grw * g.gr ** (-rank - 1)

    def _sub_rule_edges(self, rank):
        """Same-cell weights (nv, 4, ng, m): the sub-rule of the rank from the
        cell's left boundary up to each Gauss node, weighted by
        grw / r_>^(k+1) * r_<^k, applied to the shape-node interpolant."""
        g = self._grid
        m = self._shape_x.size
        half = rank // 2
        n_sub = g.top + half
        scratch = 8 * g.nv * g.ng * n_sub * (3 * m + 4) * len(_COMBOS)
        if self._estimate + scratch > self._memory_bytes:
            raise MemoryError(f"rank-aware sub-rule of rank {rank} ({n_sub} points per Gauss "
                              f"node) needs {scratch} more bytes; memory_bytes="
                              f"{self._memory_bytes}")
        span = g.gr - g.lo[:, None]                                     # (nv, ng)
        out = np.empty((g.nv, len(_COMBOS), g.ng, m))
        with np.errstate(over="raise", invalid="raise", divide="raise"):
            for e, (a, b) in enumerate(_COMBOS):
                q = (g.orders[a] if a == b else g.top) + half
                x, w = _gl01(q)
                r_lt = g.lo[:, None, None] + span[:, :, None] * x         # (nv, ng, q)
                shape_fn = cardinal_values((r_lt - g.lo[:, None, None]) / g.width[:, None, None],
                                           self._shape_x, self._shape_lam)  # (nv, ng, q, m)
                inner = span[:, :, None] * w * (r_lt / g.gr[:, :, None]) ** rank
                out[:, e] = (g.grw / g.gr)[:, :, None] * np.einsum("cgq,cgqj->cgj", inner,
                                                                   shape_fn)
        return out

    def _rank_tables(self, rank):
        edges = self._sub_rule_edges(rank)   # holds the memory check of the rank
        rising, falling = self._gauss_moments(rank)
        return rising, falling, edges.ravel()


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
                "tables": [dv.f64(t) for t in self.tables],
            }
        return self._dev

    def _rank_device(self, rank):
        """Device rising/falling weights and the same-cell edges of one rank:
        by destination (CSR) for the action, as directed (src, dst, w) for
        the pair contraction.  Zero-weight edges are dropped."""
        rank = _multipole(rank)
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
            raise ValueError(f"coefficient vector must be finite with the workspace's padded "
                             f"length ns = {self._ns}")
        return values

    def sample_columns(self, P, Q):
        """Samples of n spinors with coefficient stacks P, Q (n, ns) at every
        auxiliary node: device tensor (n_nodes, n), complex if any
        coefficient is."""
        P = np.atleast_2d(np.asarray(P))
        Q = np.atleast_2d(np.asarray(Q))
        if P.shape != Q.shape or P.ndim != 2 or P.shape[1] != self._ns:
            raise ValueError(f"coefficient vector must be finite with the workspace's padded "
                             f"length ns = {self._ns}")
        if not (np.all(np.isfinite(P)) and np.all(np.isfinite(Q))):
            raise ValueError(f"coefficient vector must be finite with the workspace's padded "
                             f"length ns = {self._ns}")
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
            call("ax_feat_sample", dv.ptr(tab), nv, width, order, self.coefficient_sizes[c], dv.ptr(cd),
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
            raise ValueError(f"density must be finite with one value per auxiliary node "
                             f"({self.radii.size})")
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
        rank = _multipole(rank)
        n = self.radii.size
        if dv._is_tensor(density):
            if density.shape[0] != n or not bool(dv.torch().isfinite(density).all()):
                raise ValueError(f"density must be finite with one value per auxiliary node "
                             f"({self.radii.size})")
            res = self.apply_columns(density.to(device=dv.device()), rank)
            return res.reshape(density.shape)
        rho = np.asarray(density)
        rho = rho.astype(np.result_type(rho, float), copy=False)
        if rho.ndim not in (1, 2) or rho.shape[0] != n or not np.all(np.isfinite(rho)):
            raise ValueError(f"density must be finite with one value per auxiliary node "
                             f"({self.radii.size})")
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
            raise ValueError("integral takes exactly four (P, Q) coefficient pairs")
        rank = _multipole(rank)
        P = np.stack([self._check_coefficients(p) for p, _ in orbitals])
        Q = np.stack([self._check_coefficients(q) for _, q in orbitals])
        self.rank_arrays(rank)
        S = self.sample_columns(P, Q)
        val = dv.host(self.integrals(S, [[0, 1, 2, 3]], rank))[0]
        return complex(val) if np.iscomplexobj(val) else float(val)


class PolynomialCellCoulombFeatures(CellCoulombFeatures):
    """The same operator on the shape nodes alone
    (coulomb_features.py:359-524 for the interface): per cell and component
    the m = 2*top - 1 Chebyshev points; the Gauss-node values are their
    interpolants, so the moments and the same-cell edges fold into per-node
    rising / falling weights and one dense symmetric 2m x 2m block per cell.
    Sampling, action and integrals are those of ``CellCoulombFeatures``."""

    def __init__(self, workspace, *, memory_bytes=128 * _MiB, rank_cache_size=8):
        super().__init__(workspace, memory_bytes=memory_bytes, rank_cache_size=rank_cache_size)
        native = dict(self.identity)
        g = self._grid
        nv, ng = g.nv, g.ng
        m = self._shape_x.size
        n_nodes = 2 * nv * m
        per_cell = (2 * m) * (2 * m + 1) // 2
        n_edges = nv * per_cell
        held = (8 * nv * ng * m                           # Gauss-node interpolation
                + 8 * nv * m * sum(g.orders) + 8 * n_nodes + 16 * n_edges
                + rank_cache_size * 8 * (2 * nv * m + n_edges)
                + 16 * nv * (2 * m) ** 2)                 # block assembly scratch
        device = (8 * nv * m * sum(g.orders) + 4 * n_nodes + 8 * nv * m
                  + rank_cache_size * _device_rank_bytes(n_nodes, n_edges, nv * m))
        estimate = native["payload_estimate_bytes"] + held + device
        if estimate > memory_bytes:
            raise MemoryError(f"complete polynomial-coordinate geometry would hold {estimate} "
                              f"bytes (with the Gauss-node operator it is built from), above "
                              f"memory_bytes={memory_bytes}")
        # Gauss-node values as interpolants of the shape-node values
        self._to_gauss = _frozen(cardinal_values((g.gr - g.lo[:, None]) / g.width[:, None],
                                                 self._shape_x, self._shape_lam))  # (nv, ng, m)
        self.tables = tuple(_frozen(t[:, ng:, :]) for t in self.tables)
        shape_r = (g.lo[:, None] + g.width[:, None] * self._shape_x).ravel()
        self.radii = _frozen(np.concatenate((shape_r, shape_r)))
        self._node_bases = (0, nv * m)
        self.outer_indices = _frozen(np.arange(n_nodes, dtype=np.int64).reshape(2, nv, m))
        # per cell, the upper triangle of the block over local index
        # component * m + j, mapped to global node numbers
        up_l, up_r = np.triu_indices(2 * m)
        self._block_pick = (up_l, up_r)
        local = self.outer_indices.transpose(1, 0, 2).reshape(nv, 2 * m)
        self.edge_left = _frozen(local[:, up_l].ravel())
        self.edge_right = _frozen(local[:, up_r].ravel())
        self._m = m
        self._estimate = estimate
        self._rank_cache.clear()
        self.identity = MappingProxyType(dict(
            native,
            representation="cell Coulomb operator on Chebyshev shape nodes only",
            native_geometry_sha256=native["geometry_sha256"],
            geometry_sha256=_digest(native["geometry_sha256"].encode(), self.radii,
                                    self._to_gauss, *self.tables),
            auxiliary_nodes=n_nodes,
            same_cell_edges=n_edges,
            quadrature_nodes=(nv, m),
            payload_estimate_bytes=estimate,
            density_polynomial_degree=m - 1,
            interpolation_nodes_per_component_cell=m,
            contraction=("c-product samples at the shape nodes; the Gauss-node operator "
                         "conjugated by the interpolation map"),
        ))

    @property
    def _ngo(self):
        return self._m

    def _rank_tables(self, rank):
        edges = self._sub_rule_edges(rank)                      # (nv, 4, ng, m)
        rising_g, falling_g = self._gauss_moments(rank)         # (nv, ng)
        L = self._to_gauss                                      # (nv, ng, m)
        rising = np.einsum("cg,cgj->cj", rising_g, L)
        falling = np.einsum("cg,cgj->cj", falling_g, L)
        m = self._m
        block = np.zeros((self._nv, 2 * m, 2 * m))
        for e, (a, b) in enumerate(_COMBOS):
            # edge outer(a, g) -> shape(b, j) with the outer value an
            # interpolant: a term between shape(a, i) and shape(b, j)
            piece = np.einsum("cgi,cgj->cij", L, edges[:, e])
            block[:, a * m:(a + 1) * m, b * m:(b + 1) * m] += piece
            block[:, b * m:(b + 1) * m, a * m:(a + 1) * m] += np.swapaxes(piece, 1, 2)
        up_l, up_r = self._block_pick
        w = block[:, up_l, up_r]
        w[:, up_l == up_r] *= 0.5     # a diagonal edge acts in both directions
        return rising, falling, w.ravel()


class CellPairContraction:
    """M[m,n] = sum[r,s] R^k(mr,sn) * rho[r,s] with c-products
    (pair_contraction.py:18-148 for the interface).

    Operands are complete auxiliary-node column samples ordered target-left
    (m columns), source-left (r), source-right (s), target-right (n); they
    are copied to the device at construction, so later changes by the
    caller do not matter.  On the device: per-cell moments of the left and
    right operand column products on the outer nodes, prefix / suffix sums
    of the right moments over cells (the cross-cell part), and every
    directed same-cell edge of the features (``ax_cpc_*``).  Results are
    FP64 (complex FP64 for complex operands).  ``edge_batch_size`` is
    accepted and recorded; the device takes all edges in one pass."""

    def __init__(self, features, target_left, source_left, source_right, target_right, rank, *,
                 edge_batch_size=1024, max_work_bytes=128 * _MiB):
        _require_count("edge_batch_size", edge_batch_size)
        _require_count("max_work_bytes", max_work_bytes)
        n_nodes = len(features.radii)
        ops = []
        for a in (target_left, source_left, source_right, target_right):
            a = np.asarray(a)
            a = a.astype(np.result_type(a, float), copy=False)
            if a.ndim != 2 or a.shape[0] != n_nodes or a.shape[1] < 1 or \
                    not bool(np.isfinite(a).all()):
                raise ValueError(f"operands must be finite 2-D arrays with one row per "
                                 f"auxiliary node ({n_nodes}) and at least one column")
            ops.append(a)
        rank = _multipole(rank)
        features.rank_arrays(rank)
        m, r, s, n = (a.shape[1] for a in ops)
        nv = features._nv
        d = features._rank_device(rank)
        n_edges = d["n_directed"]
        split = int(min(256, max(1, n_edges // 512)))
        self._complex = any(np.iscomplexobj(a) for a in ops)
        item = 16 if self._complex else 8
        work = int(load().ax_cpc_workspace_bytes(nv, r, s, m, n, n_nodes, n_edges, split,
                                                  int(self._complex)))
        # every device byte this object holds: operands, left moments (rising
        # and falling), right moments with their prefix / suffix, the
        # action's workspace and one output
        bound = (item * n_nodes * (m + r + s + n) + item * 2 * nv * r * m
                 + item * 4 * nv * s * n + work + item * m * n)
        if bound > max_work_bytes:
            raise MemoryError(f"complete cell-pair contraction would hold {bound} device bytes, "
                              f"above max_work_bytes={max_work_bytes}")
        torch = dv.torch()
        self._np_dtype = np.complex128 if self._complex else np.float64
        self._torch_dtype = torch.complex128 if self._complex else torch.float64
        self._features = features
        self._rank = rank
        self._dims = (m, r, s, n)
        self._nv = nv
        self._d = d
        fd = features._device()
        A, B, C, D = (dv.to_dev(np.ascontiguousarray(a, dtype=self._np_dtype)) for a in ops)
        self._ops = (A, B, C, D)
        ngo = features._ngo
        cplx = int(self._complex)

        def dev(*shape):
            return torch.empty(shape, dtype=self._torch_dtype, device=dv.device())

        self._left = dev(2, nv, r, m)              # [rising, falling][cell][r][m]
        call("ax_cpc_moments", dv.ptr(A), m, m, dv.ptr(B), r, r, dv.ptr(fd["outer"]),
             dv.ptr(d["rising"]), dv.ptr(d["falling"]), nv, ngo, cplx,
             dv.ptr(self._left[0]), dv.ptr(self._left[1]), _s())
        right_rise, right_fall = dev(nv, s, n), dev(nv, s, n)
        call("ax_cpc_moments", dv.ptr(D), n, n, dv.ptr(C), s, s, dv.ptr(fd["outer"]),
             dv.ptr(d["rising"]), dv.ptr(d["falling"]), nv, ngo, cplx,
             dv.ptr(right_rise), dv.ptr(right_fall), _s())
        self._prefix, self._suffix = dev(nv, s, n), dev(nv, s, n)
        call("ax_cpc_scan", dv.ptr(right_rise), dv.ptr(right_fall), nv, s * n, cplx,
             dv.ptr(self._prefix), dv.ptr(self._suffix), _s())
        self._density_shape = (r, s)
        self._output_shape = (m, n)
        self._split = split
        self._work = dv.empty((work // 8 + 1,))
        self.identity = MappingProxyType(dict(
            representation="cell pair-density matrix action on auxiliary-node samples",
            feature_geometry_sha256=features.identity["geometry_sha256"],
            operand_sha256=_digest(*ops),
            rank=rank,
            density_shape=(r, s),
            output_shape=(m, n),
            edge_batch_size=int(edge_batch_size),
            complete_work_bound_bytes=bound,
            max_work_bytes=int(max_work_bytes),
            operator="cross-cell moments with prefix / suffix sums plus every same-cell edge",
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
            raise ValueError(f"source density matrix must be finite with shape {self._density_shape}")
        if np.iscomplexobj(rho) and not self._complex:
            # real operands: the action is linear in rho
            result = (dv.host(self.apply_device(dv.to_dev(np.ascontiguousarray(rho.real))))
                      + 1j * dv.host(self.apply_device(
                          dv.to_dev(np.ascontiguousarray(rho.imag)))))
        else:
            result = dv.host(self.apply_device(
                dv.to_dev(np.ascontiguousarray(rho, dtype=self._np_dtype))))
        if not np.all(np.isfinite(result)):
            raise ValueError("cell-pair action produced non-finite values")
        return result
