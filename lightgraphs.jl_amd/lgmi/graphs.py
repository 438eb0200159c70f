"""SimpleGraph / SimpleDiGraph mirrors backed by a sorted Int32 CSR.

The reference stores ``fadjlist::Vector{Vector{T}}`` (and ``badjlist`` for
digraphs) with every row sorted ascending
(src/SimpleGraphs/simplegraph.jl:8-16, simpledigraph.jl:8-22).  Here the same
adjacency is a 0-based CSR (``row_ptr``, ``col``) built by liblgmi's
``lgmi_csr_build``; ``T`` is the reference's eltype (Int64, UInt8, Int16 ...)
and decides the dtype and ``typemax`` sentinel of traversal results.
Vertex ids in the public API are 1-based, as in Julia.
"""
from __future__ import annotations

import atexit
import weakref

import numpy as np

from . import _lib

# Graphs that hold a device handle.  At interpreter exit every handle still
# alive is destroyed while the HIP runtime (and a profiler attached to it)
# is fully up, instead of whenever a late __del__ or the runtime's own
# teardown reaches it.
_live: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _release_all():
    for g in list(_live):
        try:
            g._release()
        except Exception:
            pass


def _as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def csr_from_edges(n: int, src, dst, symmetrize: bool, drop_self_loops: bool = False):
    """Sorted, de-duplicated CSR from 0-based endpoint arrays."""
    lib = _lib.load()
    src, dst = _as_i32(src), _as_i32(dst)
    ne = len(src)
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    cap = max(1, 2 * ne if symmetrize else ne)
    col = np.empty(cap, dtype=np.int32)
    nnz = np.zeros(1, dtype=np.int64)
    rc = lib.lgmi_csr_build(n, ne, _lib.i32p(src), _lib.i32p(dst), int(symmetrize), int(drop_self_loops),
                            _lib.i32p(row_ptr), _lib.i32p(col), cap, _lib.i64p(nnz))
    if rc != 0:
        raise ValueError(f"lgmi_csr_build failed ({rc})")
    return row_ptr, col[: int(nnz[0])].copy()


def csr_transpose(row_ptr, col):
    lib = _lib.load()
    n = len(row_ptr) - 1
    t_rp = np.zeros(n + 1, dtype=np.int32)
    t_col = np.empty(max(1, len(col)), dtype=np.int32)
    rc = lib.lgmi_csr_transpose(n, _lib.i32p(_as_i32(row_ptr)), _lib.i32p(_as_i32(col)), _lib.i32p(t_rp),
                                _lib.i32p(t_col))
    if rc != 0:
        raise ValueError(f"lgmi_csr_transpose failed ({rc})")
    return t_rp, t_col[: len(col)].copy()


class _GraphBase:
    directed = False

    def __init__(self, n: int = 0, T=np.int64):
        self.T = np.dtype(T)
        self._n = int(n)
        self._src: list[int] = []
        self._dst: list[int] = []
        self._row_ptr = np.zeros(self._n + 1, dtype=np.int32)
        self._col = np.zeros(0, dtype=np.int32)
        self._in_row_ptr = None
        self._in_col = None
        self._dirty = False
        self._handle = None

    # -- construction -------------------------------------------------
    @classmethod
    def from_csr(cls, row_ptr, col, T=np.int64, in_row_ptr=None, in_col=None):
        g = cls(len(row_ptr) - 1, T=T)
        g._row_ptr = _as_i32(row_ptr)
        g._col = _as_i32(col)
        if cls.directed:
            if in_row_ptr is None:
                in_row_ptr, in_col = csr_transpose(g._row_ptr, g._col)
            g._in_row_ptr, g._in_col = _as_i32(in_row_ptr), _as_i32(in_col)
        return g

    @classmethod
    def from_edge_list(cls, src, dst, T=np.int64, one_based: bool = True, nv_min: int = 0):
        """SimpleGraph(edge_list) / SimpleDiGraph(edge_list)
        (src/SimpleGraphs/simplegraph.jl:208-244, simpledigraph.jl:212-253)
        built on the device (lgmi_graph_from_edges): nv is the largest id
        used (at least nv_min), edges with an id below 1 are skipped, rows
        are sorted and de-duplicated.  Returns the graph, whose device handle
        is already built, and the reference's ne."""
        src = np.asarray(src, dtype=np.int64)
        dst = np.asarray(dst, dtype=np.int64)
        if one_based:
            src, dst = src - 1, dst - 1
        from . import tuning
        tuning.apply_once()
        lib = _lib.load()
        h = _lib.ctypes.c_void_p()
        ne = np.zeros(1, dtype=np.int64)
        s32 = _as_i32(np.where(src < 0, -1, src))
        d32 = _as_i32(np.where(dst < 0, -1, dst))
        _lib.check(lib.lgmi_graph_from_edges(_lib.i32p(_nonempty(s32)), _lib.i32p(_nonempty(d32)), len(s32),
                                             int(cls.directed), int(nv_min), _lib.ctypes.byref(h), _lib.i64p(ne)))
        nv = np.zeros(1, dtype=np.int32)
        nnz = np.zeros(1, dtype=np.int64)
        _lib.check(lib.lgmi_graph_size(h, _lib.i32p(nv), _lib.i64p(nnz)))
        n, m = int(nv[0]), int(nnz[0])
        rp = np.zeros(n + 1, dtype=np.int32)
        col = np.zeros(max(m, 1), dtype=np.int32)
        irp = np.zeros(n + 1, dtype=np.int32) if cls.directed else None
        icol = np.zeros(max(m, 1), dtype=np.int32) if cls.directed else None
        _lib.check(lib.lgmi_graph_csr(h, _lib.i32p(rp), _lib.i32p(col), _lib.i32p(irp) if cls.directed else None,
                                      _lib.i32p(icol) if cls.directed else None))
        g = cls.from_csr(rp, col[:m], T=T, in_row_ptr=irp, in_col=icol[:m] if cls.directed else None)
        g._handle = h
        _live.add(g)
        return g, int(ne[0])

    @classmethod
    def from_edges(cls, n: int, src, dst, T=np.int64, one_based: bool = False, drop_self_loops=False):
        src = np.asarray(src, dtype=np.int64)
        dst = np.asarray(dst, dtype=np.int64)
        if one_based:
            src, dst = src - 1, dst - 1
        rp, col = csr_from_edges(n, src, dst, symmetrize=not cls.directed, drop_self_loops=drop_self_loops)
        return cls.from_csr(rp, col, T=T)

    def add_edge(self, u: int, v: int) -> bool:
        """add_edge!(g, u, v) with 1-based ids (simplegraph.jl:445-461)."""
        if not (1 <= u <= self._n and 1 <= v <= self._n):
            return False
        if self.has_edge(u, v):
            return False
        self._src.append(u - 1)
        self._dst.append(v - 1)
        self._dirty = True
        self._release()
        return True

    def _flush(self):
        if not self._dirty:
            return
        rp, col = self._row_ptr, self._col
        # existing edges + pending edges -> CSR
        src_old = np.repeat(np.arange(self._n, dtype=np.int32), np.diff(rp))
        src = np.concatenate([src_old, np.asarray(self._src, dtype=np.int32)])
        dst = np.concatenate([col, np.asarray(self._dst, dtype=np.int32)])
        self._row_ptr, self._col = csr_from_edges(self._n, src, dst, symmetrize=not self.directed)
        if self.directed:
            self._in_row_ptr, self._in_col = csr_transpose(self._row_ptr, self._col)
        self._src, self._dst = [], []
        self._dirty = False

    # -- graph interface ------------------------------------------------
    def nv(self) -> int:
        return self._n

    def ne(self) -> int:
        self._flush()
        if self.directed:
            return int(len(self._col))
        loops = int(np.sum(self._col == np.repeat(np.arange(self._n), np.diff(self._row_ptr))))
        return (int(len(self._col)) - loops) // 2 + loops

    def is_directed(self) -> bool:
        return self.directed

    def csr(self):
        self._flush()
        return self._row_ptr, self._col

    def in_csr(self):
        self._flush()
        if self.directed:
            if self._in_row_ptr is None:
                self._in_row_ptr, self._in_col = csr_transpose(self._row_ptr, self._col)
            return self._in_row_ptr, self._in_col
        return self._row_ptr, self._col

    def outneighbors(self, v: int):
        rp, col = self.csr()
        return (col[rp[v - 1]:rp[v]] + 1).astype(self.T)

    def inneighbors(self, v: int):
        rp, col = self.in_csr()
        return (col[rp[v - 1]:rp[v]] + 1).astype(self.T)

    def has_edge(self, u: int, v: int) -> bool:
        rp, col = self.csr()
        row = col[rp[u - 1]:rp[u]]
        i = np.searchsorted(row, v - 1)
        if i < len(row) and row[i] == v - 1:
            return True
        for a, b in zip(self._src, self._dst):
            if (a, b) == (u - 1, v - 1) or (not self.directed and (a, b) == (v - 1, u - 1)):
                return True
        return False

    def degree_array(self):
        """degree(g, v) for all v (src/core.jl:124-128): out-degree for
        undirected graphs, in + out for directed ones."""
        rp, _ = self.csr()
        d = np.diff(rp).astype(np.int64)
        if self.directed:
            d = d + np.diff(self.in_csr()[0]).astype(np.int64)
        return d

    def with_eltype(self, T):
        """Graph{T}(g): the same graph with another eltype."""
        rp, col = self.csr()
        if self.directed:
            irp, icol = self.in_csr()
            return type(self).from_csr(rp, col, T=T, in_row_ptr=irp, in_col=icol)
        return type(self).from_csr(rp, col, T=T)

    # -- device handle ----------------------------------------------------
    def handle(self):
        """Device CSR for liblgmi (created once, dropped on mutation)."""
        self._flush()
        if self._handle is None:
            from . import tuning
            tuning.apply_once()
            lib = _lib.load()
            h = _lib.ctypes.c_void_p()
            rp, col = self._row_ptr, self._col
            if self.directed:
                irp, icol = self.in_csr()
                rc = lib.lgmi_graph_create(self._n, len(col), _lib.i32p(rp), _lib.i32p(_nonempty(col)),
                                           len(icol), _lib.i32p(irp), _lib.i32p(_nonempty(icol)), 1,
                                           _lib.ctypes.byref(h))
            else:
                rc = lib.lgmi_graph_create(self._n, len(col), _lib.i32p(rp), _lib.i32p(_nonempty(col)), 0,
                                           None, None, 0, _lib.ctypes.byref(h))
            _lib.check(rc)
            self._handle = h
            _live.add(self)
        return self._handle

    def _release(self):
        if self._handle is not None:
            try:
                _lib.load().lgmi_graph_destroy(self._handle)
            finally:
                self._handle = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def __repr__(self):
        kind = "directed" if self.directed else "undirected"
        return f"{{{self.nv()}, {self.ne()}}} {kind} simple {self.T.name} graph"


def _nonempty(a):
    return a if len(a) else np.zeros(1, dtype=np.int32)


class SimpleGraph(_GraphBase):
    """Undirected simple graph (src/SimpleGraphs/simplegraph.jl:8-16)."""
    directed = False

    @classmethod
    def from_digraph(cls, g: "SimpleDiGraph", T=None):
        """SimpleGraph(g::SimpleDiGraph): every arc becomes an edge."""
        rp, col = g.csr()
        src = np.repeat(np.arange(g.nv(), dtype=np.int32), np.diff(rp))
        return cls.from_edges(g.nv(), src, col, T=T or g.T)


class SimpleDiGraph(_GraphBase):
    """Directed simple graph (src/SimpleGraphs/simpledigraph.jl:8-22)."""
    directed = True
