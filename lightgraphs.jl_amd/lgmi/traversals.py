"""gdistances / gdistances! / bfs_parents / bfs_tree on the GPU.

Mirrors src/traversals/bfs.jl:10-141 with unchanged argument meaning:
vertex ids are 1-based, distances of unreached vertices are typemax(T),
unreached parents are 0 and parent[s] = s.  The search itself runs in
liblgmi's direction-optimising BFS; there is no CPU path.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .graphs import SimpleDiGraph, _GraphBase


def typemax(T) -> int:
    return int(np.iinfo(np.dtype(T)).max)


def _sources(g: _GraphBase, source) -> np.ndarray:
    """`source` may be an Integer or a collection (bfs.jl:103)."""
    if isinstance(source, (int, np.integer)):
        arr = np.array([int(source)], dtype=np.int64)
    else:
        arr = np.asarray(list(source), dtype=np.int64)
    n = g.nv()
    if arr.size and (arr.min() < 1 or arr.max() > n):
        raise IndexError(f"source vertex out of range 1:{n}")
    return np.ascontiguousarray(arr - 1, dtype=np.int32)


def _device_dist(g: _GraphBase, srcs0: np.ndarray) -> np.ndarray:
    n = g.nv()
    dist = np.empty(max(n, 1), dtype=np.int32)
    if n == 0:
        return dist[:0]
    lib = _lib.load()
    _lib.check(lib.lgmi_gdistances(g.handle(), _lib.i32p(srcs0 if len(srcs0) else np.zeros(1, np.int32)),
                                   len(srcs0), _lib.i32p(dist)))
    return dist


def gdistances_(g: _GraphBase, source, vert_level: np.ndarray, sort_alg=None) -> np.ndarray:
    """gdistances!(g, source, vert_level; sort_alg) -- bfs.jl:95-124.
    Writes only the reached entries (sources get 0); `sort_alg` is accepted
    and changes nothing."""
    srcs = _sources(g, source)
    dist = _device_dist(g, srcs)
    reached = dist >= 0
    vert_level[reached] = dist[reached].astype(vert_level.dtype)
    return vert_level


def gdistances(g: _GraphBase, source, sort_alg=None) -> np.ndarray:
    """gdistances(g, source; sort_alg) -- bfs.jl:141."""
    out = np.full(g.nv(), typemax(g.T), dtype=g.T)
    return gdistances_(g, source, out, sort_alg=sort_alg)


def _parents0(g: _GraphBase, s: int, dir_in: int) -> np.ndarray:
    n = g.nv()
    if not (1 <= int(s) <= n):
        raise IndexError(f"source vertex {s} out of range 1:{n}")
    lib = _lib.load()
    par = np.empty(n, dtype=np.int32)
    _lib.check(lib.lgmi_bfs_parents(g.handle(), int(s) - 1, int(dir_in), _lib.i32p(par)))
    return par


def bfs_parents(g: _GraphBase, s: int, dir="out") -> np.ndarray:
    """bfs_parents(g, s; dir=:out) -- bfs.jl:35-66.  `dir` other than
    "out" follows in-neighbours, as the reference's ternary does."""
    par = _parents0(g, s, 0 if dir in ("out", ":out") else 1)
    return np.where(par >= 0, par + 1, 0).astype(g.T)


def _device_many(g: _GraphBase, roots0: np.ndarray, dir_in: int, parents: bool, chunk: int = 256):
    """lgmi_bfs_run_many: 0-based distances and parents (-1 unreached) of
    one single-source search per root, rows in root order.  Up to `chunk`
    roots per call, into device arrays reused from call to call."""
    roots0 = np.ascontiguousarray(roots0, dtype=np.int32)
    n, k = g.nv(), len(roots0)
    lib = _lib.load()
    h = g.handle()
    if k == 0:
        _lib.check(lib.lgmi_bfs_run_many(h, None, 0, int(dir_in), None, None))
        return np.empty((0, n), np.int32), (np.empty((0, n), np.int32) if parents else None)
    m = max(1, min(k, chunk))
    d = _lib.DeviceBuffer(4 * n * m)
    p = _lib.DeviceBuffer(4 * n * m) if parents else None
    dist = np.empty((k, n), np.int32)
    par = np.empty((k, n), np.int32) if parents else None
    try:
        for s0 in range(0, k, m):
            part = np.ascontiguousarray(roots0[s0:s0 + m])
            _lib.check(lib.lgmi_bfs_run_many(h, _lib.i32p(part), len(part), int(dir_in), d.ptr,
                                             p.ptr if p is not None else None))
            dist[s0:s0 + len(part)] = d.to_host(np.int32, (len(part), n))
            if p is not None:
                par[s0:s0 + len(part)] = p.to_host(np.int32, (len(part), n))
    finally:
        d.free()
        if p is not None:
            p.free()
    return dist, par


def bfs_many(g: _GraphBase, sources, dir="out"):
    """The loop over sources of gdistances(g, s) and bfs_parents(g, s; dir)
    (bfs.jl:35-66, 95-141) in one call, several searches at a time on the
    GPU.  Returns (dists, parents), one row per source: distances as
    gdistances gives them (typemax(T) unreached) and parents as bfs_parents
    does (0 unreached)."""
    srcs = np.atleast_1d(np.asarray(sources, dtype=np.int64))
    n = g.nv()
    if srcs.size and (srcs.min() < 1 or srcs.max() > n):
        raise IndexError(f"source vertex out of range 1:{n}")
    dist, par = _device_many(g, srcs - 1, 0 if dir in ("out", ":out") else 1, True)
    dists = dist.astype(g.T)
    dists[dist < 0] = typemax(g.T)
    parents = np.where(par >= 0, par.astype(g.T) + 1, 0).astype(g.T)
    return dists, parents


def tree(parents: np.ndarray) -> SimpleDiGraph:
    """tree(parents) -- bfs.jl:10-19: arcs u -> v for every parent u > 0,
    u != v."""
    parents = np.asarray(parents)
    n = len(parents)
    v = np.arange(1, n + 1)
    keep = (parents > 0) & (parents != v)
    return SimpleDiGraph.from_edges(n, parents[keep].astype(np.int64) - 1, v[keep] - 1, T=parents.dtype)


def bfs_tree(g: _GraphBase, s: int, dir="out") -> SimpleDiGraph:
    """bfs_tree(g, s; dir) -- bfs.jl:76."""
    return tree(bfs_parents(g, s, dir=dir))
