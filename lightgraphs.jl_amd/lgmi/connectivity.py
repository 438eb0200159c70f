"""connected_components, is_connected, has_path and period on the GPU.

Mirrors src/connectivity.jl:13-179 and 495-516 and
src/traversals/bfs.jl:150-172 (1-based vertex ids).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .graphs import _GraphBase


def connected_components_(label: np.ndarray, g: _GraphBase) -> np.ndarray:
    """connected_components!(label, g) (connectivity.jl:13-34): every vertex
    gets the smallest (1-based) vertex id of its component -- weak
    components for a directed graph."""
    n = g.nv()
    if n:
        lab = np.empty(n, dtype=np.int32)
        _lib.check(_lib.load().lgmi_connected_components(g.handle(), _lib.i32p(lab)))
        label[:] = lab.astype(label.dtype) + 1
    return label


def components(labels) -> tuple[list, dict]:
    """components(labels) (connectivity.jl:52-64): component vertex lists
    in order of first appearance, and the map label -> list index
    (1-based)."""
    labels = np.asarray(labels)
    d: dict = {}
    c: list = []
    if len(labels) == 0:
        return c, d
    order = np.argsort(labels, kind="stable")
    first = {}
    for v, l in enumerate(labels.tolist()):
        if l not in first:
            first[l] = v
    keys = sorted(first, key=lambda k: first[k])
    groups = np.split(order, np.nonzero(np.diff(labels[order]))[0] + 1)
    by_label = {int(labels[gr[0]]): gr for gr in groups}
    for i, k in enumerate(keys, start=1):
        d[k] = i
        c.append((by_label[int(k)] + 1).astype(np.int64).tolist())
    return c, d


def components_dict(labels) -> dict:
    """components_dict(labels) (connectivity.jl:37-47)."""
    c, d = components(labels)
    return {k: c[i - 1] for k, i in d.items()}


def connected_components(g: _GraphBase) -> list:
    """connected_components(g) (connectivity.jl:89-95)."""
    label = connected_components_(np.zeros(g.nv(), dtype=np.int64), g)
    return components(label)[0]


def is_connected(g: _GraphBase) -> bool:
    """is_connected(g) (connectivity.jl:131-134): one (weak) component; an
    empty graph has none and is not connected."""
    n = g.nv()
    if n == 0:
        return False
    label = connected_components_(np.zeros(n, dtype=np.int64), g)
    return bool((label == 1).all())


def is_weakly_connected(g: _GraphBase) -> bool:
    """is_weakly_connected(g) (connectivity.jl:179)."""
    return is_connected(g)


def has_path(g: _GraphBase, u: int, v: int, exclude_vertices=()) -> bool:
    """has_path(g, u, v; exclude_vertices) (traversals/bfs.jl:150-172)."""
    n = g.nv()
    for x in (u, v):
        if not 1 <= int(x) <= n:
            raise IndexError(f"vertex {x} out of range 1:{n}")
    ex = np.asarray(list(exclude_vertices), dtype=np.int64)
    if ex.size and (ex.min() < 1 or ex.max() > n):
        raise IndexError(f"excluded vertex out of range 1:{n}")
    ex0 = np.ascontiguousarray(ex - 1, dtype=np.int32)
    res = ctypes.c_int(0)
    _lib.check(_lib.load().lgmi_has_path(g.handle(), int(u) - 1, int(v) - 1,
                                         _lib.i32p(ex0 if ex0.size else np.zeros(1, np.int32)), len(ex0),
                                         ctypes.byref(res)))
    return bool(res.value)


def period(g: _GraphBase) -> int:
    """period(g) for a strongly connected directed graph
    (connectivity.jl:495-516); ValueError (the reference's ArgumentError)
    otherwise."""
    if not g.is_directed():
        raise TypeError("period is defined for directed graphs")
    out = np.zeros(1, dtype=np.int64)
    _lib.check(_lib.load().lgmi_period(g.handle(), _lib.i64p(out)))
    return int(out[0])


def neighborhood_dists(g: _GraphBase, v: int, d, distmx=None, dir: str = "out") -> list:
    """neighborhood_dists(g, v, d; dir) (connectivity.jl:681-702) with unit
    lengths: (vertex, distance) for every vertex within distance d of v, in
    the reference's FIFO queue order; dir=:in follows in-neighbours."""
    import math
    if distmx is not None:
        raise NotImplementedError("weighted distmx is not on the accelerated path")
    n = g.nv()
    if not 1 <= int(v) <= n:
        raise IndexError(f"vertex {v} out of range 1:{n}")
    if d < 0:
        return []
    dd = n if math.isinf(d) else int(math.floor(d))
    verts = np.empty(n, dtype=np.int32)
    dists = np.empty(n, dtype=np.int32)
    cnt = np.zeros(1, dtype=np.int32)
    _lib.check(_lib.load().lgmi_neighborhood(g.handle(), int(v) - 1, int(min(dd, n)), int(dir in ("in", ":in")),
                                             _lib.i32p(verts), _lib.i32p(dists), _lib.i32p(cnt)))
    k = int(cnt[0])
    return list(zip((verts[:k].astype(np.int64) + 1).tolist(), dists[:k].astype(np.int64).tolist()))


def neighborhood(g: _GraphBase, v: int, d, distmx=None, dir: str = "out") -> list:
    """neighborhood(g, v, d; dir) (connectivity.jl:634-635)."""
    return [x for x, _ in neighborhood_dists(g, v, d, distmx, dir)]
