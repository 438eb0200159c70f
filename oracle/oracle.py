"""CPU oracle wrapper (TEST INFRASTRUCTURE ONLY).

ctypes bindings for ``oracle/liblgoracle.so``, the plain-C restatement of the
reference's BFS and exact-Brandes functions (LightGraphs.jl v1.3.5):

* ``gdistances``            src/traversals/bfs.jl:95-141
* ``bfs_parents``           src/traversals/bfs.jl:35-66
* ``par_gdistances``        src/Parallel/traversals/gdistances.jl:36-139
* ``par_bfs_tree``          src/Parallel/traversals/bfs.jl:49-104
* ``dijkstra_allpaths``     src/shortestpaths/dijkstra.jl:59-141
* ``betweenness``           src/centrality/betweenness.jl:46-160
* ``betweenness_threaded``  src/Parallel/centrality/betweenness.jl:42-75

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker or the timed CPU
baseline.  Vertices are 0-based; unreached distances/parents are -1.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblgoracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_gdistances.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, ctypes.c_int32, _i32p]
    lib.orc_bfs_parents.argtypes = [ctypes.c_int32, _i32p, _i32p, ctypes.c_int32, _i32p]
    lib.orc_par_gdistances.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, ctypes.c_int32, _i32p]
    lib.orc_par_bfs_tree.argtypes = [ctypes.c_int32, _i32p, _i32p, ctypes.c_int32, _i32p]
    lib.orc_par_gdistances_threaded.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, ctypes.c_int32,
                                                ctypes.c_int, ctypes.c_int32, _i32p]
    lib.orc_par_bfs_tree_threaded.argtypes = [ctypes.c_int32, _i32p, _i32p, ctypes.c_int32, ctypes.c_int, _i32p]
    lib.orc_dijkstra_allpaths.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, ctypes.c_int32,
                                          _i64p, _f64p, _i32p, _i32p, _i32p]
    lib.orc_betweenness.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, ctypes.c_int,
                                    _i32p, ctypes.c_int32, ctypes.c_int, ctypes.c_int, _f64p]
    lib.orc_betweenness_threaded.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, ctypes.c_int,
                                             _i32p, ctypes.c_int32, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, _f64p]
    lib.orc_rescale.argtypes = [_f64p, ctypes.c_int32, ctypes.c_int, ctypes.c_int, ctypes.c_int64]
    lib.orc_stress.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, ctypes.c_int, _i32p, ctypes.c_int32,
                               ctypes.c_int, _i64p]
    lib.orc_bfs_dist_parent.argtypes = [ctypes.c_int32, _i32p, _i32p, ctypes.c_int32, _i32p, _i32p]
    lib.orc_bfs_dist_parent.restype = ctypes.c_int64
    _lib = lib
    return lib


def _p32(a):
    return a.ctypes.data_as(_i32p)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _check(rc):
    if rc == -2:
        raise IndexError("vertex out of range")
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")


def gdistances(row_ptr, col, srcs, prefill=-1):
    """Serial gdistances (bfs.jl:141): pre-filled with `prefill`, then
    gdistances! writes only reached entries."""
    lib = _load()
    row_ptr, col, srcs = _c32(row_ptr), _c32(col), _c32(np.atleast_1d(srcs))
    nv = len(row_ptr) - 1
    out = np.full(nv, prefill, dtype=np.int32)
    _check(lib.orc_gdistances(nv, _p32(row_ptr), _p32(col), _p32(srcs), len(srcs), _p32(out)))
    return out


def gdistances_inplace(row_ptr, col, srcs, vert_level):
    """gdistances! (bfs.jl:95-124) on a caller-owned int32 buffer."""
    lib = _load()
    row_ptr, col, srcs = _c32(row_ptr), _c32(col), _c32(np.atleast_1d(srcs))
    assert vert_level.dtype == np.int32 and vert_level.flags.c_contiguous
    _check(lib.orc_gdistances(len(row_ptr) - 1, _p32(row_ptr), _p32(col), _p32(srcs), len(srcs),
                              _p32(vert_level)))
    return vert_level


def bfs_parents(row_ptr, col, src):
    """_bfs_parents over the given CSR (out- or in-neighbours)."""
    lib = _load()
    row_ptr, col = _c32(row_ptr), _c32(col)
    nv = len(row_ptr) - 1
    out = np.empty(nv, dtype=np.int32)
    _check(lib.orc_bfs_parents(nv, _p32(row_ptr), _p32(col), int(src), _p32(out)))
    return out


def par_gdistances(row_ptr, col, srcs):
    lib = _load()
    row_ptr, col, srcs = _c32(row_ptr), _c32(col), _c32(np.atleast_1d(srcs))
    nv = len(row_ptr) - 1
    out = np.empty(nv, dtype=np.int32)
    _check(lib.orc_par_gdistances(nv, _p32(row_ptr), _p32(col), _p32(srcs), len(srcs), _p32(out)))
    return out


def par_bfs_tree(row_ptr, col, src):
    lib = _load()
    row_ptr, col = _c32(row_ptr), _c32(col)
    nv = len(row_ptr) - 1
    out = np.empty(nv, dtype=np.int32)
    _check(lib.orc_par_bfs_tree(nv, _p32(row_ptr), _p32(col), int(src), _p32(out)))
    return out


def par_gdistances_threaded(row_ptr, col, srcs, nthreads, queue_segment_size=20):
    """Parallel.gdistances on `nthreads` threads (the reference's schedule;
    timed as the CPU baseline).  Same output as par_gdistances."""
    lib = _load()
    row_ptr, col, srcs = _c32(row_ptr), _c32(col), _c32(np.atleast_1d(srcs))
    nv = len(row_ptr) - 1
    out = np.empty(nv, dtype=np.int32)
    _check(lib.orc_par_gdistances_threaded(nv, _p32(row_ptr), _p32(col), _p32(srcs), len(srcs), int(nthreads),
                                           int(queue_segment_size), _p32(out)))
    return out


def par_bfs_tree_threaded(row_ptr, col, src, nthreads):
    """Parallel.bfs_tree on `nthreads` threads: one legal outcome of the
    reference's first-CAS-wins race (timed as the CPU baseline)."""
    lib = _load()
    row_ptr, col = _c32(row_ptr), _c32(col)
    nv = len(row_ptr) - 1
    out = np.empty(nv, dtype=np.int32)
    _check(lib.orc_par_bfs_tree_threaded(nv, _p32(row_ptr), _p32(col), int(src), int(nthreads), _p32(out)))
    return out


def dijkstra_allpaths(row_ptr, col, in_row_ptr, src):
    """Returns (dists, pathcounts, preds list-of-lists, closest_vertices)."""
    lib = _load()
    row_ptr, col, in_row_ptr = _c32(row_ptr), _c32(col), _c32(in_row_ptr)
    nv = len(row_ptr) - 1
    dists = np.empty(nv, dtype=np.int64)
    pc = np.empty(nv, dtype=np.float64)
    cnt = np.empty(nv, dtype=np.int32)
    preds = np.empty(max(int(in_row_ptr[-1]), 1), dtype=np.int32)
    closest = np.empty(nv, dtype=np.int32)
    _check(lib.orc_dijkstra_allpaths(nv, _p32(row_ptr), _p32(col), _p32(in_row_ptr), int(src),
                                     dists.ctypes.data_as(_i64p), pc.ctypes.data_as(_f64p),
                                     _p32(cnt), _p32(preds), _p32(closest)))
    plist = [preds[in_row_ptr[v]:in_row_ptr[v] + cnt[v]].tolist() for v in range(nv)]
    return dists, pc, plist, closest


def betweenness(row_ptr, col, in_row_ptr, directed, vs, normalize=True, endpoints=False):
    """Serial betweenness_centrality (betweenness.jl:46-75)."""
    lib = _load()
    row_ptr, col, in_row_ptr = _c32(row_ptr), _c32(col), _c32(in_row_ptr)
    vs = _c32(np.atleast_1d(vs) if len(np.atleast_1d(vs)) else np.zeros(0, np.int32))
    nv = len(row_ptr) - 1
    out = np.empty(nv, dtype=np.float64)
    _check(lib.orc_betweenness(nv, _p32(row_ptr), _p32(col), _p32(in_row_ptr), int(directed),
                               _p32(vs), len(vs), int(normalize), int(endpoints),
                               out.ctypes.data_as(_f64p)))
    return out


def betweenness_threaded(row_ptr, col, in_row_ptr, directed, vs, normalize=True,
                         endpoints=False, nthreads=None, index_quirk=True):
    """threaded_betweenness_centrality (Parallel/centrality/betweenness.jl:42-75).
    index_quirk=True restates the findall-positions-as-sources behaviour."""
    lib = _load()
    row_ptr, col, in_row_ptr = _c32(row_ptr), _c32(col), _c32(in_row_ptr)
    vs = _c32(np.atleast_1d(vs) if len(np.atleast_1d(vs)) else np.zeros(0, np.int32))
    nv = len(row_ptr) - 1
    out = np.empty(nv, dtype=np.float64)
    nt = nthreads or os.cpu_count() or 1
    _check(lib.orc_betweenness_threaded(nv, _p32(row_ptr), _p32(col), _p32(in_row_ptr),
                                        int(directed), _p32(vs), len(vs), int(normalize),
                                        int(endpoints), int(nt), int(index_quirk),
                                        out.ctypes.data_as(_f64p)))
    return out


def stress(row_ptr, col, in_row_ptr, directed, vs, nthreads=1):
    """stress_centrality(g, vs) (stress.jl:33-77); Int64, wrapping.  With
    nthreads > 1 the Parallel form (Parallel/centrality/stress.jl), which
    sums the same integers."""
    lib = _load()
    row_ptr, col, in_row_ptr = _c32(row_ptr), _c32(col), _c32(in_row_ptr)
    vs = _c32(np.atleast_1d(vs) if len(np.atleast_1d(vs)) else np.zeros(0, np.int32))
    nv = len(row_ptr) - 1
    out = np.empty(nv, dtype=np.int64)
    _check(lib.orc_stress(nv, _p32(row_ptr), _p32(col), _p32(in_row_ptr), int(directed), _p32(vs), len(vs),
                          int(nthreads), out.ctypes.data_as(_i64p)))
    return out


def _dists_int64(row_ptr, col, s):
    """dijkstra_shortest_paths(g, s).dists with unit lengths: BFS distances,
    typemax(Int) where unreached."""
    d = gdistances(row_ptr, col, [int(s)]).astype(np.int64)
    return np.where(d < 0, np.iinfo(np.int64).max, d)


def eccentricity(row_ptr, col, vs):
    """eccentricity(g, vs) (src/distance.jl:74-84): maximum(dists)."""
    return np.array([int(_dists_int64(row_ptr, col, s).max()) for s in np.atleast_1d(vs)], dtype=np.int64)


def closeness(row_ptr, col, in_row_ptr, directed, normalize=True):
    """closeness_centrality(g; normalize) (src/centrality/closeness.jl:32-57),
    one BFS per vertex."""
    nv = len(row_ptr) - 1
    deg = np.diff(np.asarray(row_ptr)) + (np.diff(np.asarray(in_row_ptr)) if directed else 0)
    out = np.zeros(nv, dtype=np.float64)
    tmax = np.iinfo(np.int64).max
    for u in range(nv):
        if deg[u] == 0:
            continue
        d = _dists_int64(row_ptr, col, u)
        delta = d[d != tmax]
        sigma = int(delta.sum())
        l = len(delta) - 1
        if sigma > 0:
            out[u] = l / sigma
            if normalize:
                out[u] *= l / (nv - 1)
    return out


def radiality(row_ptr, col):
    """radiality_centrality(g) (src/centrality/radiality.jl:34-47), one BFS
    per vertex; sum(d.dists) wraps in Int64 as Julia's does."""
    nv = len(row_ptr) - 1
    meandists = np.zeros(nv, dtype=np.float64)
    dmtr = 0.0
    for v in range(nv):
        d = _dists_int64(row_ptr, col, v)
        dmtr = max(dmtr, float(d.max()))
        total = int(d.astype(np.uint64).sum(dtype=np.uint64).view(np.int64)) if nv else 0
        meandists[v] = total / (nv - 1)
    meandists = (dmtr + 1) - meandists
    return meandists / dmtr


def connected_component_labels(row_ptr, col, in_row_ptr, in_col, directed):
    """connected_components!(label, g) (src/connectivity.jl:13-34), 0-based
    labels: the outer loop over vertices in order, a FIFO over
    all_neighbors (out- and in-neighbours for a directed graph)."""
    nv = len(row_ptr) - 1
    label = [-1] * nv
    for u in range(nv):
        if label[u] >= 0:
            continue
        label[u] = u
        q = [u]
        head = 0
        while head < len(q):
            src = q[head]
            head += 1
            nbrs = list(col[row_ptr[src]:row_ptr[src + 1]])
            if directed:
                nbrs += list(in_col[in_row_ptr[src]:in_row_ptr[src + 1]])
            for x in nbrs:
                if label[x] < 0:
                    q.append(int(x))
                    label[x] = u
    return np.array(label, dtype=np.int64)


def has_path(row_ptr, col, u, v, exclude=()):
    """has_path(g, u, v; exclude_vertices) (src/traversals/bfs.jl:150-172),
    0-based."""
    nv = len(row_ptr) - 1
    seen = [False] * nv
    for x in exclude:
        seen[x] = True
    if seen[u] or seen[v]:
        return False
    if u == v:
        return True
    q = [u]
    seen[u] = True
    head = 0
    while head < len(q):
        src = q[head]
        head += 1
        for x in col[row_ptr[src]:row_ptr[src + 1]]:
            if x == v:
                return True
            if not seen[x]:
                q.append(int(x))
                seen[x] = True
    return False


def neighborhood_dists(row_ptr, col, v, d):
    """_neighborhood (src/connectivity.jl:685-702) with unit lengths over the
    given CSR (out- or in-neighbours), 0-based: the FIFO queue of
    (vertex, distance)."""
    nv = len(row_ptr) - 1
    q = []
    if d < 0:
        return q
    q.append((v, 0))
    seen = [False] * nv
    seen[v] = True
    i = 0
    while i < len(q):
        src, cur = q[i]
        i += 1
        if cur >= d:
            continue
        for dst in col[row_ptr[src]:row_ptr[src + 1]]:
            if not seen[dst]:
                seen[dst] = True
                if cur + 1 <= d:
                    q.append((int(dst), cur + 1))
    return q


def period(row_ptr, col, in_row_ptr, in_col):
    """period(g) of a directed graph (src/connectivity.jl:497-516): raises
    ValueError unless strongly connected; 1 with a self-loop; else the gcd
    over the edges outside bfs_tree(g, 1) of levels[s] - levels[d] + 1."""
    import math
    nv = len(row_ptr) - 1
    fwd = gdistances(row_ptr, col, [0])
    bwd = gdistances(in_row_ptr, in_col, [0])
    if nv == 0 or (fwd < 0).any() or (bwd < 0).any():
        raise ValueError("Graph must be strongly connected")
    src = np.repeat(np.arange(nv), np.diff(np.asarray(row_ptr)))
    if (src == np.asarray(col)).any():
        return 1
    par = bfs_parents(row_ptr, col, 0)
    tree = {(int(par[v]), v) for v in range(nv) if v != 0}
    # levels of the tree from vertex 1 are the BFS distances
    divisor = 0
    for s, d in zip(src.tolist(), np.asarray(col).tolist()):
        if (s, d) in tree:
            continue
        divisor = math.gcd(divisor, int(fwd[s]) - int(fwd[d]) + 1)
        if divisor == 1:
            return 1
    return divisor


def rescale(bc, n, normalize, directed, k):
    lib = _load()
    bc = np.ascontiguousarray(bc, dtype=np.float64).copy()
    lib.orc_rescale(bc.ctypes.data_as(_f64p), int(n), int(normalize), int(directed), int(k))
    return bc


def bfs_dist_parent(row_ptr, col, src):
    """Serial single-source BFS with distances and parents (CPU baseline
    leg).  Returns (dist, parents, edges_scanned)."""
    lib = _load()
    row_ptr, col = _c32(row_ptr), _c32(col)
    nv = len(row_ptr) - 1
    d = np.empty(nv, dtype=np.int32)
    p = np.empty(nv, dtype=np.int32)
    n = lib.orc_bfs_dist_parent(nv, _p32(row_ptr), _p32(col), int(src), _p32(d), _p32(p))
    if n < 0:
        _check(int(n))
    return d, p, int(n)
