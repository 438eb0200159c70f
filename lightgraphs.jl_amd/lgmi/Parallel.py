"""LightGraphs.Parallel mirrors: gdistances(!), bfs_tree(!), betweenness,
stress.

src/Parallel/traversals/gdistances.jl:36-139, src/Parallel/traversals/
bfs.jl:16-104, src/Parallel/centrality/betweenness.jl:1-75 and
src/Parallel/centrality/stress.jl:1-51.  The GPU runs
every variant; the thread/process knobs (``queue_segment_size``,
``parallel``) are accepted with the reference's meaning for results:

* Parallel.gdistances! pre-fills ``typemax(T)`` (gdistances.jl:46).
* Parallel.bfs_tree! sets ``parents[v]`` to a first-discoverer; the GPU
  gives the smallest-id one, which is one of the outcomes the threaded CAS
  (bfs.jl:59) can produce.
* parallel=:threads takes as sources the *positions* in ``vs`` of the
  vertices with degree > 0 (``findall``, betweenness.jl:53-56), exactly as
  the reference does; parallel=:distributed uses the vertices themselves.
"""
from __future__ import annotations

import numpy as np

from . import traversals as _tr
from . import centrality as _bc
from . import distances as _dist
from .graphs import _GraphBase


class ThreadQueue:
    """ThreadQueue(T, maxlength) -- Parallel/traversals/bfs.jl:16-46."""

    def __init__(self, T, maxlength: int):
        self.T = np.dtype(T)
        self.data = np.zeros(int(maxlength), dtype=self.T)
        self.head = 1
        self.tail = 1

    def push(self, val):
        off = self.tail
        self.tail += 1
        self.data[off - 1] = val
        return off

    def popfirst(self):
        off = self.head
        self.head += 1
        return self.data[off - 1]

    def isempty(self) -> bool:
        # The reference compares the Atomic object (not its value) with 1
        # (Parallel/traversals/bfs.jl:40-42), which is never equal, so only
        # head == tail counts.
        return self.head == self.tail

    def __getitem__(self, i):
        if isinstance(i, slice):
            return self.data[i.start - 1:i.stop]
        return self.data[i - 1]

    def __len__(self):
        return len(self.data)


def gdistances_(g: _GraphBase, sources, vert_level: np.ndarray, queue_segment_size: int = 20) -> np.ndarray:
    """Parallel.gdistances!(g, sources, vert_level; queue_segment_size)."""
    vert_level[:] = _tr.typemax(vert_level.dtype)
    return _tr.gdistances_(g, sources, vert_level)


def gdistances(g: _GraphBase, sources, queue_segment_size: int = 20) -> np.ndarray:
    """Parallel.gdistances(g, sources | source; queue_segment_size)."""
    out = np.empty(g.nv(), dtype=g.T)
    return gdistances_(g, sources, out, queue_segment_size=queue_segment_size)


def bfs_tree_(next_q: ThreadQueue, g: _GraphBase, source: int, parents: np.ndarray) -> np.ndarray:
    """Parallel.bfs_tree!(next, g, source, parents): fills ``parents`` (the
    reference's Array{Atomic{T}}) in place and leaves every reached vertex
    in ``next`` in breadth-first order."""
    par0 = _tr._parents0(g, source, 0)
    reached = par0 >= 0
    parents[:] = np.where(reached, par0 + 1, 0).astype(parents.dtype)
    dist = _tr._device_dist(g, np.array([int(source) - 1], dtype=np.int32))
    order = np.lexsort((np.arange(g.nv()), dist))
    order = order[dist[order] >= 0]
    for v in order:
        next_q.push(v + 1)
    next_q.head = next_q.tail
    return parents


def bfs_tree(g: _GraphBase, source: int, nv: int | None = None):
    """Parallel.bfs_tree(g, source[, nv])."""
    par = _tr.bfs_parents(g, source, dir="out")
    return _tr.tree(par)


def betweenness_centrality(g: _GraphBase, vs=None, distmx=None, normalize=True, endpoints=False,
                           parallel="distributed", rng=None) -> np.ndarray:
    """Parallel.betweenness_centrality(g[, vs | k][, distmx]; normalize,
    endpoints, parallel=:distributed | :threads)."""
    _bc._check_distmx(distmx)
    srcs = _bc._vs(g, vs, rng=rng)
    k = len(srcs)
    if parallel in ("distributed", ":distributed"):
        return _bc._run(g, srcs, k, normalize, endpoints)
    # threaded_betweenness_centrality: findall returns positions in vs
    deg = g.degree_array()
    if srcs.size and (srcs.min() < 1 or srcs.max() > g.nv()):
        raise IndexError("vertex out of range")
    pos = np.nonzero(deg[srcs - 1] > 0)[0].astype(np.int64) + 1 if srcs.size else np.zeros(0, np.int64)
    return _bc._run(g, pos, k, normalize, endpoints, skip_degree0=False)


def stress_centrality(g: _GraphBase, vs=None, parallel="distributed", rng=None) -> np.ndarray:
    """Parallel.stress_centrality(g[, vs | k]; parallel=:distributed |
    :threads) (src/Parallel/centrality/stress.jl:1-51).  Both modes sum the
    per-source Int64 vectors of the vertices in vs with degree > 0, which is
    the serial result."""
    return _bc._run_stress(g, _bc._vs(g, vs, rng=rng))


# Parallel/centrality/{closeness,radiality}.jl and Parallel/distance.jl: the
# same results as the serial functions for unit distances.
def closeness_centrality(g: _GraphBase, distmx=None, normalize=True, parallel="distributed") -> np.ndarray:
    """Parallel.closeness_centrality(g; normalize, parallel)
    (src/Parallel/centrality/closeness.jl:1-56)."""
    return _dist.closeness_centrality(g, distmx, normalize=normalize)


def radiality_centrality(g: _GraphBase, parallel="distributed") -> np.ndarray:
    """Parallel.radiality_centrality(g; parallel)
    (src/Parallel/centrality/radiality.jl:1-38)."""
    return _dist.radiality_centrality(g)


def eccentricity(g: _GraphBase, vs=None, distmx=None) -> np.ndarray:
    """Parallel.eccentricity(g[, vs]) (src/Parallel/distance.jl:3-18)."""
    _dist._check_distmx(distmx)
    return _dist.ecc_from_stats(_dist.distance_stats(g, vs), g.nv())


def diameter(g: _GraphBase, distmx=None):
    return _dist.diameter(eccentricity(g, None, distmx))


def radius(g: _GraphBase, distmx=None):
    return _dist.radius(eccentricity(g, None, distmx))


def periphery(g: _GraphBase, distmx=None):
    return _dist.periphery(eccentricity(g, None, distmx))


def center(g: _GraphBase, distmx=None):
    return _dist.center(eccentricity(g, None, distmx))
