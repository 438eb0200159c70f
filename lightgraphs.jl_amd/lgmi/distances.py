"""Distance-based measures from one forward pass per 64 sources on the GPU.

The reference runs one dijkstra_shortest_paths per source for each of
closeness_centrality (src/centrality/closeness.jl:32-57),
radiality_centrality (src/centrality/radiality.jl:34-47) and eccentricity
(src/distance.jl:74-84), and derives diameter, radius, center and periphery
(src/distance.jl:86-174) from the eccentricities.  With unit edge lengths
each of these depends on a source's distance vector only through three
integers -- the sum of the finite distances, the number of vertices reached
(the source included) and the largest finite distance -- which
lgmi_distance_stats returns for every source.  The floating-point formulas
below follow the reference's operation order; a vertex a source does not
reach is typemax(Int) in the reference's distance vector, so the Int64 sums
that include it wrap exactly as Julia's do.  Weighted distance matrices are
not on the accelerated path (NotImplementedError, for the caller to fall
back to the reference method).
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _lib
from .graphs import _GraphBase

TYPEMAX_INT = np.iinfo(np.int64).max


def _check_distmx(distmx):
    if distmx is not None:
        raise NotImplementedError("weighted distmx is not on the accelerated path")


def distance_stats(g: _GraphBase, vs=None) -> np.ndarray:
    """Per source (1-based ids, all vertices by default): rows of (sum of
    the finite distances, vertices reached including the source, largest
    finite distance), Int64."""
    n = g.nv()
    srcs1 = np.arange(1, n + 1, dtype=np.int64) if vs is None else np.asarray(list(np.atleast_1d(vs)), np.int64)
    out = np.zeros((max(len(srcs1), 1), 3), dtype=np.int64)
    if len(srcs1) == 0:
        return out[:0]
    if srcs1.min() < 1 or srcs1.max() > n:
        raise IndexError(f"vertex out of range 1:{n}")
    srcs0 = np.ascontiguousarray(srcs1 - 1, dtype=np.int32)
    _lib.check(_lib.load().lgmi_distance_stats(g.handle(), _lib.i32p(srcs0), len(srcs0), _lib.i64p(out)))
    return out


def wrapped_sum(stats: np.ndarray, n: int) -> np.ndarray:
    """sum(d.dists) in Int64 with typemax(Int) for every unreached vertex
    (wrapping arithmetic, as Julia's Int)."""
    unreached = (n - stats[:, 1]).astype(np.uint64)
    total = stats[:, 0].astype(np.uint64) + unreached * np.uint64(TYPEMAX_INT)
    return total.view(np.int64)


def ecc_from_stats(stats: np.ndarray, n: int) -> np.ndarray:
    """maximum(d.dists): typemax(Int) when some vertex is unreached."""
    return np.where(stats[:, 1] == n, stats[:, 2], TYPEMAX_INT).astype(np.int64)


def eccentricity(g: _GraphBase, vs=None, distmx=None):
    """eccentricity(g[, v | vs]) (src/distance.jl:74-84): Int64 for a single
    vertex, a vector for a vertex list (all vertices by default)."""
    _check_distmx(distmx)
    single = isinstance(vs, (int, np.integer))
    ecc = ecc_from_stats(distance_stats(g, [vs] if single else vs), g.nv())
    if (ecc == TYPEMAX_INT).any():
        warnings.warn("Infinite path length detected")
    return int(ecc[0]) if single else ecc


def _eccs(g_or_ecc, distmx=None):
    if isinstance(g_or_ecc, _GraphBase):
        return eccentricity(g_or_ecc, None, distmx)
    return np.asarray(g_or_ecc)


def diameter(g_or_ecc, distmx=None):
    """diameter(eccentricities) / diameter(g) (src/distance.jl:100-103)."""
    return np.max(_eccs(g_or_ecc, distmx))


def radius(g_or_ecc, distmx=None):
    """radius(eccentricities) / radius(g) (src/distance.jl:149-152)."""
    return np.min(_eccs(g_or_ecc, distmx))


def periphery(g_or_ecc, distmx=None) -> np.ndarray:
    """periphery (src/distance.jl:124-129): 1-based vertices whose
    eccentricity equals the diameter."""
    e = _eccs(g_or_ecc, distmx)
    return np.nonzero(e == np.max(e))[0].astype(np.int64) + 1


def center(g_or_ecc, distmx=None) -> np.ndarray:
    """center (src/distance.jl:170-174): 1-based vertices whose
    eccentricity equals the radius."""
    e = _eccs(g_or_ecc, distmx)
    return np.nonzero(e == np.min(e))[0].astype(np.int64) + 1


def closeness_from_stats(stats: np.ndarray, n: int, active: np.ndarray, normalize: bool) -> np.ndarray:
    """closeness_centrality's per-vertex formula (closeness.jl:40-54) for the
    vertices `active` (0-based, degree > 0) with their distance statistics:
    with sigma the sum of the finite distances and l the number of other
    vertices reached, l / sigma when sigma > 0, times l / (nv - 1) when
    normalize; 0 otherwise."""
    out = np.zeros(n, dtype=np.float64)
    for i, u in enumerate(active):
        sigma, l = int(stats[i, 0]), int(stats[i, 1]) - 1
        if sigma > 0:
            c = float(l) / float(sigma)
            if normalize:
                c *= float(l) / float(n - 1)
            out[u] = c
    return out


def closeness_centrality(g: _GraphBase, distmx=None, normalize=True) -> np.ndarray:
    """closeness_centrality(g; normalize) (src/centrality/closeness.jl:32-57);
    a vertex of degree 0 gets 0."""
    _check_distmx(distmx)
    n = g.nv()
    if n == 0:
        return np.zeros(0, dtype=np.float64)
    active = np.nonzero(g.degree_array() > 0)[0]
    if len(active) == 0:
        return np.zeros(n, dtype=np.float64)
    return closeness_from_stats(distance_stats(g, active + 1), n, active, normalize)


def radiality_from_stats(stats: np.ndarray, n: int) -> np.ndarray:
    """radiality_centrality's formula (radiality.jl:37-46) from the
    statistics of every vertex: ((D + 1) - mean_v) / D with mean_v =
    sum(d.dists) / (nv - 1) over the whole distance vector and D the largest
    maximum(d.dists), as Float64."""
    with np.errstate(divide="ignore", invalid="ignore"):
        mean = wrapped_sum(stats, n).astype(np.float64) / np.float64(n - 1)
        dmtr = max(0.0, float(np.max(ecc_from_stats(stats, n).astype(np.float64))))
        return (np.float64(dmtr) + 1.0 - mean) / np.float64(dmtr)


def radiality_centrality(g: _GraphBase) -> np.ndarray:
    """radiality_centrality(g) (src/centrality/radiality.jl:34-47)."""
    n = g.nv()
    if n == 0:
        return np.zeros(0, dtype=np.float64)
    return radiality_from_stats(distance_stats(g), n)
