"""betweenness_centrality (exact Brandes, FP64) and stress_centrality
(Int64) on the GPU.

Mirrors src/centrality/betweenness.jl:46-78: ``vs`` defaults to all
vertices, an integer ``k`` samples k distinct vertices, sources with
degree 0 are skipped while ``k = length(vs)`` still counts them, and
``normalize`` / ``endpoints`` follow _rescale! / _accumulate_endpoints!.
Only the unit distance matrix (``distmx=None``, the reference's
DefaultDistance) is on the accelerated path; a weighted ``distmx`` raises
NotImplementedError so that a caller can fall back to the reference method.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .graphs import _GraphBase


def sample(n: int, k: int, rng=None) -> np.ndarray:
    """sample(vertices(g), k): k distinct vertices (src/utils.jl:43-45),
    1-based.  The reference draws from Julia's GLOBAL_RNG; this draws from
    ``rng`` (a numpy Generator or seed)."""
    if k > n:
        raise ValueError("cannot sample more vertices than the graph has")
    gen = rng if isinstance(rng, np.random.Generator) else np.random.default_rng(rng)
    return gen.choice(n, size=k, replace=False).astype(np.int64) + 1


def _vs(g: _GraphBase, vs, rng=None) -> np.ndarray:
    n = g.nv()
    if vs is None:
        return np.arange(1, n + 1, dtype=np.int64)
    if isinstance(vs, (int, np.integer)):
        return sample(n, int(vs), rng=rng)
    arr = np.asarray(list(vs), dtype=np.int64)
    return arr


def _check_distmx(distmx):
    if distmx is not None:
        raise NotImplementedError("weighted distmx is not on the accelerated path")


def _run(g: _GraphBase, srcs1: np.ndarray, k: int, normalize: bool, endpoints: bool,
         skip_degree0: bool = True) -> np.ndarray:
    n = g.nv()
    if srcs1.size and (srcs1.min() < 1 or srcs1.max() > n):
        raise IndexError(f"source vertex out of range 1:{n}")
    bc = np.zeros(max(n, 1), dtype=np.float64)
    if n == 0:
        return bc[:0]
    srcs0 = np.ascontiguousarray(srcs1 - 1, dtype=np.int32)
    lib = _lib.load()
    _lib.check(lib.lgmi_betweenness(g.handle(), _lib.i32p(srcs0 if len(srcs0) else np.zeros(1, np.int32)),
                                    len(srcs0), int(k), int(bool(skip_degree0)), int(bool(normalize)),
                                    int(bool(endpoints)),
                                    _lib.f64p(bc)))
    return bc


def betweenness_centrality(g: _GraphBase, vs=None, distmx=None, normalize=True, endpoints=False,
                           rng=None) -> np.ndarray:
    """betweenness_centrality(g[, vs | k][, distmx]; normalize, endpoints)."""
    _check_distmx(distmx)
    srcs = _vs(g, vs, rng=rng)
    return _run(g, srcs, len(srcs), normalize, endpoints)


def _run_stress(g: _GraphBase, srcs1: np.ndarray) -> np.ndarray:
    n = g.nv()
    if srcs1.size and (srcs1.min() < 1 or srcs1.max() > n):
        raise IndexError(f"source vertex out of range 1:{n}")
    out = np.zeros(max(n, 1), dtype=np.int64)
    if n == 0:
        return out[:0]
    srcs0 = np.ascontiguousarray(srcs1 - 1, dtype=np.int32)
    lib = _lib.load()
    _lib.check(lib.lgmi_stress(g.handle(), _lib.i32p(srcs0 if len(srcs0) else np.zeros(1, np.int32)), len(srcs0),
                               _lib.i64p(out)))
    return out


def stress_centrality(g: _GraphBase, vs=None, rng=None) -> np.ndarray:
    """stress_centrality(g[, vs | k]) (src/centrality/stress.jl:33-51): the
    Int64 stress of every vertex over the sources vs (all vertices by
    default; an integer k samples k distinct vertices); sources of degree 0
    are skipped.  Exact (wrapping on Int64 overflow like the reference)."""
    return _run_stress(g, _vs(g, vs, rng=rng))


def betweenness_centrality_batch(graphs, normalize=True, endpoints=False, return_ms=False):
    """betweenness_centrality(g) for each undirected graph of `graphs` (all
    sources, DefaultDistance), computed in one persistent launch.  Returns a
    list of FP64 vectors (and the device milliseconds with return_ms)."""
    lib = _lib.load()
    ng = len(graphs)
    nv = np.array([g.nv() for g in graphs], dtype=np.int32)
    for g in graphs:
        if g.is_directed():
            raise ValueError("betweenness_centrality_batch takes undirected graphs")
    rps, cols = zip(*[g.csr() for g in graphs]) if ng else ((), ())
    row_off = np.zeros(ng, dtype=np.int64)
    col_off = np.zeros(ng, dtype=np.int64)
    bc_off = np.zeros(ng, dtype=np.int64)
    if ng:
        row_off[1:] = np.cumsum([len(r) for r in rps])[:-1]
        col_off[1:] = np.cumsum([len(c) for c in cols])[:-1]
        bc_off[1:] = np.cumsum(nv.astype(np.int64))[:-1]
    rp_cat = np.ascontiguousarray(np.concatenate(rps) if ng else np.zeros(1), dtype=np.int32)
    col_cat = np.ascontiguousarray(np.concatenate(cols) if ng and sum(len(c) for c in cols) else np.zeros(1),
                                   dtype=np.int32)
    out = np.zeros(max(1, int(nv.astype(np.int64).sum())), dtype=np.float64)
    ms = np.zeros(1, dtype=np.float64)
    _lib.check(lib.lgmi_betweenness_batch(ng, _lib.i32p(nv if ng else np.zeros(1, np.int32)), _lib.i64p(row_off),
                                          _lib.i32p(rp_cat), _lib.i64p(col_off), _lib.i32p(col_cat),
                                          _lib.i64p(bc_off), int(bool(normalize)), int(bool(endpoints)),
                                          _lib.f64p(out), _lib.f64p(ms)))
    res = [out[bc_off[i]:bc_off[i] + nv[i]].copy() for i in range(ng)]
    return (res, float(ms[0])) if return_ms else res


def batch_tuning(wave_max_n=-1, task_work_target=0.0, lds_max_n=-1, lane_group=0, slots=0):
    """Launch thresholds of betweenness_centrality_batch (process-wide):
    wave_max_n -- graphs up to this size run one source per wavefront;
    task_work_target -- n * (nnz + n) above which a graph's sources are split;
    lds_max_n -- graphs above this size keep their search state in global
    memory; lane_group -- lanes per adjacency row (1, 2, 4 or 8); slots --
    entries per lane per round (2, 4, 8 or 16).  wave_max_n = -1,
    task_work_target = 0, lds_max_n = -1, lane_group = 0 and slots = 0 select
    the defaults."""
    _lib.check(_lib.load().lgmi_betweenness_batch_tuning(int(wave_max_n), float(task_work_target), int(lds_max_n),
                                                         int(lane_group), int(slots)))


def set_launch(mode: int) -> None:
    """0: all lane groups in one persistent (graph-captured) launch when they
    fit, level-per-launch otherwise (default); 1: level-per-launch only.
    Both give bit-identical results."""
    _lib.check(_lib.load().lgmi_betweenness_set_launch(int(mode)))


def last_stats(g: _GraphBase) -> dict:
    st = _lib.BcStats()
    _lib.check(_lib.load().lgmi_betweenness_stats(g.handle(), _lib.ctypes.byref(st)))
    return {f: getattr(st, f) for f, _ in st._fields_}
