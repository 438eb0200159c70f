"""Graph generators for the workloads (restated, seeded with this
project's SplitMix64 RNG; see include/lgmi_gen.h for the reference file:line
each one follows) plus the small named graphs the reference's tests use."""
from __future__ import annotations

import numpy as np

from . import _lib
from .graphs import SimpleDiGraph, SimpleGraph, csr_from_edges


def _edges(n_cap: int):
    return np.empty(max(1, n_cap), dtype=np.int32), np.empty(max(1, n_cap), dtype=np.int32)


def kronecker_edges(scale: int, edgefactor: int, A=0.57, B=0.19, C=0.19, seed: int = 1):
    """kronecker(SCALE, edgefactor, A, B, C; seed) sample pairs (0-based)."""
    m = edgefactor << scale
    src, dst = _edges(m)
    got = _lib.load().lgmi_gen_kronecker(scale, edgefactor, A, B, C, seed, _lib.i32p(src), _lib.i32p(dst))
    if got != m:
        raise ValueError("kronecker generation failed")
    return src, dst


def rmat(scale: int, edgefactor: int = 16, A=0.57, B=0.19, C=0.19, seed: int = 1, directed=False, T=np.int64):
    """Symmetrised RMAT (self-loops and duplicates removed) or, with
    directed=True, the kronecker digraph as the reference builds it
    (duplicates removed, self-loops kept)."""
    src, dst = kronecker_edges(scale, edgefactor, A, B, C, seed)
    n = 1 << scale
    if directed:
        rp, col = csr_from_edges(n, src, dst, symmetrize=False, drop_self_loops=False)
        return SimpleDiGraph.from_csr(rp, col, T=T)
    rp, col = csr_from_edges(n, src, dst, symmetrize=True, drop_self_loops=True)
    return SimpleGraph.from_csr(rp, col, T=T)


def erdos_renyi(n: int, ne: int, seed: int = 1, T=np.int64) -> SimpleGraph:
    """SimpleGraph(n, ne; seed) / erdos_renyi(n, ne)."""
    src, dst = _edges(ne)
    got = _lib.load().lgmi_gen_erdos_renyi(n, ne, seed, _lib.i32p(src), _lib.i32p(dst))
    if got < 0:
        raise ValueError("erdos_renyi: bad arguments")
    return SimpleGraph.from_edges(n, src[:got], dst[:got], T=T)


def barabasi_albert(n: int, k: int, seed: int = 1, T=np.int64) -> SimpleGraph:
    src, dst = _edges(k + (n - k - 1) * k)
    got = _lib.load().lgmi_gen_barabasi_albert(n, k, seed, _lib.i32p(src), _lib.i32p(dst))
    if got < 0:
        raise ValueError("barabasi_albert: bad arguments")
    return SimpleGraph.from_edges(n, src[:got], dst[:got], T=T)


def watts_strogatz(n: int, k: int, beta: float, seed: int = 1, T=np.int64) -> SimpleGraph:
    cap = n * (n - 1) // 2 if k == n - 1 else n * k // 2
    src, dst = _edges(cap)
    got = _lib.load().lgmi_gen_watts_strogatz(n, k, beta, seed, _lib.i32p(src), _lib.i32p(dst))
    if got < 0:
        raise ValueError("watts_strogatz: bad arguments")
    return SimpleGraph.from_edges(n, src[:got], dst[:got], T=T)


def grid(dims, T=np.int64) -> SimpleGraph:
    """grid([r, c]) (staticgraphs.jl:452-474); 1-D grids are path graphs."""
    dims = list(dims)
    if len(dims) == 1:
        return path_graph(dims[0], T=T)
    if len(dims) != 2:
        raise NotImplementedError("only 1-D and 2-D grids")
    r, c = int(dims[0]), int(dims[1])
    src, dst = _edges(2 * r * c)
    got = _lib.load().lgmi_gen_grid2d(r, c, _lib.i32p(src), _lib.i32p(dst))
    return SimpleGraph.from_edges(r * c, src[:got], dst[:got], T=T)


def uneven_batch(count: int = 4096, nmin: int = 200, nmax: int = 20_000, seed: int = 1):
    """The uneven small-graph batch: sizes log-uniform in [nmin, nmax],
    even positions barabasi_albert(n, 4), odd ones watts_strogatz(n, 8, 0.1),
    each seeded per graph from `seed`."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(np.int64)
    out = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            out.append(barabasi_albert(int(n), 4, seed=seed * 1_000_003 + i))
        else:
            out.append(watts_strogatz(int(n), 8, 0.1, seed=seed * 1_000_003 + i))
    return out


def path_graph(n: int, T=np.int64) -> SimpleGraph:
    a = np.arange(max(n - 1, 0))
    return SimpleGraph.from_edges(n, a, a + 1, T=T)


def cycle_graph(n: int, T=np.int64) -> SimpleGraph:
    a = np.arange(n)
    return SimpleGraph.from_edges(n, a, (a + 1) % n, T=T)


def star_graph(n: int, T=np.int64) -> SimpleGraph:
    a = np.arange(1, n)
    return SimpleGraph.from_edges(n, np.zeros_like(a), a, T=T)


def complete_graph(n: int, T=np.int64) -> SimpleGraph:
    iu = np.triu_indices(n, 1)
    return SimpleGraph.from_edges(n, iu[0], iu[1], T=T)


def house_graph(T=np.int64) -> SimpleGraph:
    """smallgraph(:house) (smallgraphs.jl:202-203)."""
    e = np.array([(1, 2), (1, 3), (2, 4), (3, 4), (3, 5), (4, 5)])
    return SimpleGraph.from_edges(5, e[:, 0], e[:, 1], T=T, one_based=True)
