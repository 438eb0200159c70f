"""Parity at the measured workloads' full sizes.

* BFS on symmetrised RMAT-22 (4.19M vertices, 128.3M adjacency entries), the
  throughput workload: distances and parents of four of its 64 seeded roots
  against the oracle's serial BFS, bit for bit, plus the invariants every
  BFS result has (edges join levels at most one apart; the parent is a
  neighbour one level up and the smallest such id) on a further root.  All
  64 roots also go through lgmi_bfs_run_many, five of its rows against the
  oracle bit for bit.
* Exact betweenness on RMAT-20 with the 1024 seeded sources of the
  betweenness workload, through a size-independent identity: for a source s,
  the dependencies sum to sum_t (d(s, t) - 1) over the vertices t != s that s
  reaches (each shortest s-t path has d(s, t) - 1 interior vertices), so
  with normalize=false on an undirected graph (_rescale!, betweenness.jl:
  136-160: 0.5 * n / k) the betweenness vector sums to
  0.5 * n / k * sum_s (sum of distances - reached + 1).  The distance sums
  come from the oracle's BFS on a sample of sources and from
  lgmi_distance_stats for all of them.  Bar: 1e-12 relative.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rmat22():
    import lgmi
    say("building rmat22")
    g = lgmi.generators.rmat(22, 16, seed=1)
    return g


def _bench_roots(deg, n, seed):
    rng = np.random.default_rng(seed)
    return rng.choice(np.nonzero(deg > 0)[0], size=n, replace=False).astype(np.int64)


def test_bfs_rmat22_vs_oracle(oracle, rmat22):
    from lgmi import traversals
    g = rmat22
    rp, col = g.csr()
    roots = _bench_roots(np.diff(rp), 64, 20260101)     # the throughput workload's roots
    for r in roots[:4].tolist():
        say(f"rmat22 root {r}")
        d_or, p_or, _ = oracle.bfs_dist_parent(rp, col, r)
        d_gpu = traversals._device_dist(g, np.array([r], dtype=np.int32))
        p_gpu = traversals._parents0(g, r + 1, 0)
        assert np.array_equal(d_gpu, d_or), f"dist root {r}"
        assert np.array_equal(p_gpu, p_or), f"parents root {r}"


def test_bfs_many_rmat22_vs_oracle(oracle, rmat22):
    """All 64 roots of the throughput workload in one lgmi_bfs_run_many call
    (eight searches at a time), the entry point the throughput is measured
    on: five rows bit for bit against the oracle, and every row a search of
    its own root over the same component."""
    from lgmi import traversals
    g = rmat22
    rp, col = g.csr()
    roots = _bench_roots(np.diff(rp), 64, 20260101)
    say("rmat22 run_many, 64 roots")
    dist, par = traversals._device_many(g, roots, 0, True)
    assert dist.shape == (64, g.nv()) and par.shape == (64, g.nv())
    for i in (0, 1, 2, 3, 63):
        r = int(roots[i])
        say(f"rmat22 run_many row {i}, root {r}")
        d_or, p_or, _ = oracle.bfs_dist_parent(rp, col, r)
        assert np.array_equal(dist[i], d_or), f"dist row {i} root {r}"
        assert np.array_equal(par[i], p_or), f"parents row {i} root {r}"
    idx = np.arange(64)
    assert np.all(dist[idx, roots] == 0) and np.all(par[idx, roots] == roots)
    reached = dist >= 0
    assert np.all(reached == reached[0]), "the roots share one component"
    assert np.array_equal(par >= 0, reached)


def test_bfs_rmat22_invariants(rmat22):
    from lgmi import traversals
    g = rmat22
    rp, col = g.csr()
    deg = np.diff(rp)
    r = int(_bench_roots(deg, 64, 20260101)[10])
    say(f"rmat22 invariants, root {r}")
    d = traversals._device_dist(g, np.array([r], dtype=np.int32)).astype(np.int64)
    p = traversals._parents0(g, r + 1, 0).astype(np.int64)
    row = np.repeat(np.arange(g.nv(), dtype=np.int64), deg)
    dr, dc = d[row], d[col.astype(np.int64)]
    # reached sets are closed under adjacency, and edges join levels <= 1 apart
    assert not np.any((dr >= 0) != (dc >= 0))
    both = dr >= 0
    assert np.abs(dr[both] - dc[both]).max() <= 1
    assert d[r] == 0 and p[r] == r
    # parent: the smallest neighbour one level up (the graph is undirected,
    # so a vertex's own row lists its in-neighbours)
    big = np.iinfo(np.int64).max
    cand = np.where(both & (dc == dr - 1), col.astype(np.int64), big)
    nz = deg > 0
    best = np.full(g.nv(), big, dtype=np.int64)
    best[nz] = np.minimum.reduceat(cand, rp[:-1][nz].astype(np.int64))
    reached = d > 0
    assert np.array_equal(p[reached], best[reached])
    assert np.all(p[d < 0] == -1)


def test_betweenness_rmat20_1024_distance_identity(oracle):
    import lgmi
    from lgmi.distances import distance_stats
    say("building rmat20")
    g = lgmi.generators.rmat(20, 16, seed=1)
    rp, col = g.csr()
    n = g.nv()
    srcs = _bench_roots(np.diff(rp), 1024, 4242)          # the betweenness workload's sources
    say("betweenness, 1024 sources, normalize=false")
    bc = lgmi.betweenness_centrality(g, srcs + 1, normalize=False)
    st = distance_stats(g, srcs + 1)
    # the device's distance statistics against the oracle BFS on a sample
    for i in range(0, 1024, 128):
        d_or, _, _ = oracle.bfs_dist_parent(rp, col, int(srcs[i]))
        reach = d_or >= 0
        assert int(st[i, 0]) == int(d_or[reach].sum()) and int(st[i, 1]) == int(reach.sum())
    inner = int(np.sum(st[:, 0] - st[:, 1] + 1))
    want = 0.5 * n / 1024 * inner
    got = float(np.sum(bc))
    assert abs(got - want) <= 1e-12 * abs(want), (got, want)


def test_betweenness_rmat22_1024_one_launch_identity(oracle, rmat22):
    """Exact betweenness on RMAT-22 with 1024 sources: the 16 lane groups
    fit one persistent launch (the state is about 7.6 GB per group), and the
    result meets the same dependency-sum identity as on RMAT-20."""
    import lgmi
    from lgmi.distances import distance_stats
    g = rmat22
    rp, col = g.csr()
    n = g.nv()
    srcs = _bench_roots(np.diff(rp), 1024, 4242)
    say("betweenness on rmat22, 1024 sources, normalize=false")
    bc = lgmi.betweenness_centrality(g, srcs + 1, normalize=False)
    assert lgmi.centrality.last_stats(g)["launch"] == 2, "expected one persistent, graph-captured launch"
    st = distance_stats(g, srcs + 1)
    for i in (0, 511, 1023):
        d_or, _, _ = oracle.bfs_dist_parent(rp, col, int(srcs[i]))
        reach = d_or >= 0
        assert int(st[i, 0]) == int(d_or[reach].sum()) and int(st[i, 1]) == int(reach.sum())
    inner = int(np.sum(st[:, 0] - st[:, 1] + 1))
    want = 0.5 * n / 1024 * inner
    got = float(np.sum(bc))
    assert abs(got - want) <= 1e-12 * abs(want), (got, want)
