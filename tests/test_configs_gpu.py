"""Parity on the measured workloads at their full sizes, beside the
full-size BFS and betweenness checks of test_fullsize_gpu.py.

* Exact betweenness on symmetrised RMAT-20 over 128 of the 1024 seeded
  sources of the betweenness workload (two full 64-source lane groups),
  vertex by vertex against the oracle's threaded Brandes
  (src/Parallel/centrality/betweenness.jl:42-75 with the explicit source
  list of test/parallel/centrality/betweenness.jl:10-30), 1e-12 relative.
* The uneven 4096-graph batch (n log-uniform in [200, 20000], half
  barabasi_albert(n, 4), half watts_strogatz(n, 8, 0.1)) in one launch:
  every graph's vector passes the dependency-sum identity (for each source
  s the dependencies sum to sum_t (d(s, t) - 1) over the vertices s
  reaches, so with normalize = false on an undirected graph, where
  _rescale! scales by 0.5 * n / k with k = n (betweenness.jl:136-160), the
  vector sums to 0.5 * sum_s (sum of distances - reached + 1)), with the
  distance sums from lgmi_distance_stats (a separate kernel, itself checked
  against the oracle's BFS); 8 seeded graphs of the batch are compared
  vertex by vertex with the oracle (test/centrality/betweenness.jl:8-13
  compares the fixture vector element-wise).
* The learned direction switch on the 4096 x 4096 grid (16.8M vertices,
  8190 levels from a corner): distances and parents of two roots under the
  learned network, the alpha/beta rule and the oracle's serial BFS, bit
  for bit.
"""
import math
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel_err(got, want):
    floor = 1e-15 * max(1.0, float(np.abs(want).max()))
    return float(np.max(np.abs(got - want) / np.maximum(np.abs(want), floor)))


def _threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(env))) if env and env.isdigit() else n


def test_config3_rmat20_two_lane_groups_per_vertex(oracle):
    import lgmi
    say("config 3: building rmat20")
    g = lgmi.generators.rmat(20, 16, seed=1)
    rp, col = g.csr()
    rng = np.random.default_rng(4242)        # the betweenness workload's sources
    srcs = rng.choice(np.nonzero(np.diff(rp) > 0)[0], size=1024, replace=False).astype(np.int64)[:128]
    say("config 3: betweenness, 128 sources")
    got = lgmi.betweenness_centrality(g, srcs + 1)
    want = oracle.betweenness_threaded(rp, col, rp, False, srcs, nthreads=_threads(), index_quirk=False)
    err = _rel_err(got, want)
    say(f"config 3: max rel err {err:.2e}")
    assert err <= 1e-12


def test_config4_full_batch_identity_and_oracle(oracle):
    import lgmi
    from lgmi.distances import distance_stats
    say("config 4: building 4096 graphs")
    graphs = lgmi.generators.uneven_batch(4096, 200, 20_000, seed=1)
    say("config 4: one batch launch, normalize=false")
    res = lgmi.betweenness_centrality_batch(graphs, normalize=False)
    assert len(res) == len(graphs)
    worst = 0.0
    for i, g in enumerate(graphs):
        n = g.nv()
        st = distance_stats(g)
        inner = int(np.sum(st[:, 0] - st[:, 1] + 1))
        want = 0.5 * inner
        got = math.fsum(res[i].tolist())
        assert np.all(res[i] >= 0.0), i
        err = abs(got - want) / max(abs(want), 1.0)
        worst = max(worst, err)
        assert err <= 1e-12, (i, n, got, want)
        g._release()
        if i % 512 == 511:
            say(f"config 4: identity checked on {i + 1} graphs, worst {worst:.2e}")
    rng = np.random.default_rng(99)
    for i in rng.choice(len(graphs), size=8, replace=False).tolist():
        rp, col = graphs[i].csr()
        want = oracle.betweenness_threaded(rp, col, rp, False, np.arange(graphs[i].nv()), normalize=False,
                                           nthreads=_threads(), index_quirk=False)
        err = _rel_err(res[i], want)
        say(f"config 4: graph {i} (n={graphs[i].nv()}) vs oracle {err:.2e}")
        assert err <= 1e-12


def test_config5_grid4096_learned_switch_bit_exact(oracle):
    import lgmi
    from lgmi import _lib, policy
    params, _ = policy.load_weights()
    if params is None:
        pytest.skip("no committed policy weights")
    say("config 5: building the 4096 x 4096 grid")
    g = lgmi.generators.grid((4096, 4096))
    rp, col = g.csr()
    lib = _lib.load()
    h = g.handle()
    n = g.nv()
    for root in (0, 4096 * 1500 + 2900):
        outs = []
        for mode in ("heuristic", "learned"):
            if mode == "learned":
                policy.set_learned(params)
            else:
                policy.set_heuristic()
            say(f"config 5: root {root}, {mode}")
            try:
                _lib.check(lib.lgmi_bfs_run(h, root, 0, 1, None))
                d = np.empty(n, np.int32)
                p = np.empty(n, np.int32)
                _lib.check(lib.lgmi_bfs_fetch(h, _lib.i32p(d), _lib.i32p(p)))
            finally:
                policy.set_heuristic()
            outs.append((d, p))
        d_or, p_or, _ = oracle.bfs_dist_parent(rp, col, root)
        for d, p in outs:
            assert np.array_equal(d, d_or), root
            assert np.array_equal(p, p_or), root
        if root == 0:
            assert int(d_or.max()) == 8190
