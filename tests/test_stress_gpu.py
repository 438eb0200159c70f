"""GPU parity of stress_centrality (serial and Parallel) against the
reference fixture and the oracle.  Bar: exact (Int64)."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_stress_reference_fixture():
    """test/centrality/stress.jl:2-7, test/parallel/centrality/stress.jl:2-10."""
    import lgmi
    z = load_golden("ref_graph_50_500.npz")
    g = lgmi.SimpleDiGraph.from_edges(int(z["n"]), z["src"], z["dst"])
    say("stress graph-50-500")
    assert lgmi.stress_centrality(g).tolist() == z["sc"].tolist()
    assert lgmi.Parallel.stress_centrality(g, parallel="distributed").tolist() == z["sc"].tolist()
    assert lgmi.Parallel.stress_centrality(g, parallel="threads").tolist() == z["sc"].tolist()
    assert len(lgmi.stress_centrality(g, 3)) == 50
    assert len(lgmi.stress_centrality(g, list(range(1, 21)))) == 50


def test_stress_small_reference_cases():
    import lgmi
    # stress.jl test :17-24 and doctests :18-30
    g = lgmi.SimpleGraph.from_edges(5, [1, 2, 3, 1, 4], [2, 3, 4, 4, 5], one_based=True)
    assert lgmi.stress_centrality(g).tolist() == [4, 2, 4, 10, 0]
    g = lgmi.SimpleGraph.from_edges(3, [1, 1], [2, 3], one_based=True)
    assert lgmi.stress_centrality(g).tolist() == [2, 0, 0]
    g = lgmi.SimpleGraph.from_edges(4, [1, 2, 3, 1], [2, 3, 4, 4], one_based=True)
    assert lgmi.stress_centrality(g).tolist() == [2, 2, 2, 2]
    g = lgmi.SimpleGraph.from_edges(4, [1, 2], [2, 3], one_based=True)
    assert lgmi.stress_centrality(g, [4]).tolist() == [0, 0, 0, 0]
    assert lgmi.stress_centrality(g, []).tolist() == [0, 0, 0, 0]


@pytest.mark.parametrize("mode", [0, 1])
def test_stress_vs_oracle(oracle, mode):
    """Undirected RMAT with heavy rows (more than 1024 neighbours), a directed
    graph with heavy in- and out-rows, more than one 64-lane group, duplicate
    and degree-0 sources; persistent launch (mode 0) and one launch per level
    (mode 1)."""
    import lgmi
    from lgmi import centrality as C
    from test_bc_gpu import _directed_hubs
    try:
        C.set_launch(mode)
        g = lgmi.generators.rmat(15, 16, seed=21)
        rp, col = g.csr()
        rng = np.random.default_rng(22)
        iso = np.nonzero(np.diff(rp) == 0)[0]
        vs0 = np.concatenate([rng.integers(0, g.nv(), 300), iso[:3], [5, 5]])
        say(f"stress rmat15 mode {mode}")
        got = lgmi.stress_centrality(g, vs0 + 1)
        assert C.last_stats(g)["launch"] in ((1, 2) if mode == 0 else (0,))
        want = oracle.stress(rp, col, rp, False, vs0, nthreads=8)
        assert got.tolist() == want.tolist()
        g = _directed_hubs()
        rp, col = g.csr()
        irp, _ = g.in_csr()
        vs0 = np.concatenate([[0, 1, 4999], rng.integers(0, 5000, 200)])
        say(f"stress directed hubs mode {mode}")
        got = lgmi.stress_centrality(g, vs0 + 1)
        want = oracle.stress(rp, col, irp, True, vs0, nthreads=8)
        assert got.tolist() == want.tolist()
    finally:
        C.set_launch(0)


def test_stress_then_betweenness_same_graph(oracle):
    """Stress and betweenness alternate on one graph handle (shared work
    buffers, separately captured launches)."""
    import lgmi
    g = lgmi.generators.barabasi_albert(3000, 3, seed=23)
    rp, col = g.csr()
    vs0 = np.arange(0, 3000, 7)
    for _ in range(2):
        s = lgmi.stress_centrality(g, vs0 + 1)
        assert s.tolist() == oracle.stress(rp, col, rp, False, vs0, nthreads=8).tolist()
        b = lgmi.betweenness_centrality(g, vs0 + 1)
        want = oracle.betweenness(rp, col, rp, False, vs0)
        assert np.max(np.abs(b - want) / np.maximum(np.abs(want), 1e-300)) < 1e-12
