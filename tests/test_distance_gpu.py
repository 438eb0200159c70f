"""GPU parity of the distance measures (closeness, radiality, eccentricity,
diameter, radius, center, periphery; serial and Parallel) against the
reference fixture, the reference's own vectors and the oracle.  Bar: exact
(the three per-source integers are exact and the host formulas follow the
reference's operation order)."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_radiality_reference_fixture():
    import lgmi
    z = load_golden("ref_graph_50_500.npz")
    g = lgmi.SimpleDiGraph.from_edges(int(z["n"]), z["src"], z["dst"])
    say("radiality graph-50-500")
    assert lgmi.radiality_centrality(g).tolist() == z["rc"].tolist()
    assert lgmi.Parallel.radiality_centrality(g, parallel="distributed").tolist() == z["rc"].tolist()
    assert lgmi.Parallel.radiality_centrality(g, parallel="threads").tolist() == z["rc"].tolist()


def test_reference_vectors():
    import lgmi
    g = lgmi.SimpleGraph.from_edges(5, [1, 2, 3, 1, 4], [2, 3, 4, 4, 5], one_based=True)
    assert np.allclose(lgmi.radiality_centrality(g), [5 / 6, 3 / 4, 5 / 6, 11 / 12, 2 / 3], rtol=1e-15)
    g = lgmi.SimpleDiGraph.from_edges(4, [1, 2, 1, 3], [2, 3, 3, 4], one_based=True)
    assert lgmi.closeness_centrality(g, normalize=False).tolist() == [0.75, 0.6666666666666666, 1.0, 0.0]
    assert lgmi.closeness_centrality(g).tolist() == [0.75, 0.4444444444444444, 0.3333333333333333, 0.0]
    assert lgmi.Parallel.closeness_centrality(g, parallel="threads").tolist() == [
        0.75, 0.4444444444444444, 0.3333333333333333, 0.0]
    g = lgmi.SimpleGraph.from_edges(5, [1], [2], one_based=True)
    assert lgmi.closeness_centrality(g).tolist() == [0.25, 0.25, 0.0, 0.0, 0.0]
    star5 = lgmi.SimpleGraph.from_edges(5, [1, 1, 1, 1], [2, 3, 4, 5], one_based=True)
    assert lgmi.closeness_centrality(star5).tolist() == [1.0] + [0.5714285714285714] * 4
    path4 = lgmi.SimpleGraph.from_edges(4, [1, 2, 3], [2, 3, 4], one_based=True)
    assert lgmi.closeness_centrality(path4).tolist() == [0.5, 0.75, 0.75, 0.5]
    star4 = lgmi.SimpleGraph.from_edges(4, [1, 1, 1], [2, 3, 4], one_based=True)
    assert lgmi.radiality_centrality(star4).tolist() == [1.0] + [0.6666666666666666] * 3
    path3 = lgmi.SimpleGraph.from_edges(3, [1, 2], [2, 3], one_based=True)
    assert lgmi.radiality_centrality(path3).tolist() == [0.75, 1.0, 0.75]
    assert lgmi.eccentricity(path3, 1) == 2
    assert lgmi.eccentricity(path3, [1, 2]).tolist() == [2, 1]
    assert (lgmi.diameter(star5), lgmi.radius(star5)) == (2, 1)
    assert lgmi.periphery(star5).tolist() == [2, 3, 4, 5] and lgmi.center(star5).tolist() == [1]
    path5 = lgmi.SimpleGraph.from_edges(5, [1, 2, 3, 4], [2, 3, 4, 5], one_based=True)
    assert (lgmi.Parallel.diameter(path5), lgmi.Parallel.radius(path5)) == (4, 2)
    assert lgmi.Parallel.periphery(path5).tolist() == [1, 5] and lgmi.Parallel.center(path5).tolist() == [3]
    two = lgmi.SimpleGraph.from_edges(2, [], [])
    with pytest.warns(UserWarning):
        assert lgmi.eccentricity(two).tolist() == [np.iinfo(np.int64).max] * 2


@pytest.mark.parametrize("mode", [0, 1])
def test_distance_stats_vs_oracle(oracle, mode):
    """The three per-source integers against the oracle's BFS distances:
    undirected RMAT with heavy rows and isolated vertices, a directed graph
    with heavy rows, more than one lane group, duplicate sources; both
    launch paths."""
    import lgmi
    from lgmi import centrality as C
    from lgmi import distances as D
    from test_bc_gpu import _directed_hubs
    from test_oracle import stats_from_oracle
    try:
        C.set_launch(mode)
        g = lgmi.generators.rmat(14, 16, seed=31)
        rp, col = g.csr()
        rng = np.random.default_rng(32)
        vs0 = np.concatenate([rng.integers(0, g.nv(), 200), [3, 3]])
        say(f"distance stats rmat14 mode {mode}")
        got = D.distance_stats(g, vs0 + 1)
        assert C.last_stats(g)["launch"] in ((1, 2) if mode == 0 else (0,))
        assert got.tolist() == stats_from_oracle(oracle, rp, col, vs0).tolist()
        g = _directed_hubs()
        rp, col = g.csr()
        vs0 = np.concatenate([[0, 1, 4999], rng.integers(0, 5000, 100)])
        say(f"distance stats directed hubs mode {mode}")
        assert D.distance_stats(g, vs0 + 1).tolist() == stats_from_oracle(oracle, rp, col, vs0).tolist()
    finally:
        C.set_launch(0)


def test_measures_vs_oracle(oracle):
    """closeness, radiality, eccentricity and Parallel forms on disconnected
    random graphs (undirected and directed) against the oracle: exact."""
    import lgmi
    rng = np.random.default_rng(33)
    for directed in (False, True):
        n = 700
        src = rng.integers(0, n, 900)
        dst = rng.integers(0, n, 900)
        cls = lgmi.SimpleDiGraph if directed else lgmi.SimpleGraph
        g = cls.from_edges(n, src, dst, drop_self_loops=True)
        rp, col = g.csr()
        irp = g.in_csr()[0] if directed else rp
        say(f"measures directed={directed}")
        for normalize in (True, False):
            want = oracle.closeness(rp, col, irp, directed, normalize=normalize)
            assert lgmi.closeness_centrality(g, normalize=normalize).tolist() == want.tolist()
        want = oracle.radiality(rp, col)
        assert np.array_equal(lgmi.radiality_centrality(g), want, equal_nan=True)
        assert np.array_equal(lgmi.Parallel.radiality_centrality(g, parallel="threads"), want, equal_nan=True)
        with pytest.warns(UserWarning):
            e = lgmi.eccentricity(g)
        assert e.tolist() == oracle.eccentricity(rp, col, range(n)).tolist()


def test_long_path_many_levels(oracle):
    """A 5000-vertex path (4999 levels, far beyond any level capacity): the
    distances-only pass keeps two levels."""
    import lgmi
    n = 5000
    g = lgmi.SimpleGraph.from_edges(n, np.arange(n - 1), np.arange(1, n))
    vs = np.array([1, 2500, 5000, 17])
    say("path 5000")
    assert lgmi.eccentricity(g, vs).tolist() == [4999, 2500, 4999, 4983]
    from lgmi import distances as D
    st = D.distance_stats(g, vs)
    assert st[:, 1].tolist() == [n] * 4
    assert st[:, 0].tolist() == [n * (n - 1) // 2, 2499 * 2500 // 2 + 2500 * 2501 // 2, n * (n - 1) // 2,
                                 16 * 17 // 2 + 4983 * 4984 // 2]
