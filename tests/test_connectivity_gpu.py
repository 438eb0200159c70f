"""GPU parity of connected_components, is_connected, has_path and period
against the reference's own cases and the oracle.  Bar: exact."""
import sys

import numpy as np
import pytest

from conftest import csr_of

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_reference_cases():
    import lgmi
    edges = [(1, 2), (2, 3), (3, 4), (5, 6), (6, 7), (8, 9), (10, 9)]
    e = np.array(edges)
    g = lgmi.SimpleGraph.from_edges(14, e[:, 0], e[:, 1], one_based=True)
    assert not lgmi.is_connected(g)
    label = lgmi.connected_components_(np.zeros(14, dtype=np.int64), g)
    assert label[:10].tolist() == [1, 1, 1, 1, 5, 5, 5, 8, 8, 8]
    cc = lgmi.connected_components(g)
    assert len(cc) >= 3 and sorted(cc[2]) == [8, 9, 10]
    house = lgmi.SimpleGraph.from_edges(5, [1, 1, 2, 3, 3, 4], [2, 3, 4, 4, 5, 5], one_based=True)
    assert lgmi.is_connected(house)
    # doctests connectivity.jl:69-86, 98-111
    g = lgmi.SimpleGraph.from_edges(5, [1, 2, 4], [2, 3, 5], one_based=True)
    assert lgmi.connected_components(g) == [[1, 2, 3], [4, 5]]
    assert not lgmi.is_connected(g)
    dg = lgmi.SimpleDiGraph.from_edges(3, [1, 2, 3], [2, 3, 1], one_based=True)
    assert lgmi.is_connected(dg)
    dg = lgmi.SimpleDiGraph.from_edges(3, [1, 2], [2, 1], one_based=True)
    assert not lgmi.is_weakly_connected(dg)
    # has_path: test/traversals/bfs.jl:48-72
    g = lgmi.SimpleGraph.from_edges(6, [1, 2, 2, 4, 3], [2, 3, 4, 5, 5], one_based=True)
    hp = lgmi.has_path
    assert hp(g, 1, 5) and hp(g, 1, 2) and hp(g, 1, 5, [3]) and hp(g, 1, 5, [4]) and not hp(g, 1, 5, [3, 4])
    assert hp(g, 5, 1) and hp(g, 5, 1, [3]) and hp(g, 5, 1, [4]) and not hp(g, 5, 1, [3, 4])
    assert not hp(g, 1, 6) and not hp(g, 6, 1) and hp(g, 1, 1)
    assert not hp(g, 1, 2, [2]) and not hp(g, 1, 2, [1])
    # period: test/connectivity.jl:206-222
    ring = np.array([(i, i % 10 + 1) for i in range(1, 11)])
    assert lgmi.period(lgmi.SimpleDiGraph.from_edges(10, ring[:, 0], ring[:, 1], one_based=True)) == 10
    ring2 = np.vstack([ring, [[1, 4]]])
    assert lgmi.period(lgmi.SimpleDiGraph.from_edges(10, ring2[:, 0], ring2[:, 1], one_based=True)) == 2
    with pytest.raises(ValueError):
        lgmi.period(lgmi.SimpleDiGraph.from_edges(3, [1, 2], [2, 3], one_based=True))
    loop = np.vstack([ring, [[3, 3]]])
    assert lgmi.period(lgmi.SimpleDiGraph.from_edges(10, loop[:, 0], loop[:, 1], one_based=True)) == 1


@pytest.mark.parametrize("directed", [False, True])
def test_components_vs_oracle(oracle, directed):
    import lgmi
    rng = np.random.default_rng(51)
    n = 3000
    src = rng.integers(0, n, 2600)
    dst = rng.integers(0, n, 2600)
    cls = lgmi.SimpleDiGraph if directed else lgmi.SimpleGraph
    g = cls.from_edges(n, src, dst)
    rp, col = g.csr()
    irp, icol = g.in_csr() if directed else (rp, col)
    say(f"components directed={directed}")
    want = oracle.connected_component_labels(rp, col, irp, icol, directed) + 1
    got = lgmi.connected_components_(np.zeros(n, dtype=np.int64), g)
    assert got.tolist() == want.tolist()
    # a hub with more than 64 neighbours goes through the wavefront path
    g = lgmi.generators.rmat(14, 8, seed=52, directed=directed)
    rp, col = g.csr()
    irp, icol = g.in_csr() if directed else (rp, col)
    want = oracle.connected_component_labels(rp, col, irp, icol, directed) + 1
    got = lgmi.connected_components_(np.zeros(g.nv(), dtype=np.int64), g)
    assert got.tolist() == want.tolist()
    assert lgmi.is_connected(g) == bool((want == 1).all())


def test_components_rmat20_label_invariants():
    """RMAT-20 (31M adjacency entries): every edge joins equal labels, every
    label is the smallest id carrying it, and a label's vertex labels
    itself."""
    import lgmi
    g = lgmi.generators.rmat(20, 16, seed=1)
    rp, col = g.csr()
    say("components rmat20")
    lab = lgmi.connected_components_(np.zeros(g.nv(), dtype=np.int64), g) - 1
    src = np.repeat(np.arange(g.nv()), np.diff(rp))
    assert (lab[src] == lab[col]).all()
    assert (lab <= np.arange(g.nv())).all()
    assert (lab[lab] == lab).all()
    first = np.full(g.nv(), g.nv(), dtype=np.int64)
    np.minimum.at(first, lab, np.arange(g.nv()))
    assert (first[lab] == lab).all()


def test_has_path_vs_oracle(oracle):
    import lgmi
    rng = np.random.default_rng(53)
    for directed in (False, True):
        n = 1500
        src = rng.integers(0, n, 1700)
        dst = rng.integers(0, n, 1700)
        cls = lgmi.SimpleDiGraph if directed else lgmi.SimpleGraph
        g = cls.from_edges(n, src, dst)
        rp, col = g.csr()
        say(f"has_path directed={directed}")
        for _ in range(25):
            u, v = (int(x) for x in rng.integers(0, n, 2))
            ex = rng.integers(0, n, int(rng.integers(0, 40))).tolist()
            assert lgmi.has_path(g, u + 1, v + 1, [x + 1 for x in ex]) == oracle.has_path(rp, col, u, v, ex)


def test_period_vs_oracle(oracle):
    import lgmi
    rng = np.random.default_rng(54)
    for k, n in [(3, 300), (5, 1000), (1, 200), (7, 7 * 150)]:
        ring = [(i, (i + 1) % n) for i in range(n)]
        chords = [(int(a), int((a + k * rng.integers(1, n // k) + 1) % n)) for a in rng.integers(0, n, 20)]
        e = np.array(ring + chords)
        g = lgmi.SimpleDiGraph.from_edges(n, e[:, 0], e[:, 1], drop_self_loops=True)
        rp, col = g.csr()
        irp, icol = g.in_csr()
        say(f"period k={k} n={n}")
        assert lgmi.period(g) == oracle.period(rp, col, irp, icol)


def test_neighborhood_reference_cases():
    import lgmi
    e = np.array([(1, 2), (2, 3), (3, 1), (3, 4), (4, 5), (5, 4)])
    g = lgmi.SimpleDiGraph.from_edges(5, e[:, 0], e[:, 1], one_based=True)
    assert lgmi.neighborhood_dists(g, 1, 3) == [(1, 0), (2, 1), (3, 2), (4, 3)]
    assert lgmi.neighborhood(g, 1, 2) == [1, 2, 3]
    assert lgmi.neighborhood_dists(g, 4, 3) == [(4, 0), (5, 1)]
    assert lgmi.neighborhood_dists(g, 4, 3, dir="in") == [(4, 0), (3, 1), (5, 1), (2, 2), (1, 3)]
    star = lgmi.SimpleGraph.from_edges(10, [1] * 9, list(range(2, 11)), one_based=True)
    assert lgmi.neighborhood_dists(star, 1, 0) == [(1, 0)]
    assert [len(lgmi.neighborhood(star, v, dd)) for v, dd in [(1, 1), (2, 1), (1, 2), (2, 2), (2, -1)]] == [
        10, 2, 10, 10, 0]
    dstar = lgmi.SimpleDiGraph.from_edges(10, [1] * 9, list(range(2, 11)), one_based=True)
    assert [len(lgmi.neighborhood(dstar, v, dd)) for v, dd in [(1, 1), (2, 1), (1, 2), (2, 2)]] == [10, 1, 10, 1]
    assert [len(lgmi.neighborhood(dstar, v, dd, dir="in")) for v, dd in [(1, 1), (2, 1), (1, 2), (2, 2)]] == [
        1, 2, 1, 2]
    assert lgmi.neighborhood_dists(dstar, 1, 2.5) == lgmi.neighborhood_dists(dstar, 1, 2)


def test_neighborhood_vs_oracle(oracle):
    """Queue order against the oracle: random directed graph (both
    directions) and RMAT-16 with hubs, several depths."""
    import lgmi
    rng = np.random.default_rng(62)
    n = 2000
    src = rng.integers(0, n, 6000)
    dst = rng.integers(0, n, 6000)
    g = lgmi.SimpleDiGraph.from_edges(n, src, dst)
    rp, col = g.csr()
    irp, icol = g.in_csr()
    say("neighborhood random digraph")
    for v in (0, 7, 1999):
        for dd in (0, 1, 3, 50):
            want = [(x + 1, y) for x, y in oracle.neighborhood_dists(rp, col, v, dd)]
            assert lgmi.neighborhood_dists(g, v + 1, dd) == want
            want = [(x + 1, y) for x, y in oracle.neighborhood_dists(irp, icol, v, dd)]
            assert lgmi.neighborhood_dists(g, v + 1, dd, dir="in") == want
    g = lgmi.generators.rmat(16, 16, seed=63)
    rp, col = g.csr()
    hub = int(np.argmax(np.diff(rp)))
    say("neighborhood rmat16")
    for v in (hub, 3):
        for dd in (1, 2, 6):
            want = [(x + 1, y) for x, y in oracle.neighborhood_dists(rp, col, v, dd)]
            assert lgmi.neighborhood_dists(g, v + 1, dd) == want
