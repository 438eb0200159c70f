"""CPU tests of the host-side mirror (argument handling that needs no GPU)."""
import numpy as np
import pytest


def test_typemax_and_tree():
    import lgmi
    assert lgmi.typemax(np.uint8) == 255 and lgmi.typemax(np.int16) == 32767
    # bfs.jl:10-19: parents [1,1,1,3] -> arcs 1->2, 1->3, 3->4
    t = lgmi.tree(np.array([1, 1, 1, 3], dtype=np.int64))
    assert t.is_directed() and t.nv() == 4 and t.ne() == 3
    assert not t.has_edge(2, 3) and t.has_edge(3, 4)
    t = lgmi.tree(np.array([1, 0, 1], dtype=np.uint8))
    assert t.ne() == 1 and t.T == np.uint8


def test_threadqueue_semantics():
    # test/parallel/traversals/bfs.jl:2-13
    from lgmi.Parallel import ThreadQueue
    q = ThreadQueue(np.int64, 5)
    assert q.isempty() is True
    q.push(1)
    assert q[1] == 1
    for i in range(2, 6):
        q.push(i)
    assert set(q[1:5].tolist()) == {1, 2, 3, 4, 5}
    assert q.popfirst() == 1


def test_sample_distinct():
    import lgmi
    s = lgmi.sample(50, 20, rng=3)
    assert len(set(s.tolist())) == 20 and s.min() >= 1 and s.max() <= 50
    with pytest.raises(ValueError):
        lgmi.sample(5, 6)


def test_graph_mutation_and_eltypes():
    import lgmi
    g = lgmi.SimpleDiGraph(4)
    for u, v in [(1, 2), (2, 3), (1, 3), (3, 4)]:
        assert g.add_edge(u, v)
    assert not g.add_edge(1, 2)
    assert g.ne() == 4 and g.outneighbors(1).tolist() == [2, 3] and g.inneighbors(3).tolist() == [1, 2]
    g8 = g.with_eltype(np.uint8)
    assert g8.T == np.uint8 and g8.ne() == 4
    u = lgmi.SimpleGraph.from_digraph(g)
    assert u.ne() == 4 and u.outneighbors(3).tolist() == [1, 2, 4]
    # self-loop stored once (s1 of test/centrality/betweenness.jl:3-5)
    s = lgmi.SimpleGraph(3)
    s.add_edge(1, 2); s.add_edge(2, 3); s.add_edge(3, 3)
    assert s.ne() == 3 and s.outneighbors(3).tolist() == [2, 3]
    assert s.degree_array().tolist() == [1, 2, 2]


def test_weighted_distmx_is_not_accelerated():
    import lgmi
    g = lgmi.generators.path_graph(3)
    with pytest.raises(NotImplementedError):
        lgmi.betweenness_centrality(g, None, np.ones((3, 3)))
