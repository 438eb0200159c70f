"""loadgraph / loadgraphs / savegraph with the graphs built on the device
(lgmi_graph_from_edges), against the reference's persistence tests
(test/persistence/persistence.jl:6-52) and the host-built graphs.  Bar:
exact."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu
LG = os.path.join(GOLDEN, "lg")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same(a, b):
    if (a.nv(), a.ne(), a.is_directed()) != (b.nv(), b.ne(), b.is_directed()):
        return False
    ra, ca = a.csr()
    rb, cb = b.csr()
    return np.array_equal(ra, rb) and np.array_equal(ca, cb)


def test_loadgraphs_reference_file():
    import lgmi
    from lgmi.persistence import loadgraph, loadgraphs, sprint_savegraph
    say("loadgraphs tutte-pathdigraph.jgz")
    d = loadgraphs(os.path.join(LG, "tutte-pathdigraph.jgz"))
    p1, p2 = d["Tutte"], d["pathdigraph"]
    assert (p2.ne(), p2.nv()) == (9, 10)
    assert (p1.nv(), p1.ne()) == (46, 69) and (np.diff(p1.csr()[0]) == 3).all()
    assert len(sprint_savegraph(p1)) == 421 and len(sprint_savegraph(p2)) == 70
    # the Tutte graph's diameter is 8 and its radius 5, computed on the device
    assert lgmi.diameter(p1) == 8 and lgmi.radius(p1) == 5
    gs = loadgraph(os.path.join(LG, "tutte-pathdigraph.jgz"), "pathdigraph")
    assert same(gs, p2)
    assert lgmi.gdistances(gs, 1).tolist() == list(range(10))
    with pytest.raises(ValueError):
        loadgraph(os.path.join(LG, "tutte-pathdigraph.jgz"), "badname")


def test_save_load_round_trip(tmp_path):
    from lgmi.persistence import loadgraph, loadgraphs, savegraph
    d = loadgraphs(os.path.join(LG, "tutte-pathdigraph.jgz"))
    p1, p2 = d["Tutte"], d["pathdigraph"]
    f = str(tmp_path / "g.lg")
    for g in (p1, p2):
        say(f"round trip nv={g.nv()}")
        assert savegraph(f, g) == 1
        assert same(loadgraph(f), g) and same(loadgraph(f, "graph"), g)
    assert savegraph(f, {"p1": p1, "p2": p2}) == 2
    back = loadgraphs(f)
    assert sorted(back) == ["p1", "p2"] and same(back["p1"], p1) and same(back["p2"], p2)


def test_graph_50_500_betweenness():
    """The betweenness fixture graph loaded from its LG file on the device;
    betweenness against graph-50-500-bc.txt (test/centrality/betweenness.jl:
    8-13), at the FP64 bar."""
    import lgmi
    from lgmi.persistence import loadgraph
    say("loadgraph graph-50-500.jgz")
    g = loadgraph(os.path.join(LG, "graph-50-500.jgz"), "graph-50-500")
    z = load_golden("ref_graph_50_500.npz")
    bc = lgmi.betweenness_centrality(g)
    np.testing.assert_allclose(bc, z["bc"], rtol=1e-12, atol=1e-15)


def test_add_edge_semantics_and_types():
    """Out-of-range endpoints and duplicates dropped, self-loops kept, blank
    lines counted (lg.jl:39-50 with add_edge!); the header's element type;
    a value beyond it is an OverflowError."""
    from lgmi.persistence import loadgraph, loadlg_mult
    g = loadgraph(b"4,5,u,g\n1,2\n2,1\n3,3\n1,9\n\n", "g")
    assert (g.nv(), g.ne()) == (4, 2) and g.csr()[1].tolist() == [1, 0, 2]
    g = loadgraph(b"6,2,d,graph,2,Int32,simplegraph\n2,1\n6,1\n")
    assert g.T == np.int32 and (g.nv(), g.ne()) == (6, 2)
    assert g.in_csr()[1].tolist() == [1, 5]
    empty = loadgraph(b"3,0,u,graph\n")
    assert (empty.nv(), empty.ne()) == (3, 0)
    with pytest.raises(OverflowError):
        loadlg_mult(b"3,1,u,g,2,UInt8,simplegraph\n1,300\n")


def test_larger_file_vs_host(tmp_path):
    """A generated graph written with savegraph and read back through the
    device ingest equals the original, gzip or not."""
    import gzip

    import lgmi
    from lgmi.persistence import loadgraph, savegraph
    g = lgmi.generators.rmat(14, 8, seed=3)
    f = str(tmp_path / "r.lg")
    savegraph(f, g, "rmat")
    say("loadgraph rmat14")
    assert same(loadgraph(f, "rmat"), g)
    with open(f, "rb") as src, gzip.open(str(tmp_path / "r.lg.gz"), "wb") as dst:
        dst.write(src.read())
    assert same(loadgraph(str(tmp_path / "r.lg.gz"), "rmat"), g)
