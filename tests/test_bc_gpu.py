"""GPU parity of betweenness_centrality (serial and Parallel) against the
reference fixture, networkx vectors and the oracle.

Bar: within 1e-12 relative error of the reference's FP64 values (the only
difference is floating-point summation order); exact where the value is a
small dyadic/rational the reference tests compare with ==.
"""
import sys

import numpy as np
import pytest

from conftest import GOLDEN_NX, load_golden

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def close(got, want, rtol=RTOL):
    want = np.asarray(want, dtype=np.float64)
    scale = np.maximum(np.abs(want), 1e-300)
    err = np.abs(got - want)
    ok = (err <= rtol * scale) | (err <= 1e-15 * max(1.0, float(np.abs(want).max(initial=0.0))))
    return bool(ok.all())


def test_reference_fixture_graph_50_500():
    import lgmi
    z = load_golden("ref_graph_50_500.npz")
    n = int(z["n"])
    for T in (np.int64, np.uint8, np.int16):
        say(f"graph-50-500 T={np.dtype(T).name}")
        g = lgmi.SimpleDiGraph.from_edges(n, z["src"], z["dst"], T=T)
        bc = lgmi.betweenness_centrality(g)
        # test/centrality/betweenness.jl:13 (Float32 equality) and FP64 1e-12
        assert (bc.astype(np.float32) == z["bc"].astype(np.float32)).all()
        assert close(bc, z["bc"])
        y = lgmi.betweenness_centrality(g, endpoints=True, normalize=False)
        ref = [122.10760591498584, 159.0072453120582, 176.39547945994505]
        assert np.round(y[:3], 4).tolist() == np.round(ref, 4).tolist()
        assert close(y[:3], ref)
        x = lgmi.betweenness_centrality(g, 3, rng=1)
        x2 = lgmi.betweenness_centrality(g, list(range(1, 21)))
        assert len(x) == 50 and len(x2) == 50
        # Parallel, both modes (test/parallel/centrality/betweenness.jl:10-30)
        P = lgmi.Parallel
        assert close(P.betweenness_centrality(g, parallel="threads"), bc)
        assert close(P.betweenness_centrality(g, parallel="distributed"), bc)
        assert close(P.betweenness_centrality(g, endpoints=True, normalize=False, parallel="threads"), y)
        assert len(P.betweenness_centrality(g, 3, parallel="threads", rng=2)) == 50
        assert len(P.betweenness_centrality(g, list(range(1, 21)), parallel="distributed")) == 50


def test_small_reference_cases():
    import lgmi
    gen = lgmi.generators
    s2 = lgmi.SimpleDiGraph(3)
    s2.add_edge(1, 2); s2.add_edge(2, 3); s2.add_edge(3, 3)
    s1 = lgmi.SimpleGraph.from_digraph(s2)
    # test/centrality/betweenness.jl:20-21
    assert lgmi.betweenness_centrality(s1).tolist() == [0, 1, 0]
    assert lgmi.betweenness_centrality(s2).tolist() == [0, 0.5, 0]
    for mode in ("threads", "distributed"):
        assert lgmi.Parallel.betweenness_centrality(s1, parallel=mode).tolist() == [0, 1, 0]
        assert lgmi.Parallel.betweenness_centrality(s2, parallel=mode).tolist() == [0, 0.5, 0]
    # :23-29
    k2 = gen.path_graph(2)
    z = lgmi.betweenness_centrality(k2, normalize=True)
    assert z[0] == z[1] == 0.0
    assert (lgmi.betweenness_centrality(k2, [1, 2]) == z).all()
    # :32-33
    z = lgmi.betweenness_centrality(gen.path_graph(5), normalize=False)
    assert z[0] == z[4] == 0.0 and z.tolist() == [0.0, 3.0, 4.0, 3.0, 0.0]
    # doctests betweenness.jl:32-43
    assert lgmi.betweenness_centrality(gen.star_graph(3)).tolist() == [1.0, 0.0, 0.0]
    assert close(lgmi.betweenness_centrality(gen.path_graph(4)), [0.0, 2 / 3, 2 / 3, 0.0])
    # endpoints=true on a disconnected graph: hand-derived closed form
    g = lgmi.SimpleGraph.from_edges(4, [0, 2], [1, 3])
    assert lgmi.betweenness_centrality(g, normalize=False, endpoints=True).tolist() == [3.0] * 4


def test_grid_50_bound(oracle):
    import lgmi
    g = lgmi.generators.grid([50, 50])
    z = lgmi.betweenness_centrality(g, normalize=False)
    assert z.max() < 2500 * 2499                  # test/centrality/betweenness.jl:75-77
    rp, col = g.csr()
    want = oracle.betweenness_threaded(rp, col, rp, False, np.arange(2500), normalize=False, nthreads=8,
                                       index_quirk=False)
    assert close(z, want)


@pytest.mark.parametrize("name", GOLDEN_NX)
def test_networkx_golden(name):
    import lgmi
    z = load_golden(name)
    n, directed = int(z["n"]), bool(z["directed"])
    say(f"golden {name}")
    cls = lgmi.SimpleDiGraph if directed else lgmi.SimpleGraph
    g = cls.from_edges(n, z["src"], z["dst"])
    assert close(lgmi.betweenness_centrality(g), z["bc_norm"])
    assert close(lgmi.betweenness_centrality(g, normalize=False), z["bc_raw"])
    assert close(lgmi.betweenness_centrality(g, z["sub_src"] + 1), z["bc_sub_norm"])
    if "bc_end_raw" in z:
        assert close(lgmi.betweenness_centrality(g, normalize=False, endpoints=True), z["bc_end_raw"])


def test_sources_quirks(oracle):
    """Degree-0 and duplicate sources, k = length(vs), the threaded
    findall-position quirk, and more than one 64-lane group."""
    import lgmi
    g = lgmi.generators.rmat(10, 4, seed=3)
    rp, col = g.csr()
    n = g.nv()
    deg = np.diff(rp)
    iso = np.nonzero(deg == 0)[0]
    rng = np.random.default_rng(4)
    vs0 = np.concatenate([rng.integers(0, n, 150), iso[:5], [7, 7, 7]])
    rng.shuffle(vs0)
    for endpoints in (False, True):
        for normalize in (True, False):
            say(f"quirks endpoints={endpoints} normalize={normalize}")
            got = lgmi.betweenness_centrality(g, vs0 + 1, normalize=normalize, endpoints=endpoints)
            want = oracle.betweenness(rp, col, rp, False, vs0, normalize=normalize, endpoints=endpoints)
            assert close(got, want)
            got_t = lgmi.Parallel.betweenness_centrality(g, vs0 + 1, normalize=normalize, endpoints=endpoints,
                                                         parallel="threads")
            want_t = oracle.betweenness_threaded(rp, col, rp, False, vs0, normalize=normalize,
                                                 endpoints=endpoints, nthreads=4, index_quirk=True)
            assert close(got_t, want_t)


def test_er10k_all_sources(oracle):
    """Config: ER(10_000, 50_000, seed=1) all-source BC, normalize=true,
    endpoints=false; and gdistances from vertex 1."""
    import lgmi
    g = lgmi.generators.erdos_renyi(10_000, 50_000, seed=1)
    rp, col = g.csr()
    say("er10k bc")
    bc = lgmi.betweenness_centrality(g)
    want = oracle.betweenness_threaded(rp, col, rp, False, np.arange(10_000), nthreads=16, index_quirk=False)
    assert close(bc, want)
    d = lgmi.gdistances(g, 1)
    dref = oracle.gdistances(rp, col, [0])
    assert (np.where(dref < 0, np.int64(lgmi.typemax(np.int64)), dref.astype(np.int64)) == d).all()


def test_directed_rmat_subset(oracle):
    import lgmi
    g = lgmi.generators.rmat(12, 8, seed=5, directed=True)
    rp, col = g.csr()
    irp, _ = g.in_csr()
    rng = np.random.default_rng(8)
    vs0 = rng.integers(0, g.nv(), 200)
    got = lgmi.betweenness_centrality(g, vs0 + 1)
    want = oracle.betweenness(rp, col, irp, True, vs0)
    assert close(got, want)
    got = lgmi.betweenness_centrality(g, vs0 + 1, normalize=False, endpoints=True)
    want = oracle.betweenness(rp, col, irp, True, vs0, normalize=False, endpoints=True)
    assert close(got, want)


def test_bc_deterministic():
    import lgmi
    g = lgmi.generators.barabasi_albert(3000, 4, seed=6)
    a = lgmi.betweenness_centrality(g)
    b = lgmi.betweenness_centrality(g)
    assert (a.view(np.uint64) == b.view(np.uint64)).all()


def test_rmat16_sources_vs_oracle(oracle):
    import lgmi
    g = lgmi.generators.rmat(16, 16, seed=7)
    rp, col = g.csr()
    deg = np.diff(rp)
    rng = np.random.default_rng(9)
    vs0 = rng.choice(np.nonzero(deg > 0)[0], size=256, replace=False)
    say("rmat16 256 sources")
    got = lgmi.betweenness_centrality(g, vs0 + 1)
    want = oracle.betweenness_threaded(rp, col, rp, False, vs0, nthreads=16, index_quirk=False)
    assert close(got, want)


def _bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def _directed_hubs():
    """A directed graph whose out-row of vertex 0 and in-row of vertex 4999
    are longer than the heavy-row threshold (rows cut into segments)."""
    import lgmi
    rng = np.random.default_rng(15)
    n = 5000
    src = np.concatenate([rng.integers(0, n, 20000), np.zeros(3000, np.int64), np.arange(1, 3001)])
    dst = np.concatenate([rng.integers(0, n, 20000), np.arange(1, 3001), np.full(3000, n - 1)])
    return lgmi.SimpleDiGraph.from_edges(n, src, dst, drop_self_loops=True)


def test_heavy_rows_directed_vs_oracle(oracle):
    """Heavy in-rows (forward) and out-rows (backward) of a directed graph,
    both launch paths, against the oracle."""
    import lgmi
    from lgmi import centrality as C
    g = _directed_hubs()
    rp, col = g.csr()
    irp, _ = g.in_csr()
    assert np.diff(rp).max() > 1024 and np.diff(irp).max() > 1024
    vs0 = np.concatenate([[0, 1, 2, 4999], np.random.default_rng(16).integers(0, 5000, 300)])
    try:
        for mode in (1, 0):
            C.set_launch(mode)
            say(f"directed hubs launch mode {mode}")
            got = lgmi.betweenness_centrality(g, vs0 + 1, normalize=False, endpoints=True)
            want = oracle.betweenness(rp, col, irp, True, vs0, normalize=False, endpoints=True)
            assert close(got, want)
    finally:
        C.set_launch(0)


@pytest.mark.parametrize("seg_len", [64, 4096])
def test_segment_length_never_changes_result(oracle, seg_len):
    """The heavy-row segment length (an autotuned threshold, fixed for a
    graph when its segments are built) changes only time: short segments
    (64 entries, so that most rows of the graph are heavy) and long ones
    give the oracle's betweenness and stress, both launch paths."""
    import lgmi
    from lgmi import centrality as C
    from lgmi import tuning
    g0 = _directed_hubs()
    rp, col = g0.csr()
    irp, icol = g0.in_csr()
    vs0 = np.concatenate([[0, 1, 2, 4999], np.random.default_rng(17).integers(0, 5000, 100)])
    try:
        tuning.set_values({"bc_seg_len": seg_len})
        g = lgmi.SimpleDiGraph.from_csr(rp, col, in_row_ptr=irp, in_col=icol)
        for mode in (1, 0):
            C.set_launch(mode)
            say(f"segment length {seg_len}, launch mode {mode}")
            got = lgmi.betweenness_centrality(g, vs0 + 1, normalize=False)
            want = oracle.betweenness(rp, col, irp, True, vs0, normalize=False)
            assert close(got, want)
        st = lgmi.stress_centrality(g, vs0 + 1)
        assert (st == oracle.stress(rp, col, irp, True, vs0)).all()
    finally:
        C.set_launch(0)
        tuning.set_values(None)
        tuning.load()


def test_persistent_launch_matches_level_launches(oracle):
    """One persistent launch (graph-captured or cooperative) against one
    launch per level: bit-identical results, on undirected and directed
    graphs, with and without endpoints, and a replay with other sources."""
    import lgmi
    from lgmi import centrality as C
    cases = [
        ("rmat14", lgmi.generators.rmat(14, 8, seed=11), 300),
        ("rmat12-directed", lgmi.generators.rmat(12, 8, seed=12, directed=True), 130),
        ("ba2000", lgmi.generators.barabasi_albert(2000, 3, seed=13), 2000),
        ("directed-hubs", _directed_hubs(), 200),
    ]
    try:
        for name, g, k in cases:
            rng = np.random.default_rng(14)
            for rep in range(2):
                vs = rng.choice(g.nv(), size=min(k, g.nv()), replace=False) + 1
                for endpoints in (False, True):
                    say(f"persistent {name} rep={rep} endpoints={endpoints}")
                    C.set_launch(1)
                    ref = lgmi.betweenness_centrality(g, vs, normalize=False, endpoints=endpoints)
                    assert C.last_stats(g)["launch"] == 0
                    C.set_launch(0)
                    got = lgmi.betweenness_centrality(g, vs, normalize=False, endpoints=endpoints)
                    st = C.last_stats(g)
                    assert st["launch"] in (1, 2), st
                    # the work-groups that took part: resident together,
                    # at least half of those launched (6 of the 7 per CU the
                    # occupancy query reports fit), and a multiple that the
                    # lane groups divide evenly (4 wavefronts each)
                    assert st["blocks"] // 2 <= st["participants"] <= st["blocks"], st
                    assert (st["participants"] * 4) % st["groups"] == 0, st
                    assert (_bits(got) == _bits(ref)).all()
        # a replay of the same call comes from the captured graph when the
        # runtime captured it, and gives the same bits
        g = cases[0][1]
        vs = np.arange(1, 200)
        a = lgmi.betweenness_centrality(g, vs)
        b = lgmi.betweenness_centrality(g, vs)
        assert (_bits(a) == _bits(b)).all()
        rp, col = g.csr()
        assert close(a, oracle.betweenness(rp, col, rp, False, vs - 1))
    finally:
        C.set_launch(0)


def test_persistent_level_overflow_falls_back(oracle):
    """A 60x60 grid needs more levels than the first allocation holds: the
    persistent launch stops, the level-per-launch path (which grows the
    level storage) gives the result, and the next call fits in one launch."""
    import lgmi
    from lgmi import centrality as C
    C.set_launch(0)
    g = lgmi.generators.grid([60, 60])
    vs = np.arange(1, 3601, 37)
    say("grid 60x60 first call")
    a = lgmi.betweenness_centrality(g, vs)
    assert C.last_stats(g)["launch"] == 0
    say("grid 60x60 second call")
    b = lgmi.betweenness_centrality(g, vs)
    assert C.last_stats(g)["launch"] in (1, 2)
    assert (_bits(a) == _bits(b)).all()
    rp, col = g.csr()
    assert close(a, oracle.betweenness(rp, col, rp, False, vs - 1))



def test_path_counts_beyond_fp32(oracle):
    """Path counts above 2^24 (the layered graph: 15^7 paths by layer 8, up
    to about 1.1e14), where a count narrowed to FP32 anywhere on the way
    would be rounded: against the oracle at 1e-12 with all sources, with
    sources whose lane groups mix small and large counts, and the persistent
    launch against one launch per level, bit for bit."""
    import lgmi
    from lgmi import centrality as C
    from conftest import layered_edges
    n, s, d = layered_edges()
    g = lgmi.SimpleGraph.from_edges(n, s, d)
    rp, col = g.csr()
    rng = np.random.default_rng(21)
    cases = [np.arange(1, n + 1), np.sort(rng.choice(n, 70, replace=False)) + 1,
             np.arange(n - 39, n + 1), np.array([1])]
    try:
        for vs in cases:
            say(f"layered graph, {len(vs)} sources")
            C.set_launch(0)
            got = lgmi.betweenness_centrality(g, vs, normalize=False)
            want = oracle.betweenness(rp, col, rp, False, vs - 1, normalize=False)
            assert close(got, want)
            C.set_launch(1)
            per_level = lgmi.betweenness_centrality(g, vs, normalize=False)
            assert (_bits(got) == _bits(per_level)).all()
    finally:
        C.set_launch(0)


def _row_length_graph(directed):
    """Vertex v has out-degree (7 v) mod 71, so every chunk of 64 vertices
    holds rows of 0 to 70 entries: rows that share a 64-entry pass, rows of
    exactly 64, rows that take a pass of their own (65-70), empty rows, and
    vertices without out-edges reached in a digraph."""
    import lgmi
    n = 900
    rng = np.random.default_rng(41)
    src, dst = [], []
    for v in range(n):
        d = (7 * v) % 71
        nb = rng.choice(n, size=d, replace=False)
        src.extend([v] * d)
        dst.extend(nb.tolist())
    cls = lgmi.SimpleDiGraph if directed else lgmi.SimpleGraph
    return cls.from_edges(n, np.array(src), np.array(dst), drop_self_loops=True)


@pytest.mark.parametrize("directed", [False, True])
def test_packed_row_lengths_vs_oracle(oracle, directed):
    """Betweenness, stress and closeness on graphs whose rows span the
    packed and one-row-per-pass forms of both sweeps, against the oracle;
    the persistent launch and one launch per level bit for bit."""
    import lgmi
    from lgmi import centrality as C
    g = _row_length_graph(directed)
    rp, col = g.csr()
    irp = g.in_csr()[0] if directed else rp
    n = g.nv()
    vs = np.arange(1, n + 1)
    try:
        say(f"row lengths, directed={directed}")
        C.set_launch(0)
        got = lgmi.betweenness_centrality(g, vs, normalize=False)
        assert close(got, oracle.betweenness(rp, col, irp, directed, vs - 1, normalize=False))
        C.set_launch(1)
        assert (_bits(lgmi.betweenness_centrality(g, vs, normalize=False)) == _bits(got)).all()
        C.set_launch(0)
        st = lgmi.stress_centrality(g)
        assert st.tolist() == oracle.stress(rp, col, irp, directed, np.arange(n), nthreads=8).tolist()
        cl = lgmi.closeness_centrality(g)
        assert cl.tolist() == oracle.closeness(rp, col, irp, directed, normalize=True).tolist()
    finally:
        C.set_launch(0)
