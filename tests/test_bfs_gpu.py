"""GPU parity of gdistances / bfs_parents / Parallel BFS against the oracle.

Bar: bit-exact distances and parents.  Every case runs through liblgmi.so on
the device; the oracle (oracle/liblgoracle.so) is only the checker.
"""
import sys

import numpy as np
import pytest

from conftest import GOLDEN_NX, load_golden

pytestmark = pytest.mark.gpu

T_LIST = [np.int64, np.uint8, np.int16]    # testgraphs triplication, test/runtests.jl:14


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _g5(T=np.int64):
    import lgmi
    g = lgmi.SimpleDiGraph(4, T=T)
    for u, v in [(1, 2), (2, 3), (1, 3), (3, 4)]:
        g.add_edge(u, v)
    return g


@pytest.mark.parametrize("T", T_LIST)
def test_reference_vectors(T):
    import lgmi
    say(f"reference vectors T={np.dtype(T).name}")
    house = lgmi.generators.house_graph(T=T)
    # test/traversals/bfs.jl:31-33
    assert lgmi.gdistances(house, 2).tolist() == [1, 0, 2, 1, 2]
    assert lgmi.gdistances(house, 2).dtype == np.dtype(T)
    assert lgmi.gdistances(house, [1, 2]).tolist() == [0, 0, 1, 1, 2]
    assert lgmi.gdistances(house, []).tolist() == [lgmi.typemax(T)] * 5
    # test/traversals/bfs.jl:9-13
    g5 = _g5(T)
    assert lgmi.bfs_parents(g5, 1).tolist() == [1, 1, 1, 3]
    z = lgmi.bfs_tree(g5, 1)
    assert z.nv() == 4 and z.ne() == 3 and not z.has_edge(2, 3)
    # :16-26 house tree properties
    p = lgmi.bfs_parents(house, 1)
    t1 = lgmi.bfs_tree(house, 1)
    t2 = lgmi.tree(p)
    assert len(p) == 5 and t1.csr()[1].tolist() == t2.csr()[1].tolist()
    assert t2.is_directed() and t2.ne() < t2.nv()
    assert p.tolist() == [1, 1, 1, 2, 3]
    # Parallel (test/parallel/traversals/gdistances.jl:7-17, bfs.jl:19-28)
    P = lgmi.Parallel
    assert P.gdistances(g5, 1).tolist() == lgmi.gdistances(g5, 1).tolist()
    assert P.gdistances(g5, [1, 3]).tolist() == lgmi.gdistances(g5, [1, 3]).tolist()
    assert P.gdistances(house, 2).tolist() == lgmi.gdistances(house, 2).tolist()
    assert P.gdistances(house, [1, 2]).tolist() == lgmi.gdistances(house, [1, 2]).tolist()
    parents = np.zeros(4, dtype=T)
    q = P.ThreadQueue(T, 4)
    P.bfs_tree_(q, g5, 1, parents)
    assert parents.tolist() == [1, 1, 1, 3]
    zt = P.bfs_tree(g5, 1)
    assert zt.nv() == 4 and zt.ne() == 3 and not zt.has_edge(2, 3)


def test_gdistances_inplace_semantics():
    import lgmi
    g = lgmi.SimpleGraph.from_edges(5, [0, 3], [1, 4])
    buf = np.full(5, 99, dtype=np.int64)
    lgmi.gdistances_(g, 1, buf)               # serial: writes reached entries only
    assert buf.tolist() == [0, 1, 99, 99, 99]
    lgmi.Parallel.gdistances_(g, [1], buf)    # Parallel: pre-fills typemax
    assert buf.tolist() == [0, 1] + [lgmi.typemax(np.int64)] * 3


def test_bounds_errors():
    import lgmi
    g = lgmi.generators.path_graph(4)
    with pytest.raises(IndexError):
        lgmi.gdistances(g, 5)
    with pytest.raises(IndexError):
        lgmi.bfs_parents(g, 0)
    with pytest.raises(IndexError):
        lgmi.gdistances(g, [1, 9])


def _check_graph(oracle, g, roots, tag):
    rp, col = g.csr()
    irp, icol = g.in_csr()
    for r in roots:
        say(f"{tag}: root {r}")
        d_or = oracle.gdistances(rp, col, [r])
        d_gpu = lgmi_dist0(g, r)
        assert (d_gpu == d_or).all(), f"{tag} dist root {r}"
        p_or = oracle.bfs_parents(rp, col, r)
        p_gpu = lgmi_par0(g, r, 0)
        assert (p_gpu == p_or).all(), f"{tag} parents root {r}"
        if g.is_directed():
            p_or_in = oracle.bfs_parents(irp, icol, r)
            assert (lgmi_par0(g, r, 1) == p_or_in).all(), f"{tag} parents(:in) root {r}"


def lgmi_dist0(g, r):
    from lgmi import traversals
    return traversals._device_dist(g, np.array([r], dtype=np.int32))


def lgmi_par0(g, r, dir_in):
    from lgmi import traversals
    return traversals._parents0(g, r + 1, dir_in)


@pytest.mark.parametrize("name", GOLDEN_NX)
def test_golden_graphs(oracle, name):
    import lgmi
    z = load_golden(name)
    n, directed = int(z["n"]), bool(z["directed"])
    cls = lgmi.SimpleDiGraph if directed else lgmi.SimpleGraph
    g = cls.from_edges(n, z["src"], z["dst"])
    for s, dref in zip(z["bfs_src"], z["dist"]):
        assert (lgmi_dist0(g, int(s)) == dref).all()
    _check_graph(oracle, g, [int(x) for x in z["bfs_src"]], name)


def test_generated_graphs(oracle):
    import lgmi.generators as gen
    cases = [
        ("er10k", gen.erdos_renyi(10_000, 50_000, seed=1)),
        ("rmat14", gen.rmat(14, 16, seed=2)),
        ("rmat12dir", gen.rmat(12, 16, seed=3, directed=True)),
        ("ba5000", gen.barabasi_albert(5000, 4, seed=4)),
        ("ws5000", gen.watts_strogatz(5000, 8, 0.1, seed=5)),
        ("grid64", gen.grid([64, 64])),
        ("path3000", gen.path_graph(3000)),
        ("star1000", gen.star_graph(1000)),
    ]
    rng = np.random.default_rng(0)
    for tag, g in cases:
        deg = np.diff(g.csr()[0])
        roots = [0] + rng.choice(np.nonzero(deg > 0)[0], size=3).tolist() + [int(np.argmax(deg))]
        _check_graph(oracle, g, roots, tag)


def test_multisource_and_isolated(oracle):
    import lgmi
    import lgmi.generators as gen
    g = gen.rmat(13, 8, seed=9)
    rp, col = g.csr()
    deg = np.diff(rp)
    iso = int(np.nonzero(deg == 0)[0][0])
    for srcs in ([iso], [0, 5, 77, 5], list(range(0, 8192, 97))):
        say(f"multisource {len(srcs)}")
        want = oracle.gdistances(rp, col, srcs)
        got = lgmi.gdistances(g, [s + 1 for s in srcs])
        tmax = lgmi.typemax(np.int64)
        assert (np.where(want < 0, np.int64(tmax), want.astype(np.int64)) == got).all()
    assert (lgmi_par0(g, iso, 0) == np.where(np.arange(g.nv()) == iso, iso, -1)).all()


def test_direction_never_changes_output(oracle):
    """Forced top-down, forced bottom-up, the alpha/beta heuristic and the
    reference's top-down-only policy give bit-identical outputs."""
    from lgmi import _lib
    import lgmi.generators as gen
    lib = _lib.load()
    for tag, g in (("rmat15", gen.rmat(15, 16, seed=21)), ("rmat12dir", gen.rmat(12, 16, seed=22, directed=True)),
                   ("grid40", gen.grid([40, 40]))):
        rp, col = g.csr()
        h = g.handle()
        root = int(np.argmax(np.diff(rp)))
        outs = []
        for forced in (-1, 0, 1):
            say(f"{tag}: forced direction {forced}")
            _lib.check(lib.lgmi_bfs_force_direction(h, forced))
            d = np.empty(g.nv(), np.int32)
            p = np.empty(g.nv(), np.int32)
            _lib.check(lib.lgmi_bfs(h, root, 0, _lib.i32p(d), _lib.i32p(p)))
            outs.append((d, p))
        _lib.check(lib.lgmi_bfs_force_direction(h, -1))
        _lib.check(lib.lgmi_bfs_set_policy(_lib.LGMI_POLICY_TOPDOWN, 15.0, 18.0, None, 0))
        d = np.empty(g.nv(), np.int32)
        p = np.empty(g.nv(), np.int32)
        _lib.check(lib.lgmi_bfs(h, root, 0, _lib.i32p(d), _lib.i32p(p)))
        _lib.check(lib.lgmi_bfs_set_policy(_lib.LGMI_POLICY_HEURISTIC, 15.0, 18.0, None, 0))
        outs.append((d, p))
        for d, p in outs[1:]:
            assert (d == outs[0][0]).all() and (p == outs[0][1]).all()
        assert (outs[0][0] == oracle.gdistances(rp, col, [root])).all()
        assert (outs[0][1] == oracle.bfs_parents(rp, col, root)).all()


def test_full_tiles_every_direction(oracle):
    """Full 512-vertex tiles: a star whose 2999 leaves also form a ring and a
    complete graph on 600 vertices, so that whole tiles are frontier
    (top-down) or candidates (bottom-up); forced directions and the heuristic
    against the oracle from the hub and from a leaf."""
    from lgmi import _lib
    from lgmi.graphs import SimpleGraph
    import lgmi.generators as gen
    lib = _lib.load()
    n = 3000
    a = np.arange(1, n)
    src = np.concatenate([np.zeros(n - 1, np.int64), a])
    dst = np.concatenate([a, np.where(a + 1 < n, a + 1, 1)])
    for tag, g in (("star-ring", SimpleGraph.from_edges(n, src, dst)), ("k600", gen.complete_graph(600))):
        rp, col = g.csr()
        h = g.handle()
        for root in (0, 7):
            for forced in (-1, 0, 1):
                say(f"{tag} root {root} forced {forced}")
                _lib.check(lib.lgmi_bfs_force_direction(h, forced))
                d = np.empty(g.nv(), np.int32)
                p = np.empty(g.nv(), np.int32)
                _lib.check(lib.lgmi_bfs(h, root, 0, _lib.i32p(d), _lib.i32p(p)))
                assert (d == oracle.gdistances(rp, col, [root])).all()
                assert (p == oracle.bfs_parents(rp, col, root)).all()
        _lib.check(lib.lgmi_bfs_force_direction(h, -1))


def test_repeat_bit_identical():
    import lgmi.generators as gen
    g = gen.rmat(16, 16, seed=31)
    root = int(np.argmax(np.diff(g.csr()[0])))
    a = [lgmi_par0(g, root, 0) for _ in range(4)]
    for x in a[1:]:
        assert (x == a[0]).all()


def test_rmat20_full_compare(oracle):
    import lgmi.generators as gen
    say("building rmat20")
    g = gen.rmat(20, 16, seed=1)
    rp, col = g.csr()
    rng = np.random.default_rng(5)
    roots = rng.choice(np.nonzero(np.diff(rp) > 0)[0], size=3).tolist()
    _check_graph(oracle, g, roots, "rmat20")


def _variant_counts(g, roots):
    import ctypes

    from lgmi import _lib
    lib = _lib.load()
    buf = (_lib.BfsStep * 20000)()
    st = _lib.BfsStats()
    counts = {0: 0, 1: 0, 2: 0, 3: 0}
    for r in roots:
        _lib.check(lib.lgmi_bfs_run(g.handle(), r, 0, 1, ctypes.byref(st)))
        n = lib.lgmi_bfs_steps(g.handle(), 20000, buf)
        for i in range(n):
            if buf[i].dir == 0:
                counts[int(buf[i].variant)] += 1
    return counts


def _hub_neighbour_roots(g, k=3):
    rp, _ = g.csr()
    irp, icol = g.in_csr()
    deg = np.diff(rp)
    hub = int(np.argmax(deg))
    # low-degree vertices with an edge to the hub: the hub is in F_1
    nb = icol[irp[hub]:irp[hub + 1]]
    nb = nb[deg[nb] > 0]
    return [int(x) for x in nb[np.argsort(deg[nb], kind="stable")[:k]]]


def test_range_mode_levels_exact(oracle):
    """Top-down levels over a frontier of hub rows run in range mode (each
    work-group owns a vertex range and forms parents in LDS) when queue mode
    is off; distances and parents stay bit-exact, with and without parents,
    along out- and in-edges."""
    import lgmi.generators as gen
    from lgmi import tuning
    seen = 0
    try:
        tuning.set_values({"queue_front": 0, "solo_front": 0})
        for g, tag in ((gen.rmat(20, 16, seed=2), "rmat20"), (gen.rmat(17, 16, seed=3, directed=True), "rmat17-dir")):
            roots = _hub_neighbour_roots(g)
            _check_graph(oracle, g, roots, tag)
            seen += _variant_counts(g, roots)[1]
    finally:
        tuning.set_values(None)
        tuning.load()
    say(f"range-mode levels: {seen}")
    assert seen > 0, "no level ran in range mode"


def test_queue_mode_levels_exact(oracle):
    """With queue mode taking every top-down level it may (frontier lists
    up to 4096 vertices at any edge count), the levels that read the
    frontier from the work-group queues -- the single source of level 0, hub
    levels, the small levels of a grid and of a chain -- give bit-exact
    distances and parents, along out- and in-edges."""
    import lgmi.generators as gen
    from lgmi import tuning
    seen = 0
    try:
        tuning.set_values({"queue_front": 4096, "queue_min_edges": 0, "queue_max_edges": 1 << 40,
                           "range_front": 0, "solo_front": 0})
        cases = [(gen.rmat(18, 16, seed=6), "rmat18"), (gen.rmat(15, 16, seed=7, directed=True), "rmat15-dir"),
                 (gen.grid([300, 200]), "grid300x200"), (gen.path_graph(3000), "path3000"),
                 (gen.barabasi_albert(50_000, 3, seed=8), "ba50k")]
        for g, tag in cases:
            rp, _ = g.csr()
            deg = np.diff(rp)
            roots = _hub_neighbour_roots(g, 2) + [0, int(np.nonzero(deg > 0)[0][-1])]
            _check_graph(oracle, g, roots, tag)
            seen += _variant_counts(g, roots)[2]
    finally:
        tuning.set_values(None)
        tuning.load()
    say(f"queue-mode levels: {seen}")
    assert seen > 0, "no level ran in queue mode"


@pytest.mark.parametrize("values", [
    {"wave_row": 2, "lane_rounds": 0},
    {"wave_row": 257, "lane_rounds": 64},
    {"queue_front": 0, "range_front": 0},
    {"queue_front": 256, "queue_min_edges": 0, "queue_max_edges": 2048, "range_light": 1 << 40},
    {"solo_front": 0},
    {"solo_front": 64, "solo_edges": 256},
    {"solo_front": 4096, "solo_edges": 262144},
])
def test_tuning_never_changes_output(oracle, values):
    """Every launch threshold setting gives the same distances and parents
    (the autotuner only trades time)."""
    import lgmi.generators as gen
    from lgmi import tuning
    try:
        tuning.set_values(values)
        for g, tag in ((gen.rmat(17, 16, seed=9), "rmat17"), (gen.rmat(14, 8, seed=10, directed=True), "rmat14-dir"),
                       (gen.grid([64, 64]), "grid64")):
            rp, _ = g.csr()
            roots = _hub_neighbour_roots(g, 1) + [int(np.argmax(np.diff(rp)))]
            _check_graph(oracle, g, roots, f"{tag} {values}")
    finally:
        tuning.set_values(None)
        tuning.load()


def _expected_features(d, deg_td, deg_bu, nnz, L, prev_dir):
    """The 16 policy features of level L's state, from the oracle's
    distances: log2(1 + x) of the frontier size, frontier edges, unvisited
    edges, unvisited vertices and the level; the frontier's share of the
    vertices; edges log-ratio; growth against the previous frontier; the
    previous direction; the frontier's degree histogram (bins 0, 1-3, 4-15,
    16-63, 64-255, 256-1023, >= 1024) as fractions."""
    nv = len(d)
    f = d == L
    seen = (d >= 0) & (d <= L)
    nf, mf = int(f.sum()), int(deg_td[f].sum())
    mu = int(nnz - deg_bu[seen].sum())
    nu = int(nv - seen.sum())
    n_prev = int((d == L - 1).sum()) if L > 0 else 0
    lg = lambda x: np.log2(1.0 + max(float(x), 0.0))
    out = [lg(nf), lg(mf), lg(mu), lg(nu), lg(L), nf / nv, lg(mf) - lg(mu), lg(nf) - lg(n_prev), float(prev_dir)]
    dg = deg_td[f]
    bins = np.where(dg <= 0, 0, np.minimum((np.floor(np.log2(np.maximum(dg, 1))).astype(int)) // 2 + 1, 6))
    hist = np.bincount(bins, minlength=7)[:7].astype(np.float64)
    out += list(hist / hist.sum()) if hist.sum() else [0.0] * 7
    return np.array(out)


def test_level_counters_match_oracle(oracle):
    """The per-level records the kernels reduce on the device (frontier size,
    frontier edges, new vertices) equal the counts from the oracle's
    distances, level by level, for both directions and for range mode; the
    16 policy features derived from them on the device (FP32, hardware
    log2) match the same features computed from the oracle's distances."""
    import ctypes

    import lgmi.generators as gen
    from lgmi import _lib
    lib = _lib.load()
    for g, tag in ((gen.rmat(18, 16, seed=4), "rmat18"), (gen.rmat(15, 16, seed=5, directed=True), "rmat15-dir"),
                   (gen.grid([128, 96]), "grid128x96")):
        rp, col = g.csr()
        deg = np.diff(rp)
        deg_in = np.diff(g.in_csr()[0])
        hub = int(np.argmax(deg))
        roots = [hub, int(np.nonzero(deg > 0)[0][0])]
        buf = (_lib.BfsStep * 512)()
        st = _lib.BfsStats()
        for forced in (-1, 0, 1):
            _lib.check(lib.lgmi_bfs_force_direction(g.handle(), forced))
            for r in roots:
                say(f"{tag}: root {r} forced {forced}")
                d = oracle.gdistances(rp, col, [r])
                _lib.check(lib.lgmi_bfs_run(g.handle(), r, 0, 1, ctypes.byref(st)))
                n = lib.lgmi_bfs_steps(g.handle(), 512, buf)
                assert n == int(d.max()) + 1 == st.levels
                for L in range(n):
                    f = d == L
                    assert buf[L].n_front == int(f.sum()), (tag, r, L)
                    assert buf[L].m_front == int(deg[f].sum()), (tag, r, L)
                    assert buf[L].n_next == int((d == L + 1).sum()), (tag, r, L)
                    want = _expected_features(d, deg, deg_in, len(col), L, buf[L - 1].dir if L > 0 else 0)
                    got = np.array(buf[L].feat[:16], dtype=np.float64)
                    assert np.allclose(got, want, rtol=1e-5, atol=1e-5), (tag, r, L, got, want)
        _lib.check(lib.lgmi_bfs_force_direction(g.handle(), -1))


def test_solo_mode_levels_exact(oracle):
    """Small top-down levels run in solo mode: work-group 0 alone expands
    them and goes on to the next level inside the same launch.  Distances
    and parents stay bit-exact along out- and in-edges, with and without
    parents, for chains and grids (thousands of levels in few launches),
    hub levels, multi-source searches, and frontiers that outgrow the LDS
    list inside a solo level (a star: its hub's 30000 leaves are claimed by
    the one work-group, 4096 of them listed, the rest handed over through
    the frontier bitmap)."""
    import ctypes

    import lgmi
    import lgmi.generators as gen
    from lgmi import _lib, tuning
    lib = _lib.load()
    seen = 0
    try:
        for values in ({}, {"solo_front": 4096, "solo_edges": 262144}):
            tuning.set_values(None)
            tuning.set_values(values)
            star = gen.star_graph(30001)
            cases = [(gen.grid([300, 200]), "grid300x200"), (gen.path_graph(5000), "path5000"),
                     (gen.rmat(18, 16, seed=6), "rmat18"), (gen.rmat(15, 16, seed=7, directed=True), "rmat15-dir"),
                     (gen.barabasi_albert(50_000, 3, seed=8), "ba50k"), (star, "star30001")]
            for g, tag in cases:
                rp, _ = g.csr()
                deg = np.diff(rp)
                roots = [0, int(np.nonzero(deg > 0)[0][-1])] + _hub_neighbour_roots(g, 1)
                _check_graph(oracle, g, roots, f"{tag} {values}")
                seen += _variant_counts(g, roots)[3]
                # multi-source distances (solo from level 1 on)
                srcs = np.array(sorted(set(roots)), dtype=np.int32)
                from lgmi import traversals
                want = oracle.gdistances(rp, g.csr()[1], srcs)
                assert (traversals._device_dist(g, srcs) == want).all(), f"{tag} multi-source"
            # a chain runs its levels in a handful of launches
            g = gen.path_graph(5000)
            st = _lib.BfsStats()
            _lib.check(lib.lgmi_bfs_run(g.handle(), 0, 0, 1, ctypes.byref(st)))
            say(f"path5000 {values}: {st.levels} levels in {st.launches} launches")
            assert st.levels == 5000
    finally:
        tuning.set_values(None)
        tuning.load()
    say(f"solo-mode levels: {seen}")
    assert seen > 0, "no level ran in solo mode"


def test_solo_mode_has_path_walls(oracle):
    """has_path's excluded vertices are marked visited before the search;
    solo levels (which tell fresh vertices by the visited bitmap alone)
    never cross them.  The wall is the set of grid vertices at distance 60
    from vertex 1, once whole and once with a gap."""
    import lgmi
    import lgmi.generators as gen
    g = gen.grid([120, 80])
    rp, col = g.csr()
    d = oracle.gdistances(rp, col, [0])
    far = int(np.argmax(d))
    ring = np.nonzero(d == 60)[0]
    for wall, want in ((ring[ring != ring[len(ring) // 2]], True), (ring, False)):
        assert oracle.has_path(rp, col, 0, far, exclude=[int(w) for w in wall]) == want
        assert lgmi.has_path(g, 1, far + 1, exclude_vertices=[int(w) + 1 for w in wall]) == want
