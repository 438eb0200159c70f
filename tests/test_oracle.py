"""Pins the CPU oracle (oracle/lgoracle.c) before it is trusted as the
checker of the GPU path.

Each oracle function is checked against (a) the reference's own test vectors
and fixtures (cited file:line), (b) networkx golden vectors, and (c)
independent brute force on tiny graphs (explicit enumeration of all
shortest paths; the min-predecessor closed form for parents).  A few tests
plant a plausible mistake in an oracle result and check that the same
comparison rejects it.
"""
import itertools

import numpy as np
import pytest

from conftest import GOLDEN_NX, csr_of, load_golden


# ----------------------------------------------------------------------
# small graphs of the reference tests (1-based edge lists)
# ----------------------------------------------------------------------
G5 = (4, True, [(1, 2), (2, 3), (1, 3), (3, 4)])            # test/traversals/bfs.jl:4-5
HOUSE = (5, False, [(1, 2), (1, 3), (2, 4), (3, 4), (3, 5), (4, 5)])   # smallgraphs.jl:202-203
S2 = (3, True, [(1, 2), (2, 3), (3, 3)])                     # test/centrality/betweenness.jl:3-5
G615 = (5, False, [(1, 2), (1, 3), (2, 5), (3, 5), (3, 4), (4, 5)])   # test/shortestpaths/dijkstra.jl:71-77


def mk(spec, drop_self_loops=False):
    n, directed, edges = spec
    e = np.array(edges, dtype=np.int64) - 1 if edges else np.zeros((0, 2), np.int64)
    return n, directed, csr_of(n, e[:, 0], e[:, 1], directed, drop_self_loops)


def path_spec(n):
    return (n, False, [(i, i + 1) for i in range(1, n)])


def star_spec(n):
    return (n, False, [(1, i) for i in range(2, n + 1)])


# ----------------------------------------------------------------------
# (a) reference vectors
# ----------------------------------------------------------------------
def test_gdistances_reference_vectors(oracle):
    n, _, (rp, col, _, _) = mk(HOUSE)
    # test/traversals/bfs.jl:31-33 (1-based values)
    assert (oracle.gdistances(rp, col, [1]) == np.array([1, 0, 2, 1, 2])).all()
    assert (oracle.gdistances(rp, col, [0, 1]) == np.array([0, 0, 1, 1, 2])).all()
    assert (oracle.gdistances(rp, col, np.zeros(0, np.int32)) == -1).all()
    # test/parallel/traversals/gdistances.jl:7-17: Parallel equals serial
    n5, _, (rp5, col5, _, _) = mk(G5)
    for srcs in ([0], [0, 2]):
        assert (oracle.par_gdistances(rp5, col5, srcs) == oracle.gdistances(rp5, col5, srcs)).all()
    for srcs in ([1], [0, 1]):
        assert (oracle.par_gdistances(rp, col, srcs) == oracle.gdistances(rp, col, srcs)).all()


def test_gdistances_inplace_writes_only_reached(oracle):
    # bfs.jl:95-124 never writes unreached entries (the caller pre-fills)
    rp, col, _, _ = csr_of(5, [0, 3], [1, 4], False)
    buf = np.full(5, 99, dtype=np.int32)
    oracle.gdistances_inplace(rp, col, [0], buf)
    assert buf.tolist() == [0, 1, 99, 99, 99]


def test_bfs_parents_reference_vectors(oracle):
    n, _, (rp, col, irp, icol) = mk(G5)
    # test/traversals/bfs.jl:12 and test/parallel/traversals/bfs.jl:26
    assert (oracle.bfs_parents(rp, col, 0) + 1).tolist() == [1, 1, 1, 3]
    assert (oracle.par_bfs_tree(rp, col, 0) + 1).tolist() == [1, 1, 1, 3]
    # bfs_parents(house, 1), worked by hand from bfs.jl:38-66
    _, _, (hrp, hcol, _, _) = mk(HOUSE)
    assert (oracle.bfs_parents(hrp, hcol, 0) + 1).tolist() == [1, 1, 1, 2, 3]
    # dir=:in on g5 from vertex 4: 4 <- 3 <- {1,2}; parent of 1 and 2 is 3
    par_in = oracle.bfs_parents(irp, icol, 3) + 1
    assert par_in.tolist() == [3, 3, 4, 4]


def _bfs_tree_ok(rp, col, src, par, dist):
    """par is a BFS tree of the graph from src: the root is its own parent,
    every reached vertex's parent is an in-neighbour one level closer, and
    exactly the reached vertices have a parent."""
    reached = dist >= 0
    if par[src] != src or not ((par >= 0) == reached).all():
        return False
    for v in np.nonzero(reached)[0]:
        if v == src:
            continue
        u = par[v]
        if dist[u] != dist[v] - 1 or v not in col[rp[u]:rp[u + 1]]:
            return False
    return True


def test_threaded_parallel_bfs_reference_vectors(oracle):
    # test/parallel/traversals/gdistances.jl:7-17 and
    # test/parallel/traversals/bfs.jl:26, on 1, 3 and 8 threads
    _, _, (rp5, col5, _, _) = mk(G5)
    _, _, (rp, col, _, _) = mk(HOUSE)
    for nt in (1, 3, 8):
        for srcs in ([0], [0, 2]):
            assert (oracle.par_gdistances_threaded(rp5, col5, srcs, nt) == oracle.gdistances(rp5, col5, srcs)).all()
        for srcs in ([1], [0, 1], []):
            assert (oracle.par_gdistances_threaded(rp, col, srcs, nt) == oracle.gdistances(rp, col, srcs)).all()
        assert (oracle.par_bfs_tree_threaded(rp5, col5, 0, nt) + 1).tolist() == [1, 1, 1, 3]


@pytest.mark.parametrize("directed", [False, True])
def test_threaded_parallel_bfs_random(oracle, directed):
    """The threaded schedules against the serial restatements on graphs large
    enough for the threads to race: distances equal for every thread count
    and segment size; the threaded tree is a BFS tree."""
    rng = np.random.default_rng(7 + directed)
    n = 30000
    src = rng.integers(0, n, 240000)
    dst = (src + rng.geometric(0.002, src.size)) % n          # a long diameter
    hub = rng.integers(0, n, 4000)
    src = np.concatenate([src, np.zeros(4000, np.int64)])
    dst = np.concatenate([dst, hub])
    rp, col, _, _ = csr_of(n, src, dst, directed, drop_self_loops=True)
    for roots in ([0], [5, 17, 29999], list(rng.integers(0, n, 64))):
        ref = oracle.gdistances(rp, col, roots)
        for nt, seg in ((2, 20), (8, 1), (16, 20), (8, 1000)):
            assert (oracle.par_gdistances_threaded(rp, col, roots, nt, seg) == ref).all()
    for root in (0, 123):
        dist = oracle.gdistances(rp, col, [root])
        for nt in (1, 4, 16):
            par = oracle.par_bfs_tree_threaded(rp, col, root, nt)
            assert _bfs_tree_ok(rp, col, root, par, dist)
    # the validity check rejects a parent that is adjacent but on the wrong
    # level, and one on the right level that is not adjacent
    par = oracle.par_bfs_tree_threaded(rp, col, 0, 4)
    dist = oracle.gdistances(rp, col, [0])
    v = next(int(w) for w in np.nonzero(dist == 3)[0] if (dist[col[rp[w]:rp[w + 1]]] >= 3).any())
    wrong_level = next(int(w) for w in col[rp[v]:rp[v + 1]] if dist[w] >= 3)   # adjacent, level >= 3
    bad = par.copy(); bad[v] = wrong_level if not directed else int(par[par[v]])
    assert not _bfs_tree_ok(rp, col, 0, bad, dist)
    same_level = np.nonzero(dist == 2)[0]
    u = next(int(w) for w in same_level if v not in col[rp[w]:rp[w + 1]])
    bad = par.copy(); bad[v] = u
    assert not _bfs_tree_ok(rp, col, 0, bad, dist)


def test_dijkstra_allpaths_reference_vectors(oracle):
    # test/shortestpaths/dijkstra.jl:93-97
    n, _, (rp, col, irp, _) = mk(G615)
    d, pc, preds, closest = oracle.dijkstra_allpaths(rp, col, irp, 0)
    assert pc.tolist() == [1.0, 1.0, 1.0, 1.0, 2.0]
    assert [sorted(p) for p in preds] == [[], [0], [0], [2], [1, 2]]
    assert (closest + 1).tolist() == [1, 2, 3, 5, 4]
    # dijkstra.jl:100-106: unreachable vertices appended in id order
    _, _, (rp2, col2, irp2, _) = mk((5, False, [(1, 2), (1, 3), (4, 5)]))
    _, _, _, closest2 = oracle.dijkstra_allpaths(rp2, col2, irp2, 0)
    assert (closest2 + 1).tolist() == [1, 2, 3, 4, 5]


def test_betweenness_reference_fixture(oracle):
    z = load_golden("ref_graph_50_500.npz")
    n = int(z["n"])
    rp, col, irp, icol = csr_of(n, z["src"], z["dst"], True)
    vs = np.arange(n)
    bc = oracle.betweenness(rp, col, irp, True, vs, normalize=True, endpoints=False)
    # test/centrality/betweenness.jl:13 compares in Float32; hold FP64 to 1e-12
    assert (bc.astype(np.float32) == z["bc"].astype(np.float32)).all()
    assert np.max(np.abs(bc - z["bc"]) / np.abs(z["bc"])) < 1e-12
    # :15-17 endpoints=true, normalize=false, first three to 4 digits
    y = oracle.betweenness(rp, col, irp, True, vs, normalize=False, endpoints=True)
    ref = [122.10760591498584, 159.0072453120582, 176.39547945994505]
    assert np.round(y[:3], 4).tolist() == np.round(ref, 4).tolist()
    assert np.max(np.abs(y[:3] - ref) / np.abs(ref)) < 1e-12
    # threaded form over 1:n equals the serial one (test/parallel/centrality/betweenness.jl:10-19)
    zt = oracle.betweenness_threaded(rp, col, irp, True, vs, nthreads=4)
    assert np.allclose(zt, bc, rtol=1e-12, atol=0)


def test_betweenness_small_reference_cases(oracle):
    # test/centrality/betweenness.jl:20-21: s1 (undirected s2) and s2
    n, d, (rp, col, irp, _) = mk(S2)
    assert oracle.betweenness(rp, col, irp, True, np.arange(3)).tolist() == [0, 0.5, 0]
    n, d, (rp, col, irp, _) = mk((3, False, S2[2]))
    assert oracle.betweenness(rp, col, irp, False, np.arange(3)).tolist() == [0, 1, 0]
    # :23-29 K2 gives zeros
    n, d, (rp, col, irp, _) = mk((2, False, [(1, 2)]))
    assert oracle.betweenness(rp, col, irp, False, np.arange(2)).tolist() == [0.0, 0.0]
    # :32-33 path_graph(5), normalize=false: ends are 0
    n, d, (rp, col, irp, _) = mk(path_spec(5))
    z = oracle.betweenness(rp, col, irp, False, np.arange(5), normalize=False)
    assert z[0] == z[4] == 0.0
    assert z.tolist() == [0.0, 3.0, 4.0, 3.0, 0.0]        # textbook path BC
    # doctests betweenness.jl:32-43
    n, d, (rp, col, irp, _) = mk(star_spec(3))
    assert oracle.betweenness(rp, col, irp, False, np.arange(3)).tolist() == [1.0, 0.0, 0.0]
    n, d, (rp, col, irp, _) = mk(path_spec(4))
    assert oracle.betweenness(rp, col, irp, False, np.arange(4)).tolist() == [0.0, 2 / 3, 2 / 3, 0.0]


def test_betweenness_grid_bound(oracle):
    # test/centrality/betweenness.jl:75-77: grid([50, 50]), normalize=false
    import lgmi.generators as gen
    g = gen.grid([50, 50])
    rp, col = g.csr()
    z = oracle.betweenness(rp, col, rp, False, np.arange(2500), normalize=False)
    assert z.max() < 2500 * 2499


def test_threaded_index_quirk(oracle):
    # Parallel/centrality/betweenness.jl:53-56 uses findall positions as
    # sources: vs = [3, 3] on path_graph(3) processes sources 1 and 2.
    n, d, (rp, col, irp, _) = mk(path_spec(3))
    quirky = oracle.betweenness_threaded(rp, col, irp, False, [2, 2], normalize=False, nthreads=2)
    as_vertices = oracle.betweenness(rp, col, irp, False, [0, 1], normalize=False)
    # rescale uses k = length(vs) = 2 in both
    assert np.allclose(quirky, as_vertices)
    honest = oracle.betweenness(rp, col, irp, False, [2, 2], normalize=False)
    assert not np.allclose(quirky, honest)


# ----------------------------------------------------------------------
# (b) networkx golden vectors
# ----------------------------------------------------------------------
@pytest.mark.parametrize("name", GOLDEN_NX)
def test_oracle_vs_networkx(oracle, name):
    z = load_golden(name)
    n, directed = int(z["n"]), bool(z["directed"])
    rp, col, irp, icol = csr_of(n, z["src"], z["dst"], directed)
    for s, dref in zip(z["bfs_src"], z["dist"]):
        assert (oracle.gdistances(rp, col, [int(s)]) == dref).all()
        assert (oracle.par_gdistances(rp, col, [int(s)]) == dref).all()
    vs = np.arange(n)
    for key, kw in (("bc_norm", dict(normalize=True)), ("bc_raw", dict(normalize=False))):
        got = oracle.betweenness(rp, col, irp, directed, vs, **kw)
        ref = z[key]
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max()))
    got = oracle.betweenness(rp, col, irp, directed, z["sub_src"], normalize=True)
    assert np.allclose(got, z["bc_sub_norm"], rtol=1e-12, atol=1e-15)
    if "bc_end_raw" in z:
        got = oracle.betweenness(rp, col, irp, directed, vs, normalize=False, endpoints=True)
        assert np.allclose(got, z["bc_end_raw"], rtol=1e-12, atol=0)


# ----------------------------------------------------------------------
# (c) brute force on tiny graphs
# ----------------------------------------------------------------------
def _random_tiny(rng, n, p, directed):
    adj = rng.random((n, n)) < p
    np.fill_diagonal(adj, False)
    if not directed:
        adj = np.triu(adj)
        adj = adj | adj.T
    s, d = np.nonzero(adj)
    return s, d, adj


def _all_shortest_paths(adj, s, t):
    """Every shortest s->t path, by enumerating simple paths (independent of BFS)."""
    n = len(adj)
    best, paths = None, []
    stack = [(s, [s])]
    while stack:
        u, path = stack.pop()
        if best is not None and len(path) - 1 > best:
            continue
        if u == t:
            L = len(path) - 1
            if best is None or L < best:
                best, paths = L, [path]
            elif L == best:
                paths.append(path)
            continue
        for v in np.nonzero(adj[u])[0]:
            if v not in path:
                stack.append((int(v), path + [int(v)]))
    return paths


def _brute_bc(adj, directed):
    n = len(adj)
    raw = np.zeros(n)
    for s, t in itertools.permutations(range(n), 2):
        paths = _all_shortest_paths(adj, s, t)
        if not paths:
            continue
        for p in paths:
            for v in p[1:-1]:
                raw[v] += 1.0 / len(paths)
    return raw


@pytest.mark.parametrize("directed", [False, True])
def test_bc_and_parents_brute_force(oracle, directed):
    rng = np.random.default_rng(123 + directed)
    for trial in range(12):
        n = int(rng.integers(3, 8))
        s, d, adj = _random_tiny(rng, n, 0.35, directed)
        rp, col, irp, icol = csr_of(n, s, d, directed)
        raw = _brute_bc(adj, directed)
        got = oracle.betweenness(rp, col, irp, directed, np.arange(n), normalize=False)
        want = raw if directed else raw * 0.5
        # sources of degree 0 contribute nothing in both
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12), (trial, adj)
        got_n = oracle.betweenness(rp, col, irp, directed, np.arange(n), normalize=True)
        if n > 2:
            assert np.allclose(got_n, raw / ((n - 1) * (n - 2)), rtol=1e-12, atol=1e-12)
        # parents: min-id predecessor one level up (closed form)
        for src in range(n):
            dist = oracle.gdistances(rp, col, [src])
            par = oracle.bfs_parents(rp, col, src)
            for v in range(n):
                if v == src:
                    assert par[v] == src
                elif dist[v] < 0:
                    assert par[v] == -1
                else:
                    cands = [u for u in range(n) if adj[u, v] and dist[u] == dist[v] - 1]
                    assert par[v] == min(cands)


def test_endpoints_disconnected_closed_form(oracle):
    # _accumulate_endpoints! (betweenness.jl:109-134) counts every vertex,
    # reachable or not: on two disjoint edges each source adds nv-1 = 3 to
    # itself and 1 to each other vertex, so raw = 6 and the undirected
    # scale 1/2 gives 3.
    rp, col, irp, _ = csr_of(4, [0, 2], [1, 3], False)
    got = oracle.betweenness(rp, col, irp, False, np.arange(4), normalize=False, endpoints=True)
    assert got.tolist() == [3.0, 3.0, 3.0, 3.0]


# ----------------------------------------------------------------------
# planted mistakes: the checks above reject them
# ----------------------------------------------------------------------
def test_planted_mistakes_are_caught(oracle):
    z = load_golden("ref_graph_50_500.npz")
    n = int(z["n"])
    rp, col, irp, _ = csr_of(n, z["src"], z["dst"], True)
    bc = oracle.betweenness(rp, col, irp, True, np.arange(n))
    # wrong k in the rescale (k = n-1 instead of length(vs))
    wrong_k = oracle.rescale(oracle.betweenness(rp, col, irp, True, np.arange(n), normalize=False),
                             n, True, True, n - 1)
    assert not np.allclose(wrong_k, z["bc"], rtol=1e-6)
    # a digraph treated as undirected (reversing *every* edge would not be a
    # mistake: betweenness is invariant under edge reversal)
    srp, scol, _, _ = csr_of(n, z["src"], z["dst"], False)
    t = oracle.betweenness(srp, scol, srp, False, np.arange(n))
    assert not np.allclose(t, z["bc"], rtol=1e-6)
    # an edge-reversed graph must give the same values
    t = oracle.betweenness(irp, _transpose_col(rp, col), rp, True, np.arange(n))
    assert np.allclose(t, z["bc"], rtol=1e-12)
    assert np.allclose(bc, z["bc"], rtol=1e-12)
    # last-discoverer parents instead of first-discoverer
    _, _, (rp5, col5, _, _) = mk(G5)
    par = oracle.bfs_parents(rp5, col5, 0) + 1
    wrong = par.copy()
    wrong[2] = 2                      # 3 discovered from 2 instead of 1
    assert par.tolist() == [1, 1, 1, 3] and wrong.tolist() != [1, 1, 1, 3]


def _transpose_col(rp, col):
    from lgmi.graphs import csr_transpose
    return csr_transpose(rp, col)[1]


# ----------------------------------------------------------------------
# stress centrality (src/centrality/stress.jl)
# ----------------------------------------------------------------------

def stress_by_levels(n, directed, rp, col, irp, icol, vs):
    """The recursion of _stress_accumulate_basic! (stress.jl:53-77) written
    from BFS levels instead of the Dijkstra state: P[w] = in-neighbours one
    level closer, delta(w) = sum over x with w in P[x] of (delta(x) + 1),
    stress[w] += delta(w) * |P[w]| for w != s.  Python ints, no wrapping."""
    out = [0] * n
    deg = np.diff(rp) + (np.diff(irp) if directed else 0)
    for s in vs:
        if deg[s] == 0:
            continue
        d = [-1] * n
        d[s] = 0
        order, q = [s], [s]
        while q:
            nq = []
            for u in q:
                for v in col[rp[u]:rp[u + 1]]:
                    if d[v] < 0:
                        d[v] = d[u] + 1
                        nq.append(int(v))
            order += nq
            q = nq
        preds = [[u for u in icol[irp[w]:irp[w + 1]] if d[u] >= 0 and d[u] == d[w] - 1] if w != s else []
                 for w in range(n)]
        delta = [0] * n
        for w in reversed(order):
            for v in preds[w]:
                delta[v] += delta[w] + 1
            delta[w] *= len(preds[w])
            if w != s:
                out[w] += delta[w]
    return np.array(out, dtype=np.int64)


def test_stress_reference_fixture(oracle):
    """test/centrality/stress.jl:2-7 and test/parallel/centrality/stress.jl:2-10:
    the directed graph-50-500 fixture, exact."""
    z = load_golden("ref_graph_50_500.npz")
    n = int(z["n"])
    rp, col, irp, icol = csr_of(n, z["src"], z["dst"], True)
    sc = oracle.stress(rp, col, irp, True, np.arange(n))
    assert sc.tolist() == z["sc"].tolist()
    assert oracle.stress(rp, col, irp, True, np.arange(n), nthreads=4).tolist() == z["sc"].tolist()
    assert stress_by_levels(n, True, rp, col, irp, icol, range(n)).tolist() == z["sc"].tolist()


def test_stress_small_reference_cases(oracle):
    # test/centrality/stress.jl:17-24: cycle_graph(4) + vertex 5 + edge 4-5
    n, d, (rp, col, irp, _) = mk((5, False, [(1, 2), (2, 3), (3, 4), (1, 4), (4, 5)]))
    assert oracle.stress(rp, col, irp, False, np.arange(5)).tolist() == [4, 2, 4, 10, 0]
    # doctests stress.jl:18-30: star_graph(3) and cycle_graph(4)
    n, d, (rp, col, irp, _) = mk(star_spec(3))
    assert oracle.stress(rp, col, irp, False, np.arange(3)).tolist() == [2, 0, 0]
    n, d, (rp, col, irp, _) = mk((4, False, [(1, 2), (2, 3), (3, 4), (1, 4)]))
    assert oracle.stress(rp, col, irp, False, np.arange(4)).tolist() == [2, 2, 2, 2]
    # degree-0 sources are skipped; an empty source list gives zeros
    n, d, (rp, col, irp, _) = mk((4, False, [(1, 2), (2, 3)]))
    assert oracle.stress(rp, col, irp, False, [3]).tolist() == [0, 0, 0, 0]
    assert oracle.stress(rp, col, irp, False, []).tolist() == [0, 0, 0, 0]


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("directed", [False, True])
def test_stress_matches_level_restatement(oracle, seed, directed):
    """The C restatement (Dijkstra state, closest-vertices order) against the
    level-based restatement above, with duplicate and degree-0 sources."""
    rng = np.random.default_rng(seed)
    n = 60
    m = 150
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    rp, col, irp, icol = csr_of(n, src, dst, directed, True)
    vs = np.concatenate([rng.integers(0, n, 25), [0, 0]])
    got = oracle.stress(rp, col, irp, directed, vs)
    want = stress_by_levels(n, directed, rp, col, irp, icol, vs)
    assert got.tolist() == want.tolist()


# ----------------------------------------------------------------------
# closeness / radiality / eccentricity (src/centrality/{closeness,
# radiality}.jl, src/distance.jl) and the host formulas of lgmi.distances
# ----------------------------------------------------------------------

def stats_from_oracle(oracle, rp, col, vs):
    """(sum of finite distances, reached, largest finite distance) per source."""
    rows = []
    for s in vs:
        d = oracle.gdistances(rp, col, [int(s)]).astype(np.int64)
        r = d[d >= 0]
        rows.append((int(r.sum()), len(r), int(r.max())))
    return np.array(rows, dtype=np.int64).reshape(-1, 3)


def test_radiality_reference_fixture(oracle):
    """test/centrality/radiality.jl:2-8 (exact ==) on the directed
    graph-50-500 fixture; the host formula gives the same bits."""
    from lgmi import distances as D
    z = load_golden("ref_graph_50_500.npz")
    n = int(z["n"])
    rp, col, irp, icol = csr_of(n, z["src"], z["dst"], True)
    r = oracle.radiality(rp, col)
    assert r.tolist() == z["rc"].tolist()
    st = stats_from_oracle(oracle, rp, col, range(n))
    assert D.radiality_from_stats(st, n).tolist() == z["rc"].tolist()


def test_radiality_small_reference_cases(oracle):
    # radiality.jl test :10-16 (cycle_graph(4) + pendant) and doctests :18-31
    n, d, (rp, col, irp, _) = mk((5, False, [(1, 2), (2, 3), (3, 4), (1, 4), (4, 5)]))
    assert np.allclose(oracle.radiality(rp, col), [5 / 6, 3 / 4, 5 / 6, 11 / 12, 2 / 3], rtol=1e-15)
    n, d, (rp, col, irp, _) = mk(star_spec(4))
    assert oracle.radiality(rp, col).tolist() == [1.0, 0.6666666666666666, 0.6666666666666666, 0.6666666666666666]
    n, d, (rp, col, irp, _) = mk(path_spec(3))
    assert oracle.radiality(rp, col).tolist() == [0.75, 1.0, 0.75]


def test_closeness_reference_cases(oracle):
    from lgmi import distances as D
    # test/centrality/closeness.jl:2-11: digraph 1->2, 2->3, 1->3, 3->4
    n, d, (rp, col, irp, _) = mk((4, True, [(1, 2), (2, 3), (1, 3), (3, 4)]))
    assert oracle.closeness(rp, col, irp, True, normalize=False).tolist() == [0.75, 0.6666666666666666, 1.0, 0.0]
    assert oracle.closeness(rp, col, irp, True).tolist() == [0.75, 0.4444444444444444, 0.3333333333333333, 0.0]
    st = stats_from_oracle(oracle, rp, col, range(4))
    assert D.closeness_from_stats(st, 4, np.arange(4), True).tolist() == [0.75, 0.4444444444444444,
                                                                          0.3333333333333333, 0.0]
    # :24-30: SimpleGraph(5) with edge 1-2
    n, d, (rp, col, irp, _) = mk((5, False, [(1, 2)]))
    assert oracle.closeness(rp, col, irp, False).tolist() == [0.25, 0.25, 0.0, 0.0, 0.0]
    # doctests closeness.jl:14-28
    n, d, (rp, col, irp, _) = mk(star_spec(5))
    assert oracle.closeness(rp, col, irp, False).tolist() == [1.0] + [0.5714285714285714] * 4
    n, d, (rp, col, irp, _) = mk(path_spec(4))
    assert oracle.closeness(rp, col, irp, False).tolist() == [0.5, 0.75, 0.75, 0.5]


def test_eccentricity_reference_cases(oracle):
    from lgmi import distances as D
    # doctests src/distance.jl:47-55, 92-99, 113-122, 140-147, 159-168
    n, d, (rp, col, irp, _) = mk(path_spec(3))
    assert oracle.eccentricity(rp, col, [0]).tolist() == [2]
    assert oracle.eccentricity(rp, col, [0, 1]).tolist() == [2, 1]
    n, d, (rp, col, irp, _) = mk(star_spec(5))
    e = oracle.eccentricity(rp, col, range(5))
    assert (D.diameter(e), D.radius(e)) == (2, 1)
    assert D.periphery(e).tolist() == [2, 3, 4, 5] and D.center(e).tolist() == [1]
    n, d, (rp, col, irp, _) = mk(path_spec(5))
    e = oracle.eccentricity(rp, col, range(5))
    assert (D.diameter(e), D.radius(e)) == (4, 2)
    assert D.periphery(e).tolist() == [1, 5] and D.center(e).tolist() == [3]
    # test/distance.jl:59-64: two isolated vertices -> typemax(Int)
    n, d, (rp, col, irp, _) = mk((2, False, []))
    assert oracle.eccentricity(rp, col, [0, 1]).tolist() == [D.TYPEMAX_INT] * 2
    st = stats_from_oracle(oracle, rp, col, [0, 1])
    assert D.ecc_from_stats(st, 2).tolist() == [D.TYPEMAX_INT] * 2


@pytest.mark.parametrize("seed", [4, 5])
@pytest.mark.parametrize("directed", [False, True])
def test_distance_host_formulas_vs_oracle(oracle, seed, directed):
    """The host formulas of lgmi.distances (from the three per-source
    integers) against the oracle's direct restatements (from the whole
    distance vectors) on disconnected random graphs: bit-identical."""
    from lgmi import distances as D
    rng = np.random.default_rng(seed)
    n = 80
    src = rng.integers(0, n, 120)
    dst = rng.integers(0, n, 120)
    rp, col, irp, icol = csr_of(n, src, dst, directed, True)
    st = stats_from_oracle(oracle, rp, col, range(n))
    rad = oracle.radiality(rp, col)
    assert np.array_equal(D.radiality_from_stats(st, n), rad, equal_nan=True)
    deg = np.diff(rp) + (np.diff(irp) if directed else 0)
    active = np.nonzero(deg > 0)[0]
    for normalize in (True, False):
        want = oracle.closeness(rp, col, irp, directed, normalize=normalize)
        assert D.closeness_from_stats(st[active], n, active, normalize).tolist() == want.tolist()
    assert D.ecc_from_stats(st, n).tolist() == oracle.eccentricity(rp, col, range(n)).tolist()


# ----------------------------------------------------------------------
# connectivity: connected_components, has_path, period
# ----------------------------------------------------------------------

def test_connected_components_reference_case(oracle):
    """test/connectivity.jl:2-22 (path_graph(4) + 10 vertices + edges)."""
    from lgmi import connectivity as CC
    edges = [(1, 2), (2, 3), (3, 4), (5, 6), (6, 7), (8, 9), (10, 9)]
    n, d, (rp, col, irp, icol) = mk((14, False, edges))
    lab = oracle.connected_component_labels(rp, col, irp, icol, False) + 1
    assert lab[:10].tolist() == [1, 1, 1, 1, 5, 5, 5, 8, 8, 8]
    cd = CC.components_dict(lab)
    assert cd[1] == [1, 2, 3, 4] and cd[5] == [5, 6, 7] and cd[8] == [8, 9, 10]
    c, _ = CC.components(lab)
    assert len(c) >= 3 and sorted(c[2]) == [8, 9, 10]


@pytest.mark.parametrize("directed", [False, True])
def test_connected_components_vs_scipy(oracle, directed):
    """Labels against scipy's (weak) components mapped to each component's
    smallest vertex id."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components as sp_cc
    rng = np.random.default_rng(41)
    n = 500
    src = rng.integers(0, n, 420)
    dst = rng.integers(0, n, 420)
    rp, col, irp, icol = csr_of(n, src, dst, directed, True)
    lab = oracle.connected_component_labels(rp, col, irp, icol, directed)
    m = csr_matrix((np.ones(len(col)), np.asarray(col), np.asarray(rp)), shape=(n, n))
    _, sl = sp_cc(m, directed=directed, connection="weak")
    mins = {}
    for v, l in enumerate(sl.tolist()):
        mins.setdefault(l, v)
    assert lab.tolist() == [mins[l] for l in sl.tolist()]


def test_has_path_reference_cases(oracle):
    """test/traversals/bfs.jl:48-72."""
    n, d, (rp, col, irp, _) = mk((6, False, [(1, 2), (2, 3), (2, 4), (4, 5), (3, 5)]))
    hp = lambda u, v, ex=(): oracle.has_path(rp, col, u - 1, v - 1, [x - 1 for x in ex])  # noqa: E731
    assert hp(1, 5) and hp(1, 2) and hp(1, 5, [3]) and hp(1, 5, [4]) and not hp(1, 5, [3, 4])
    assert hp(5, 1) and hp(5, 1, [3]) and hp(5, 1, [4]) and not hp(5, 1, [3, 4])
    assert not hp(1, 6) and not hp(6, 1) and hp(1, 1)
    assert not hp(1, 2, [2]) and not hp(1, 2, [1])


def test_period_reference_cases(oracle):
    """test/connectivity.jl:206-222: cycle_digraph(10) has period 10, with the
    shortcut 1 -> 4 period 2; a graph that is not strongly connected raises."""
    ring = [(i, i % 10 + 1) for i in range(1, 11)]
    n, d, (rp, col, irp, icol) = mk((10, True, ring))
    assert oracle.period(rp, col, irp, icol) == 10
    n, d, (rp, col, irp, icol) = mk((10, True, ring + [(1, 4)]))
    assert oracle.period(rp, col, irp, icol) == 2
    n, d, (rp, col, irp, icol) = mk((3, True, [(1, 2), (2, 3)]))
    with pytest.raises(ValueError):
        oracle.period(rp, col, irp, icol)


@pytest.mark.parametrize("seed", [42, 43, 44])
def test_period_root_invariance(oracle, seed):
    """The period of a strongly connected digraph is the gcd over all edges
    of level[s] - level[d] + 1 for levels from any root (the tree edges give
    0): the restatement from vertex 1 equals the all-edge gcd from the last
    vertex, on cycles of several lengths joined by chords."""
    import math
    rng = np.random.default_rng(seed)
    k = int(rng.integers(2, 6))
    n = k * int(rng.integers(3, 12))
    ring = [(i, (i + 1) % n) for i in range(n)]
    # chords that keep the period a multiple of k
    chords = [(int(a), int((a + k * rng.integers(1, n // k) + 1) % n)) for a in rng.integers(0, n, 5)]
    e = np.array(ring + chords)
    rp, col, irp, icol = csr_of(n, e[:, 0], e[:, 1], True, True)
    p1 = oracle.period(rp, col, irp, icol)
    d = oracle.gdistances(rp, col, [n - 1]).astype(np.int64)
    src = np.repeat(np.arange(n), np.diff(rp))
    g = 0
    for s_, d_ in zip(src.tolist(), np.asarray(col).tolist()):
        g = math.gcd(g, int(d[s_] - d[d_] + 1))
    assert p1 == g and p1 % k == 0


def test_neighborhood_reference_cases(oracle):
    """Doctests src/connectivity.jl:607-632, 646-676 and test/connectivity.jl:225-253."""
    e = [(1, 2), (2, 3), (3, 1), (3, 4), (4, 5), (5, 4)]
    n, d, (rp, col, irp, icol) = mk((5, True, e))
    nd = lambda v, dd, r=rp, c=col: [(x + 1, y) for x, y in oracle.neighborhood_dists(r, c, v - 1, dd)]  # noqa: E731
    assert nd(1, 3) == [(1, 0), (2, 1), (3, 2), (4, 3)]
    assert [x for x, _ in nd(1, 2)] == [1, 2, 3]
    assert nd(4, 3) == [(4, 0), (5, 1)]
    assert nd(4, 3, irp, icol) == [(4, 0), (3, 1), (5, 1), (2, 2), (1, 3)]
    n, d, (rp, col, irp, icol) = mk(star_spec(10))
    assert nd(1, 0, rp, col) == [(1, 0)]
    assert [len(nd(v, dd, rp, col)) for v, dd in [(1, 1), (2, 1), (1, 2), (2, 2), (2, -1)]] == [10, 2, 10, 10, 0]
    n, d, (rp, col, irp, icol) = mk((10, True, [(1, i) for i in range(2, 11)]))
    assert [len(nd(v, dd, rp, col)) for v, dd in [(1, 1), (2, 1), (1, 2), (2, 2)]] == [10, 1, 10, 1]
    assert [len(nd(v, dd, irp, icol)) for v, dd in [(1, 1), (2, 1), (1, 2), (2, 2)]] == [1, 2, 1, 2]


def test_neighborhood_order_vs_networkx(oracle):
    """The queue order restated from networkx's BFS edge order (its
    bfs_edges visits sources FIFO and neighbours in insertion order, which is
    sorted here), cut at depth d."""
    import networkx as nx
    rng = np.random.default_rng(61)
    n = 300
    src = rng.integers(0, n, 700)
    dst = rng.integers(0, n, 700)
    rp, col, irp, icol = csr_of(n, src, dst, True, True)
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    for u in range(n):
        for v in col[rp[u]:rp[u + 1]]:
            G.add_edge(u, int(v))
    for v in (0, 5, 17):
        for dd in (0, 1, 2, 4, 100):
            want = [(v, 0)]
            depth = {v: 0}
            for a, b in nx.bfs_edges(G, v):
                depth[b] = depth[a] + 1
                if depth[b] <= dd:
                    want.append((b, depth[b]))
            assert oracle.neighborhood_dists(rp, col, v, dd) == want


def _brandes_py(n, rp, col, round_sigma_f32=False):
    """Textbook Brandes (all sources, undirected, normalize=false halving),
    optionally rounding each predecessor's path count to FP32 where it is
    added -- the mistake a narrowed path-count store would make."""
    from collections import deque
    bc = [0.0] * n
    for s in range(n):
        d = [-1] * n
        sig = [0.0] * n
        d[s], sig[s] = 0, 1.0
        order, q = [], deque([s])
        while q:
            v = q.popleft()
            order.append(v)
            for w in col[rp[v]:rp[v + 1]]:
                if d[w] < 0:
                    d[w] = d[v] + 1
                    q.append(w)
                if d[w] == d[v] + 1:
                    sv = float(np.float32(sig[v])) if round_sigma_f32 else sig[v]
                    sig[w] += sv
        delta = [0.0] * n
        for w in reversed(order):
            for v in col[rp[w]:rp[w + 1]]:
                if d[v] == d[w] - 1:
                    delta[v] += sig[v] / sig[w] * (1.0 + delta[w])
            if w != s:
                bc[w] += delta[w]
    return np.array(bc) * 0.5


def test_oracle_betweenness_on_layered_counts_beyond_fp32(oracle):
    """The layered graph's path counts pass 2^24: the oracle agrees with a
    textbook Brandes at 1e-12, and the same Brandes with the counts rounded
    to FP32 where they are summed misses by far more than that bar, so a GPU
    test on this graph catches a narrowed path count."""
    from conftest import layered_edges
    n, s, d = layered_edges()
    rp, col, _, _ = csr_of(n, s, d, False)
    rp_l, col_l = rp.tolist(), col.tolist()
    exact = _brandes_py(n, rp_l, col_l)
    got = oracle.betweenness(rp, col, rp, False, np.arange(n), normalize=False)
    assert np.allclose(got, exact, rtol=1e-12, atol=0)
    narrowed = _brandes_py(n, rp_l, col_l, round_sigma_f32=True)
    rel = np.abs(narrowed - exact) / np.maximum(np.abs(exact), 1e-300)
    assert rel.max() > 1e-9, rel.max()
