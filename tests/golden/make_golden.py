"""Writes the golden fixtures under tests/golden/ (run where the reference
tree and networkx are available; the GPU test runs only read the committed
.npz files).

* ref_graph_50_500.npz -- the reference's own fixture: the directed graph of
  test/testdata/graph-50-500.jgz (LG text format, gzip), the FP64
  betweenness of test/testdata/graph-50-500-bc.txt (normalize=true,
  endpoints=false), as used by test/centrality/betweenness.jl:8-13, and the
  Int64 stress centrality of test/testdata/graph-50-500-sc.txt
  (test/centrality/stress.jl:2-7) and the FP64 radiality of
  test/testdata/graph-50-500-rc.txt (test/centrality/radiality.jl:2-8).
* nx_*.npz -- networkx 3.4.2 vectors for graphs drawn with this project's
  generators: BFS distances (single_source_shortest_path_length) and
  betweenness for the full vertex set and for explicit source lists.  For
  source lists the per-source Brandes sums come from networkx's
  _single_source_shortest_path_basic / _accumulate_basic(_endpoints) and
  the scale is the reference's _rescale! (src/centrality/betweenness.jl:
  136-160), whose conventions networkx's _rescale shares for these cases.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import gzip
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, os.path.join(REPO, "lightgraphs.jl_amd"))


def read_lg(path):
    """LG format (src/persistence/lg.jl:1-52): header 'nv,ne,d|u,name',
    then one 'u, v' line per edge, 1-based."""
    with gzip.open(path, "rt") as f:
        header = f.readline().strip().split(",")
        n, ne, kind = int(header[0]), int(header[1]), header[2].strip()
        edges = []
        for line in f:
            line = line.strip()
            if not line:
                continue
            a, b = line.split(",")
            edges.append((int(a), int(b)))
    assert len(edges) == ne
    e = np.array(edges, dtype=np.int64)
    return n, kind == "d", e[:, 0] - 1, e[:, 1] - 1


def nx_graph(n, directed, src, dst):
    import networkx as nx
    G = nx.DiGraph() if directed else nx.Graph()
    G.add_nodes_from(range(n))
    G.add_edges_from(zip(src.tolist(), dst.tolist()))
    return G


def ref_rescale(bc, n, normalize, directed, k):
    """_rescale! -- src/centrality/betweenness.jl:136-160."""
    bc = bc.copy()
    do_scale = False
    scale = 0.0
    if normalize:
        if n > 2:
            do_scale = True
            scale = 1.0 / ((n - 1) * (n - 2))
    elif not directed:
        do_scale = True
        scale = 1.0 / 2.0
    if do_scale:
        if k > 0:
            scale = scale * n / k
        bc *= scale
    return bc


def nx_bc_sources(G, n, directed, sources, normalize, endpoints):
    """Brandes sums over an explicit source list (duplicates counted,
    degree-0 sources skipped as betweenness.jl:58 does)."""
    from networkx.algorithms.centrality.betweenness import (
        _accumulate_basic, _accumulate_endpoints, _single_source_shortest_path_basic)
    bc = dict.fromkeys(G, 0.0)
    for s in sources:
        deg = G.in_degree(s) + G.out_degree(s) if directed else G.degree(s)
        if deg == 0:
            continue
        S, P, sigma, _ = _single_source_shortest_path_basic(G, s)
        if endpoints:
            bc, _ = _accumulate_endpoints(bc, S, P, sigma, s)
        else:
            bc, _ = _accumulate_basic(bc, S, P, sigma, s)
    arr = np.array([bc[v] for v in range(n)], dtype=np.float64)
    return ref_rescale(arr, n, normalize, directed, len(sources))


def write_ref_fixture():
    n, directed, src, dst = read_lg(os.path.join(REF, "test/testdata/graph-50-500.jgz"))
    bc = np.loadtxt(os.path.join(REF, "test/testdata/graph-50-500-bc.txt"), delimiter=",", dtype=np.float64)
    sc = np.loadtxt(os.path.join(REF, "test/testdata/graph-50-500-sc.txt"), delimiter=",", dtype=np.int64)
    rc = np.loadtxt(os.path.join(REF, "test/testdata/graph-50-500-rc.txt"), delimiter=",", dtype=np.float64)
    np.savez(os.path.join(HERE, "ref_graph_50_500.npz"), n=n, directed=directed, src=src, dst=dst, bc=bc, sc=sc,
             rc=rc)
    print("ref_graph_50_500.npz", n, len(src), directed)


def write_nx_fixture(name, n, directed, src, dst, seed):
    import networkx as nx
    G = nx_graph(n, directed, src, dst)
    rng = np.random.default_rng(seed)
    bfs_src = rng.choice(n, size=min(4, n), replace=False)
    dists = []
    for s in bfs_src:
        d = nx.single_source_shortest_path_length(G, int(s))
        arr = np.full(n, -1, dtype=np.int32)
        for v, x in d.items():
            arr[v] = x
        dists.append(arr)
    all_src = np.arange(n)
    sub = rng.choice(n, size=min(37, n), replace=True)        # duplicates allowed
    out = dict(n=n, directed=directed, src=np.asarray(src, np.int64), dst=np.asarray(dst, np.int64),
               bfs_src=bfs_src.astype(np.int64), dist=np.array(dists), sub_src=sub.astype(np.int64))
    out["bc_norm"] = nx_bc_sources(G, n, directed, all_src, True, False)
    out["bc_raw"] = nx_bc_sources(G, n, directed, all_src, False, False)
    out["bc_sub_norm"] = nx_bc_sources(G, n, directed, sub, True, False)
    # networkx's endpoint accumulation equals the reference's only when every
    # vertex is reachable from every source (betweenness.jl:122,130 count
    # unreachable vertices too); keep it for strongly connected graphs.
    if nx.is_strongly_connected(G) if directed else nx.is_connected(G):
        out["bc_end_raw"] = nx_bc_sources(G, n, directed, all_src, False, True)
    # cross-check against networkx's own public entry point
    pub = nx.betweenness_centrality(G, normalized=True, endpoints=False)
    pub = np.array([pub[v] for v in range(n)])
    if not directed:
        pass  # same convention as the reference: 1/((n-1)(n-2)) with no halving
    assert np.allclose(pub, out["bc_norm"], rtol=1e-12, atol=0), name
    np.savez(os.path.join(HERE, f"nx_{name}.npz"), **out)
    print(f"nx_{name}.npz", n, len(src), directed)


def main():
    import lgmi.generators as gen
    from lgmi.graphs import csr_from_edges

    write_ref_fixture()

    def und_edges(g):
        rp, col = g.csr()
        s = np.repeat(np.arange(g.nv()), np.diff(rp))
        keep = s <= col
        return s[keep], col[keep]

    g = gen.erdos_renyi(300, 1200, seed=7)
    write_nx_fixture("er_300_1200", 300, False, *und_edges(g), seed=1)
    g = gen.barabasi_albert(400, 4, seed=11)
    write_nx_fixture("ba_400_4", 400, False, *und_edges(g), seed=2)
    g = gen.watts_strogatz(400, 8, 0.1, seed=13)
    write_nx_fixture("ws_400_8", 400, False, *und_edges(g), seed=3)
    g = gen.rmat(9, 8, seed=5)
    write_nx_fixture("rmat9_sym", 512, False, *und_edges(g), seed=4)
    # directed: kronecker digraph (self-loops kept) and a sparse random digraph
    s, d = gen.kronecker_edges(8, 6, seed=9)
    rp, col = csr_from_edges(256, s, d, symmetrize=False, drop_self_loops=False)
    src = np.repeat(np.arange(256), np.diff(rp))
    write_nx_fixture("rmat8_dir", 256, True, src, col, seed=5)
    rng = np.random.default_rng(17)
    s = rng.integers(0, 200, 700)
    d = rng.integers(0, 200, 700)
    rp, col = csr_from_edges(200, s, d, symmetrize=False, drop_self_loops=True)
    src = np.repeat(np.arange(200), np.diff(rp))
    write_nx_fixture("rand_dir_200", 200, True, src, col, seed=6)


if __name__ == "__main__":
    main()
