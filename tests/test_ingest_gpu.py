"""Device-side graph ingest (lgmi_graph_from_edges) against the host CSR
builder and the reference constructors' documented results.  Bar: exact."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_constructor_doctests():
    """simplegraph.jl:198-204 ({5, 2} undirected) and simpledigraph.jl:205-210
    ({5, 3} directed)."""
    import lgmi
    g, ne = lgmi.SimpleGraph.from_edge_list([1, 1], [2, 5])
    assert (g.nv(), ne) == (5, 2)
    assert g.csr()[0].tolist() == [0, 2, 3, 3, 3, 4] and g.csr()[1].tolist() == [1, 4, 0, 0]
    g, ne = lgmi.SimpleDiGraph.from_edge_list([1, 1, 3], [3, 5, 1])
    assert (g.nv(), ne) == (5, 3)
    assert g.csr()[1].tolist() == [2, 4, 0] and g.in_csr()[1].tolist() == [2, 0, 0]


@pytest.mark.parametrize("directed", [False, True])
def test_random_edge_lists_vs_host(directed):
    """Duplicates, self-loops, ids below 1 (skipped) and unsorted input."""
    import lgmi
    from lgmi.graphs import csr_from_edges, csr_transpose
    rng = np.random.default_rng(71)
    for m, n in [(0, 1), (1, 1), (50, 10), (5000, 800), (200000, 30000)]:
        src = rng.integers(-2, n, m) + 1       # 1-based, some ids below 1
        dst = rng.integers(-2, n, m) + 1
        if m:
            src[: m // 10] = dst[: m // 10]    # self-loops
            src = np.concatenate([src, src[: m // 5]])
            dst = np.concatenate([dst, dst[: m // 5]])   # duplicates
        say(f"ingest directed={directed} m={m}")
        cls = lgmi.SimpleDiGraph if directed else lgmi.SimpleGraph
        g, ne = cls.from_edge_list(src, dst)
        ok = (src >= 1) & (dst >= 1)
        s0, d0 = src[ok] - 1, dst[ok] - 1
        nv = int(max(s0.max(initial=-1), d0.max(initial=-1)) + 1)
        rp, col = csr_from_edges(nv, s0, d0, symmetrize=not directed)
        assert g.nv() == nv
        assert g.csr()[0].tolist() == rp.tolist() and g.csr()[1].tolist() == col.tolist()
        want = cls.from_csr(rp, col)
        assert ne == want.ne()
        if directed:
            irp, icol = csr_transpose(rp, col)
            assert g.in_csr()[0].tolist() == irp.tolist() and g.in_csr()[1].tolist() == icol.tolist()


def test_rmat_edge_list_and_bfs(oracle):
    """The shuffled edge list of RMAT-18 rebuilt on the device gives the same
    CSR, and BFS on the device-built handle matches the oracle."""
    import lgmi
    ref = lgmi.generators.rmat(18, 16, seed=72)
    rp, col = ref.csr()
    src = np.repeat(np.arange(ref.nv()), np.diff(rp))
    keep = src <= col                       # one direction of each undirected edge
    s, d = src[keep], col[keep]
    perm = np.random.default_rng(73).permutation(len(s))
    say("ingest rmat18")
    g, ne = lgmi.SimpleGraph.from_edge_list(s[perm] + 1, d[perm] + 1, nv_min=ref.nv())
    assert g.nv() == ref.nv() and ne == ref.ne()
    assert g.csr()[0].tolist() == rp.tolist() and (g.csr()[1] == col).all()
    root = int(np.argmax(np.diff(rp)))
    assert (lgmi.gdistances(g, root + 1) == lgmi.gdistances(ref, root + 1)).all()
    dref = oracle.gdistances(rp, col, [root])
    got = lgmi.gdistances(g, root + 1)
    assert (np.where(dref < 0, np.int64(lgmi.typemax(np.int64)), dref.astype(np.int64)) == got).all()


@pytest.mark.parametrize("directed", [False, True])
def test_wide_keys_and_skewed_digits(directed):
    """Keys of 50 bits (nv above 2^24: seven 8-bit passes of the radix
    sort) and a skewed edge list -- ten source vertices, long runs of equal
    keys and equal digits, a length that fills no tile evenly -- against
    the host builder."""
    import lgmi
    from lgmi.graphs import csr_from_edges, csr_transpose
    rng = np.random.default_rng(74)
    n = (1 << 24) + 3
    m = 123_457
    src = rng.integers(0, 10, m)
    dst = np.where(rng.random(m) < 0.5, rng.integers(0, 64, m), rng.integers(0, n, m))
    dst[:1000] = n - 1                      # the largest id, repeated
    say(f"ingest wide keys directed={directed}")
    cls = lgmi.SimpleDiGraph if directed else lgmi.SimpleGraph
    g, ne = cls.from_edge_list(src + 1, dst + 1, nv_min=n)
    rp, col = csr_from_edges(n, src, dst, symmetrize=not directed)
    assert g.nv() == n
    assert np.array_equal(g.csr()[0], rp) and np.array_equal(g.csr()[1], col)
    if directed:
        irp, icol = csr_transpose(rp, col)
        assert np.array_equal(g.in_csr()[0], irp) and np.array_equal(g.in_csr()[1], icol)
