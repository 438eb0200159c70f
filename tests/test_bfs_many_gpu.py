"""lgmi_bfs_run_many: the loop over roots of gdistances / bfs_parents
(src/traversals/bfs.jl:35-66, 95-141), several searches at a time on
separate streams and search states, against the single-search entry point
and the oracle, bit for bit.

Covers: more roots than concurrent searches and fewer (one root, none),
out- and in-edges of a digraph, a disconnected graph with isolated roots,
deep searches (a grid; captured level sequences are not used for these
searches), distances only, the learned direction policy, a per-call
search and lgmi_bfs_fetch after a batch, and a root out of range.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _single(g, r, dir_in):
    from lgmi import traversals
    d = traversals._device_dist(g, np.array([r], dtype=np.int32)) if dir_in == 0 else None
    p = traversals._parents0(g, r + 1, dir_in)
    return d, p


def _check(oracle, g, roots, dir_in=0, n_oracle=3):
    from lgmi import traversals
    dist, par = traversals._device_many(g, roots, dir_in, True)
    assert dist.shape == (len(roots), g.nv()) and par.shape == dist.shape
    rp, col = g.csr() if dir_in == 0 else g.in_csr()
    for i, r in enumerate(roots.tolist()):
        d1, p1 = _single(g, r, dir_in)
        assert np.array_equal(par[i], p1), f"parents root {r}"
        if d1 is not None:
            assert np.array_equal(dist[i], d1), f"dist root {r}"
        if i < n_oracle:
            d_or, p_or, _ = oracle.bfs_dist_parent(rp, col, r)
            assert np.array_equal(dist[i], d_or) and np.array_equal(par[i], p_or), f"oracle root {r}"
    return dist, par


def test_many_rmat_undirected(oracle):
    import lgmi
    g = lgmi.generators.rmat(15, 16, seed=31)
    deg = np.diff(g.csr()[0])
    rng = np.random.default_rng(32)
    roots = rng.choice(np.nonzero(deg > 0)[0], 37, replace=False).astype(np.int32)
    say("rmat15, 37 roots")
    _check(oracle, g, roots)
    # distances only, and a repeat gives the same bits
    from lgmi import traversals
    d2, p2 = traversals._device_many(g, roots, 0, False)
    assert p2 is None
    d3, _ = traversals._device_many(g, roots, 0, True)
    assert np.array_equal(d2, d3)
    # in calls of 5 roots each (the device arrays reused between calls)
    d4, p4 = traversals._device_many(g, roots, 0, True, chunk=5)
    assert np.array_equal(d4, d3) and np.array_equal(p4, traversals._device_many(g, roots, 0, True)[1])


def test_many_directed_both_directions(oracle):
    import lgmi
    g = lgmi.generators.rmat(13, 8, seed=33, directed=True)
    rng = np.random.default_rng(34)
    roots = rng.choice(g.nv(), 11, replace=False).astype(np.int32)
    for dir_in in (0, 1):
        say(f"directed rmat13, dir_in={dir_in}")
        _check(oracle, g, roots, dir_in)


def test_many_disconnected_and_isolated(oracle):
    import lgmi
    # two components plus isolated vertices; roots include isolated ones
    n = 3000
    rng = np.random.default_rng(35)
    a = rng.integers(0, 1000, 4000)
    b = rng.integers(0, 1000, 4000)
    c = rng.integers(1500, 2500, 3000)
    d = rng.integers(1500, 2500, 3000)
    g = lgmi.SimpleGraph.from_edges(n, np.concatenate([a, c]), np.concatenate([b, d]))
    roots = np.array([0, 2999, 1600, 1200, 5, 2400], dtype=np.int32)
    say("disconnected graph, 6 roots")
    dist, par = _check(oracle, g, roots, n_oracle=6)
    assert (dist[1] == -1).sum() == n - 1 and par[1][2999] == 2999


def test_many_deep_grid(oracle):
    import lgmi
    g = lgmi.generators.grid([160, 160])
    roots = np.array([0, 160 * 160 - 1, 80 * 160 + 80, 159], dtype=np.int32)
    say("grid 160x160, 4 roots (deep searches)")
    _check(oracle, g, roots, n_oracle=4)


def test_many_one_and_none(oracle):
    import lgmi
    from lgmi import _lib, traversals
    g = lgmi.generators.barabasi_albert(5000, 3, seed=36)
    say("one root")
    _check(oracle, g, np.array([17], dtype=np.int32))
    lib = _lib.load()
    _lib.check(lib.lgmi_bfs_run_many(g.handle(), None, 0, 0, None, None))
    say("root out of range")
    with pytest.raises(IndexError, match="out of range"):     # LGMI_E_BOUNDS
        traversals._device_many(g, np.array([3, 5000], dtype=np.int32), 0, True)
    with pytest.raises(IndexError):
        traversals.bfs_many(g, [1, 0])


def test_many_learned_policy_and_per_call_after(oracle):
    """With the learned direction policy the rows are the same; a per-call
    search and its fetch are unaffected by a batch that ran in between."""
    import lgmi
    from lgmi import _lib, policy, traversals
    g = lgmi.generators.rmat(14, 16, seed=37)
    rp, col = g.csr()
    deg = np.diff(rp)
    roots = np.random.default_rng(38).choice(np.nonzero(deg > 0)[0], 9, replace=False).astype(np.int32)
    base, bpar = traversals._device_many(g, roots, 0, True)
    params, _ = policy.load_weights()
    if params is None:
        params = policy.init_params(np.random.default_rng(1).normal(0, 1, (64, 16)).astype(np.float32), seed=1)
    try:
        policy.set_learned(params)
        say("learned policy, 9 roots")
        d, p = traversals._device_many(g, roots, 0, True)
        assert np.array_equal(d, base) and np.array_equal(p, bpar)
    finally:
        policy.set_heuristic()
    lib = _lib.load()
    h = g.handle()
    r0 = int(roots[0])
    _lib.check(lib.lgmi_bfs_run(h, r0, 0, 1, None))
    traversals._device_many(g, roots[1:4], 0, True)
    dist = np.empty(g.nv(), np.int32)
    par = np.empty(g.nv(), np.int32)
    _lib.check(lib.lgmi_bfs_fetch(h, _lib.i32p(dist), _lib.i32p(par)))
    assert np.array_equal(dist, base[0]) and np.array_equal(par, bpar[0])
    # the public form: gdistances / bfs_parents conventions per row
    ds, ps = traversals.bfs_many(g, roots[:3] + 1)
    for i in range(3):
        assert np.array_equal(ds[i], lgmi.gdistances(g, int(roots[i]) + 1))
        assert np.array_equal(ps[i], lgmi.bfs_parents(g, int(roots[i]) + 1))
