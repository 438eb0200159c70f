"""lgmi_bfs_run_noalloc through the C ABI: the BFS whose search state lives
in the library's fixed device array, with results in caller-owned device
buffers, against lgmi_bfs (the allocating entry point) bit for bit.

Two calls in a row on one GPU, the second on a smaller graph, so that the
second search runs on what the first left in the shared array; also
dir = :in on a digraph, distances only, the learned direction policy, and
the argument errors (a source out of range, a graph above the stated
vertex limit).  The caller allocates its device buffers here with the HIP
runtime directly; the entry point itself allocates nothing.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


class _Hip:
    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 4))) == 0
        # a fill the search must overwrite everywhere; the library runs on
        # its own non-blocking stream, so the fill is complete before the
        # call (the caller's duty for any buffer it passes)
        assert self.h.hipMemset(p, 0x5A, ctypes.c_size_t(max(nbytes, 4))) == 0
        assert self.h.hipDeviceSynchronize() == 0
        return p

    def to_host(self, p, n):
        out = np.empty(n, dtype=np.int32)
        assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(4 * n), 2) == 0
        return out

    def free(self, p):
        self.h.hipFree(p)


@pytest.fixture(scope="module")
def hip():
    return _Hip()


def _noalloc(hip, g, src, dir_in=0, parents=True):
    from lgmi import _lib
    lib = _lib.load()
    n = g.nv()
    d = hip.alloc(4 * n)
    p = hip.alloc(4 * n) if parents else None
    try:
        _lib.check(lib.lgmi_bfs_run_noalloc(g.handle(), int(src), int(dir_in), d, p))
        dist = hip.to_host(d, n)
        par = hip.to_host(p, n) if parents else None
    finally:
        hip.free(d)
        if p is not None:
            hip.free(p)
    return dist, par


def _alloc_path(g, src, dir_in=0):
    from lgmi import _lib
    lib = _lib.load()
    n = g.nv()
    d = np.empty(n, dtype=np.int32)
    p = np.empty(n, dtype=np.int32)
    _lib.check(lib.lgmi_bfs(g.handle(), int(src), int(dir_in), _lib.i32p(d), _lib.i32p(p)))
    return d, p


def test_noalloc_matches_bfs_two_calls_second_smaller(hip):
    import lgmi
    big = lgmi.generators.rmat(16, 16, seed=3)
    small = lgmi.generators.rmat(11, 8, seed=4)
    for g, tag in ((big, "rmat16"), (small, "rmat11")):
        deg = np.diff(g.csr()[0])
        hub = int(np.argmax(deg))
        other = int(np.nonzero(deg > 0)[0][len(deg) // 3 % max(1, np.count_nonzero(deg > 0))])
        for src in (hub, other):
            say(f"noalloc {tag} src {src}")
            d0, p0 = _alloc_path(g, src)
            d1, p1 = _noalloc(hip, g, src)
            assert np.array_equal(d1, d0), (tag, src)
            assert np.array_equal(p1, p0), (tag, src)


def test_noalloc_directed_in_and_distances_only(hip):
    import lgmi
    g = lgmi.generators.rmat(13, 8, seed=5, directed=True)
    rp, _ = g.csr()
    src = int(np.argmax(np.diff(rp)))
    say(f"noalloc directed src {src}")
    for dir_in in (0, 1):
        d0, p0 = _alloc_path(g, src, dir_in)
        d1, p1 = _noalloc(hip, g, src, dir_in)
        assert np.array_equal(d1, d0) and np.array_equal(p1, p0), dir_in
    d2, _ = _noalloc(hip, g, src, 0, parents=False)
    assert np.array_equal(d2, _alloc_path(g, src, 0)[0])


def test_noalloc_learned_policy(hip):
    import lgmi
    from lgmi import policy
    params, _ = policy.load_weights()
    if params is None:
        pytest.skip("no committed policy weights")
    g = lgmi.generators.barabasi_albert(20_000, 4, seed=6)
    src = 17
    d0, p0 = _alloc_path(g, src)
    try:
        policy.set_learned(params)
        say("noalloc learned policy")
        d1, p1 = _noalloc(hip, g, src)
    finally:
        policy.set_heuristic()
    assert np.array_equal(d1, d0) and np.array_equal(p1, p0)


def test_noalloc_argument_errors(hip):
    import lgmi
    from lgmi import _lib
    lib = _lib.load()
    g = lgmi.generators.path_graph(50)
    d = hip.alloc(4 * 50)
    try:
        assert lib.lgmi_bfs_run_noalloc(g.handle(), 50, 0, d, None) == _lib.LGMI_E_BOUNDS
        assert lib.lgmi_bfs_run_noalloc(g.handle(), -1, 0, d, None) == _lib.LGMI_E_BOUNDS
        assert lib.lgmi_bfs_run_noalloc(g.handle(), 0, 0, None, None) == _lib.LGMI_E_ARG
    finally:
        hip.free(d)
    # one vertex above the limit (no edges): rejected before any device work
    n = _lib.LGMI_NOALLOC_MAX_NV + 1
    big = lgmi.SimpleGraph.from_csr(np.zeros(n + 1, dtype=np.int32), np.zeros(0, dtype=np.int32))
    say("noalloc size limit")
    d = hip.alloc(4)
    try:
        rc = lib.lgmi_bfs_run_noalloc(big.handle(), 0, 0, d, None)
        assert rc == _lib.LGMI_E_TOO_LARGE
        assert "LGMI_NOALLOC_MAX_NV" in _lib.last_error()
    finally:
        hip.free(d)
