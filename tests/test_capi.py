"""CPU tests of liblgmi.so's boundary: it loads, exports every symbol the
headers declare, validates its inputs on the host, and its host-only
generators / CSR builder behave.  No GPU compute happens here."""
import os
import re

import numpy as np
import pytest

from conftest import REPO


def _declared():
    names = set()
    for h in ("lgmi.h", "lgmi_gen.h"):
        txt = open(os.path.join(REPO, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(lgmi_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    from lgmi import _lib
    lib = _lib.load()
    declared = _declared()
    assert declared == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name


def test_device_count_and_nodev_error():
    from lgmi import _lib
    lib = _lib.load()
    n = lib.lgmi_device_count()
    assert n >= 0
    if n == 0:
        rp = np.array([0, 1, 2], dtype=np.int32)
        col = np.array([1, 0], dtype=np.int32)
        h = _lib.ctypes.c_void_p()
        rc = lib.lgmi_graph_create(2, 2, _lib.i32p(rp), _lib.i32p(col), 0, None, None, 0, _lib.ctypes.byref(h))
        assert rc == _lib.LGMI_E_NODEV
        assert "GPU" in _lib.last_error()


@pytest.mark.parametrize("rp,col,code", [
    ([0, 2, 2], [1, 0], -1),        # unsorted row
    ([0, 1, 2], [5, 0], -2),        # neighbour out of range
    ([0, 2, 3], [1, 1, 0], -1),     # duplicate in row
    ([0, 1, 1], [1, 0], -1),        # row_ptr[nv] != nnz
])
def test_create_rejects_bad_csr_on_host(rp, col, code):
    from lgmi import _lib
    lib = _lib.load()
    rp = np.array(rp, dtype=np.int32)
    col = np.array(col, dtype=np.int32)
    h = _lib.ctypes.c_void_p()
    rc = lib.lgmi_graph_create(2, len(col), _lib.i32p(rp), _lib.i32p(col), 0, None, None, 0, _lib.ctypes.byref(h))
    assert rc == code
    with pytest.raises(IndexError if code == -2 else ValueError):
        _lib.check(rc)


def test_too_large_code():
    from lgmi import _lib
    lib = _lib.load()
    rp = np.zeros(2, dtype=np.int32)
    h = _lib.ctypes.c_void_p()
    rc = lib.lgmi_graph_create(1, 1 << 31, _lib.i32p(rp), _lib.i32p(rp), 0, None, None, 0, _lib.ctypes.byref(h))
    assert rc == _lib.LGMI_E_TOO_LARGE


def test_policy_argument_checks():
    from lgmi import _lib
    lib = _lib.load()
    assert lib.lgmi_bfs_set_policy(7, 15.0, 18.0, None, 0) == _lib.LGMI_E_ARG
    assert lib.lgmi_bfs_set_policy(0, -1.0, 18.0, None, 0) == _lib.LGMI_E_ARG
    w = np.zeros(10, dtype=np.float32)
    assert lib.lgmi_bfs_set_policy(1, 15.0, 18.0, _lib.f32p(w), 10) == _lib.LGMI_E_ARG
    assert lib.lgmi_bfs_set_policy(0, 15.0, 18.0, None, 0) == _lib.LGMI_OK


def test_batch_tuning_argument_checks():
    from lgmi import _lib
    lib = _lib.load()
    assert lib.lgmi_betweenness_batch_tuning(-2, 0.0, -1, 0, 0) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, -1.0, -1, 0, 0) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, float("nan"), -1, 0, 0) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, 0.0, -5, 0, 0) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, 0.0, -1, 5, 0) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, 0.0, -1, 4, 3) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, 0.0, -1, 16, 4) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, 0.0, -1, 4, 32) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, 0.0, -1, 2, 16) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_batch_tuning(-1, 0.0, -1, 0, 0) == _lib.LGMI_OK


def test_bc_launch_mode_argument_checks():
    from lgmi import _lib
    lib = _lib.load()
    assert lib.lgmi_betweenness_set_launch(2) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_set_launch(-1) == _lib.LGMI_E_ARG
    assert lib.lgmi_betweenness_set_launch(1) == _lib.LGMI_OK
    assert lib.lgmi_betweenness_set_launch(0) == _lib.LGMI_OK


def test_graph_from_edges_argument_checks():
    from lgmi import _lib
    lib = _lib.load()
    h = _lib.ctypes.c_void_p()
    ne = np.zeros(1, dtype=np.int64)
    a = np.zeros(2, dtype=np.int32)
    assert lib.lgmi_graph_from_edges(_lib.i32p(a), _lib.i32p(a), -1, 0, 0, _lib.ctypes.byref(h), _lib.i64p(ne)) == \
        _lib.LGMI_E_ARG
    assert lib.lgmi_graph_from_edges(None, None, 2, 0, 0, _lib.ctypes.byref(h), _lib.i64p(ne)) == _lib.LGMI_E_ARG
    assert lib.lgmi_graph_from_edges(_lib.i32p(a), _lib.i32p(a), 2, 0, -3, _lib.ctypes.byref(h), _lib.i64p(ne)) == \
        _lib.LGMI_E_ARG


def test_generators_shapes_and_determinism():
    import lgmi.generators as gen
    s1, d1 = gen.kronecker_edges(10, 16, seed=3)
    s2, d2 = gen.kronecker_edges(10, 16, seed=3)
    s3, _ = gen.kronecker_edges(10, 16, seed=4)
    assert len(s1) == 16 << 10 and (s1 == s2).all() and (d1 == d2).all() and not (s1 == s3).all()
    assert s1.min() >= 0 and s1.max() < 1024
    # RMAT skew: the a-quadrant bias concentrates edges on few vertices
    deg = np.bincount(np.concatenate([s1, d1]), minlength=1024)
    assert deg.max() > 20 * deg.mean()
    g = gen.erdos_renyi(1000, 5000, seed=1)
    assert g.ne() == 5000 and not g.is_directed()
    g = gen.barabasi_albert(500, 4, seed=2)
    assert g.ne() == 4 + (500 - 5) * 4
    g = gen.watts_strogatz(600, 8, 0.1, seed=3)
    assert g.ne() == 600 * 4
    assert (g.degree_array() >= 1).all()
    g = gen.grid([7, 9])
    assert g.nv() == 63 and g.ne() == 2 * 63 - 7 - 9
    # grid ids: (i, j) -> i*r + j (operators.jl:536); 0-based vertex 1 neighbours 0, 2, 1+7
    assert (g.outneighbors(2) - 1).tolist() == [0, 2, 8]


def test_csr_build_sorted_dedup_selfloops():
    from lgmi.graphs import csr_from_edges
    src = np.array([3, 0, 0, 2, 2, 1, 1])
    dst = np.array([0, 3, 3, 2, 1, 2, 0])
    rp, col = csr_from_edges(4, src, dst, symmetrize=True, drop_self_loops=False)
    rows = [col[rp[v]:rp[v + 1]].tolist() for v in range(4)]
    assert rows == [[1, 3], [0, 2], [1, 2], [0]]
    rp, col = csr_from_edges(4, src, dst, symmetrize=True, drop_self_loops=True)
    rows = [col[rp[v]:rp[v + 1]].tolist() for v in range(4)]
    assert rows == [[1, 3], [0, 2], [1], [0]]
    rp, col = csr_from_edges(4, src, dst, symmetrize=False)
    rows = [col[rp[v]:rp[v + 1]].tolist() for v in range(4)]
    assert rows == [[3], [0, 2], [1, 2], [0]]


def test_tuning_set_get_and_range_checks():
    """lgmi_set_tuning / lgmi_get_tuning round-trip on the host, reject
    out-of-range thresholds, and NULL restores the defaults."""
    from lgmi import _lib, tuning
    lib = _lib.load()
    _lib.check(lib.lgmi_set_tuning(None))
    d = tuning.get()
    assert d == {"wave_row": 32, "lane_rounds": 4, "range_front": 2048, "queue_front": 4096, "range_light": 32768,
                 "queue_min_edges": 4096, "queue_max_edges": 262144, "bc_seg_len": 1024, "solo_front": 4096,
                 "solo_edges": 2048}
    try:
        tuning.set_values({"wave_row": 64, "queue_min_edges": 0, "bc_seg_len": 512})
        got = tuning.get()
        assert got["wave_row"] == 64 and got["queue_min_edges"] == 0 and got["bc_seg_len"] == 512
        for bad in ({"wave_row": 1}, {"lane_rounds": 65}, {"range_front": 4096}, {"queue_front": 5000},
                    {"bc_seg_len": 32}, {"queue_max_edges": -1}, {"solo_front": 4097}, {"solo_edges": -1},
                    {"solo_edges": 262145}):
            with pytest.raises(ValueError):
                tuning.set_values(bad)
        assert tuning.get() == got        # a rejected setting changes nothing
    finally:
        _lib.check(lib.lgmi_set_tuning(None))
    assert tuning.get() == d


@pytest.mark.gpu
def test_create_rejects_unmirrored_csr():
    """An undirected CSR that is not symmetric, or an in-CSR that is not
    the transpose of the out-CSR, is rejected on the device (the bottom-up
    steps and betweenness read the reverse adjacency)."""
    from lgmi import _lib
    lib = _lib.load()
    h = _lib.ctypes.c_void_p()
    rp = np.array([0, 1, 1, 2], dtype=np.int32)          # 0 -> 1 but not 1 -> 0; 2 -> 0
    col = np.array([1, 0], dtype=np.int32)
    rc = lib.lgmi_graph_create(3, 2, _lib.i32p(rp), _lib.i32p(col), 0, None, None, 0, _lib.ctypes.byref(h))
    assert rc == _lib.LGMI_E_ARG and "symmetric" in _lib.last_error()
    # directed 0 -> 1, 2 -> 0 with an in-CSR of 0 -> 1, 0 -> 2 (reversed arc)
    irp = np.array([0, 0, 1, 2], dtype=np.int32)
    icol = np.array([0, 2], dtype=np.int32)
    rc = lib.lgmi_graph_create(3, 2, _lib.i32p(rp), _lib.i32p(col), 2, _lib.i32p(irp), _lib.i32p(icol), 1,
                               _lib.ctypes.byref(h))
    assert rc == _lib.LGMI_E_ARG and "transpose" in _lib.last_error()
    # the right in-CSR is accepted: in(0) = [2], in(1) = [0]
    irp = np.array([0, 1, 2, 2], dtype=np.int32)
    icol = np.array([2, 0], dtype=np.int32)
    _lib.check(lib.lgmi_graph_create(3, 2, _lib.i32p(rp), _lib.i32p(col), 2, _lib.i32p(irp), _lib.i32p(icol), 1,
                                     _lib.ctypes.byref(h)))
    lib.lgmi_graph_destroy(h)


def test_create_rejects_bad_csr_threaded_check():
    """A CSR of 4M entries is checked by several threads over vertex ranges:
    every violation is found wherever it lies, and when there are several
    the one at the smallest vertex is reported, as one pass would."""
    from lgmi import _lib
    lib = _lib.load()
    nv, d = 200_000, 20
    rp = np.arange(0, (nv + 1) * d, d, dtype=np.int32)
    col = ((np.arange(nv, dtype=np.int64)[:, None] + 1 + np.arange(d)[None, :] * 7) % nv).astype(np.int32)
    col.sort(axis=1)
    base = np.ascontiguousarray(col.reshape(-1))

    def create(c):
        h = _lib.ctypes.c_void_p()
        return lib.lgmi_graph_create(nv, len(c), _lib.i32p(rp), _lib.i32p(c), 0, None, None, 0, _lib.ctypes.byref(h))

    def with_unsorted(c, v):
        c[v * d], c[v * d + 1] = c[v * d + 1], c[v * d]

    c = base.copy()
    with_unsorted(c, 150_000)
    assert create(c) == -1
    c = base.copy()
    c[190_000 * d + 3] = nv + 5
    assert create(c) == -2
    c = base.copy()                       # unsorted row first, out-of-range id later
    with_unsorted(c, 50_000)
    c[190_000 * d + 3] = nv + 5
    assert create(c) == -1
    c = base.copy()                       # out-of-range id first, unsorted row later
    c[30_000 * d + 3] = -7
    with_unsorted(c, 150_000)
    assert create(c) == -2
    rp_bad = rp.copy()                    # a row pointer past nnz in a late range
    rp_bad[180_000] = rp_bad[-1] + 100
    h = _lib.ctypes.c_void_p()
    rc = lib.lgmi_graph_create(nv, len(base), _lib.i32p(rp_bad), _lib.i32p(base), 0, None, None, 0, _lib.ctypes.byref(h))
    assert rc == -1
