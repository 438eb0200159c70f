"""The LG text format on the host: header parsing (_parse_header,
src/persistence/lg.jl:63-83), the edge-line parser lgmi_lg_parse_edges
(_lg_read_one_graph, lg.jl:35-52) and savelg's text (lg.jl:120-130),
against the reference's own test data and facts
(test/persistence/persistence.jl:6-45).  No GPU: graphs are built with the
host CSR builder here; loading through the device is in
test_persistence_gpu.py."""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

LG = os.path.join(GOLDEN, "lg")


def parse(text: bytes, nlines: int):
    from lgmi import _lib
    buf = np.frombuffer(text, dtype=np.uint8) if text else np.zeros(1, np.uint8)
    s = np.zeros(max(nlines, 1), np.int64)
    d = np.zeros(max(nlines, 1), np.int64)
    out = np.zeros(3, np.int64)
    rc = _lib.load().lgmi_lg_parse_edges(int(buf.ctypes.data), len(text), nlines, _lib.i64p(s), _lib.i64p(d),
                                         _lib.i64p(out[0:]), _lib.i64p(out[1:]), _lib.i64p(out[2:]))
    k = int(out[0])
    return rc, list(zip(s[:k].tolist(), d[:k].tolist())), int(out[1]), int(out[2])


def host_graphs(path):
    """loadgraphs on the host: the module's header and edge parsing, the
    host CSR builder instead of the device ingest."""
    import lgmi
    from lgmi import persistence as P
    r = P._Reader(P._auto_decompress(P._read_all(path)))
    out = {}
    while not r.eof():
        line = r.readline().strip()
        if line.startswith("#") or line == "":
            continue
        h = P._parse_header(line)
        s, d = P._parse_edges(r, h)
        cls = lgmi.SimpleDiGraph if h.is_directed else lgmi.SimpleGraph
        out[h.name] = cls.from_edges(h.nv, s, d, T=h.dtype, one_based=True)
    return out


def test_parse_header():
    from lgmi.persistence import _parse_header
    h = _parse_header("46,69,u,Tutte")
    assert (h.nv, h.ne, h.is_directed, h.name, h.ver, h.dtype) == (46, 69, False, "Tutte", 1, np.int64)
    h = _parse_header("10,9,d,pathdigraph")
    assert (h.nv, h.ne, h.is_directed, h.name) == (10, 9, True, "pathdigraph")
    h = _parse_header("10,9,d,graph,2,Int32,simplegraph")
    # the separator r"s*,s*" also takes the first letter of "simplegraph"
    assert (h.name, h.ver, h.dtype, h.code) == ("graph", 2, np.int32, "implegraph")
    # blanks around the numbers; anything but "u" is directed
    h = _parse_header(" 7 , 3 ,x,name")
    assert (h.nv, h.ne, h.is_directed, h.name) == (7, 3, True, "name")
    # the pattern r"s*,s*" (lg.jl:65,67): the trailing s of "graphs" is part
    # of the separator once a version field follows
    assert _parse_header("5,4,u,graphs,2,Int64,simplegraph").name == "graph"
    assert _parse_header("5,4,u,graphs").name == "graphs"
    with pytest.raises(ValueError):
        _parse_header("5,4,u")
    with pytest.raises(ValueError):
        _parse_header("5,x,u,g")
    with pytest.raises(ValueError):
        _parse_header("5,4,u,g,2,Float64,simplegraph")


def test_parse_edges_lines():
    # blank lines count among the ne; blanks around commas and numbers; a
    # third field is ignored; "\r\n" endings
    text = b"1, 3\n\n 2 ,4\r\n5,6,7\n8,9\n"
    rc, pairs, consumed, _ = parse(text, 4)
    assert rc == 0 and pairs == [(1, 3), (2, 4), (5, 6)]
    assert consumed == len(b"1, 3\n\n 2 ,4\r\n5,6,7\n")
    # fewer lines than ne: readline at the end gives "" (skipped)
    rc, pairs, consumed, _ = parse(b"1,2", 5)
    assert rc == 0 and pairs == [(1, 2)] and consumed == 3
    # integer syntax of parse(Int64, s)
    rc, pairs, _, _ = parse(b"0x10,-2\n+3,0b101\n0o17,9223372036854775807\n", 3)
    assert rc == 0 and pairs == [(16, -2), (3, 5), (15, 9223372036854775807)]
    rc, pairs, _, _ = parse(b"-9223372036854775808,1\n", 1)
    assert rc == 0 and pairs == [(-9223372036854775808, 1)]


@pytest.mark.parametrize("text,code,line", [
    (b"1,2\n3\n", -1, 2),          # one field
    (b"   \n", -1, 1),             # blanks only: not empty, one field
    (b"1,2\n1,x\n", -2, 2),
    (b"1,\n", -2, 1),
    (b"1 2,3\n", -2, 1),
    (b"0x,1\n", -2, 1),
    (b"1,9223372036854775808\n", -3, 1),
    (b"-9223372036854775809,1\n", -3, 1),
])
def test_parse_edges_errors(text, code, line):
    rc, _, _, bad = parse(text, 5)
    assert (rc, bad) == (code, line)


def test_reference_files_host():
    """test/persistence/persistence.jl:6-12 and 36-40: the two graphs of
    tutte-pathdigraph.jgz; (ne, nv) of the path digraph is (9, 10); savegraph
    writes 421 characters for Tutte and 70 for the path digraph."""
    from lgmi.persistence import sprint_savegraph
    d = host_graphs(os.path.join(LG, "tutte-pathdigraph.jgz"))
    assert sorted(d) == ["Tutte", "pathdigraph"]
    p1, p2 = d["Tutte"], d["pathdigraph"]
    assert (p2.ne(), p2.nv()) == (9, 10)
    # the Tutte graph: 46 vertices, 69 edges, cubic
    assert (p1.nv(), p1.ne()) == (46, 69)
    assert (np.diff(p1.csr()[0]) == 3).all()
    assert p2.csr()[1].tolist() == list(range(1, 10))
    assert len(sprint_savegraph(p1)) == 421
    assert len(sprint_savegraph(p2)) == 70
    assert sprint_savegraph(p2).splitlines()[0] == "10,9,d,graph,2,Int64,simplegraph"


def test_graph_50_500_matches_fixture():
    """graph-50-500.jgz, the betweenness test graph
    (test/centrality/betweenness.jl:8-13), parses to the committed edge
    fixture."""
    d = host_graphs(os.path.join(LG, "graph-50-500.jgz"))
    assert list(d) == ["graph-50-500"]
    g = d["graph-50-500"]
    z = load_golden("ref_graph_50_500.npz")
    assert g.nv() == int(z["n"]) and g.is_directed() == bool(z["directed"])
    from conftest import csr_of
    rp, col, _, _ = csr_of(int(z["n"]), z["src"], z["dst"], bool(z["directed"]))
    assert np.array_equal(g.csr()[0], rp) and np.array_equal(g.csr()[1], col)


def test_graph_decomposition_host():
    d = host_graphs(os.path.join(LG, "graph-decomposition.jgz"))
    g = d["graph"]
    assert (g.nv(), g.ne(), g.is_directed()) == (21, 29, False)


def test_add_edge_semantics_host():
    """add_edge! drops out-of-range endpoints and duplicates and keeps
    self-loops; a blank line is one of the ne."""
    text = b"4,5,u,g\n1,2\n2,1\n3,3\n1,9\n\n"
    g = host_graphs(text)["g"]   # bytes are taken as the file's contents
    assert (g.nv(), g.ne()) == (4, 2)
    assert g.csr()[1].tolist() == [1, 0, 2]


def test_skip_and_comments(tmp_path):
    """loadlg skips other graphs without parsing their lines (lg.jl:54-58,
    100-112); '#' lines and blank lines between graphs are ignored."""
    from lgmi.persistence import _Reader, _parse_header
    text = b"# c\n\n2,2,u,a\nnot,parsed\n1,2\n\n3,1,d,b\n1,3\n"
    r = _Reader(text)
    names = []
    while not r.eof():
        line = r.readline().strip()
        if line.startswith("#") or line == "":
            continue
        h = _parse_header(line)
        names.append(h.name)
        r.skip_lines(h.ne)
    assert names == ["a", "b"]
    # gzip and plain files read the same
    p = tmp_path / "x.lg.gz"
    p.write_bytes(gzip.compress(text))
    from lgmi.persistence import _auto_decompress, _read_all
    assert _auto_decompress(_read_all(str(p))) == text


def _py_parse(text: bytes, nlines: int):
    """The line rules of _lg_read_one_graph (lg.jl:39-50) in plain Python:
    (pairs, bytes consumed, first bad line or 0)."""
    pos, pairs = 0, []
    for line in range(1, nlines + 1):
        if pos >= len(text):
            break
        e = text.find(b"\n", pos)
        e = len(text) if e < 0 else e
        nxt = e + 1 if e < len(text) else e
        body = text[pos:e]
        if e < len(text) and body.endswith(b"\r"):
            body = body[:-1]
        if body:
            f = body.split(b",")
            if len(f) < 2:
                return pairs, pos, line
            try:
                pairs.append((int(f[0]), int(f[1])))
            except ValueError:
                return pairs, pos, line
        pos = nxt
    return pairs, pos, 0


@pytest.mark.parametrize("trail_newline", [True, False])
def test_parse_edges_large_threaded(trail_newline):
    """Several MB of lines, so that the parser splits them over threads:
    blank lines, "\\r\\n" endings, blanks around numbers, third fields, a
    second graph after the first one's lines, and a last line without a
    line break; against the plain-Python reading of the same rules."""
    rng = np.random.default_rng(7)
    n = 300_000
    parts = []
    for i in range(n):
        r = rng.integers(0, 20)
        a, b = rng.integers(-5, 10**9, size=2)
        if r == 0:
            parts.append(b"")
        elif r == 1:
            parts.append(b" %d , %d\r" % (a, b))
        elif r == 2:
            parts.append(b"%d,%d,%d" % (a, b, i))
        else:
            parts.append(b"%d,%d" % (a, b))
    first = b"\n".join(parts) + b"\n"
    second = b"3,1,u,next\n1,2" + (b"\n" if trail_newline else b"")
    text = first + second
    for nl in (n, n - 12_345, n + 7):
        rc, pairs, consumed, bad = parse(text, nl)
        want, wpos, wbad = _py_parse(text, nl)
        assert rc == 0 and bad == 0 and wbad == 0
        assert pairs == want and consumed == wpos
    # a bad line late in the text is the one reported
    bad_text = first[: len(first) // 2] + b"oops\n" + first[len(first) // 2:]
    bad_line = bad_text[: len(first) // 2].count(b"\n") + 1
    rc, pairs, _, bad = parse(bad_text, n + 1)
    assert rc in (-1, -2) and bad == bad_line
    want, _, wbad = _py_parse(bad_text, n + 1)
    assert wbad == bad_line and pairs == want


def test_header_promising_more_lines_than_the_text():
    """A header's ne beyond the end of the text: the missing lines read as
    empty (readline at the end of a stream), and no ne-sized buffer is
    made."""
    from lgmi import persistence as P
    text = b"5,1000000000000,u,g\n1,2\n3,4"
    r = P._Reader(text)
    h = P._parse_header(r.readline().strip())
    s, d = P._parse_edges(r, h)
    assert s.tolist() == [1, 3] and d.tolist() == [2, 4] and r.eof()
