"""loadgraph / loadgraphs / savegraph in the LG text format, with the graphs
built on the device.

Mirrors src/persistence/common.jl:1-125 and src/persistence/lg.jl:1-160:

* a file may be gzip-compressed (auto_decompress, common.jl:54-72: the
  magic bytes 0x1f 0x8b decide);
* lines starting with ``#`` and blank lines between graphs are skipped
  (loadlg_mult / loadlg, lg.jl:87-112);
* a header ``nv, ne, {d|u}, name[, ver, datatype, code]`` is split as
  _parse_header splits it (lg.jl:63-83), with its pattern ``r"s*,s*"`` as
  written -- an ``s`` just before or after a comma belongs to the separator,
  so a graph named ``graphs`` followed by a version field is read as
  ``graph``;
* the ne lines after a header are the graph's edges (_lg_read_one_graph,
  lg.jl:35-52), parsed by lgmi_lg_parse_edges (host C++); add_edge! drops
  an edge with an endpoint outside 1:nv and ignores duplicates, and keeps
  self-loops, so the graph is built from the in-range pairs by the device
  ingest (lgmi_graph_from_edges) with nv taken from the header.

Writing follows savelg (lg.jl:120-130): a version-2 header with the graph's
element type, then ``src,dst`` per edge in edges(g) order (lexicographic;
u <= v for an undirected graph).
"""
from __future__ import annotations

import gzip
import io as _io
import re

import numpy as np

from . import _lib
from .graphs import SimpleDiGraph, SimpleGraph, _GraphBase

# Julia element type names a version-2 header may carry
_DTYPES = {
    "Int8": np.int8, "Int16": np.int16, "Int32": np.int32, "Int64": np.int64, "Int": np.int64,
    "UInt8": np.uint8, "UInt16": np.uint16, "UInt32": np.uint32, "UInt64": np.uint64, "UInt": np.uint64,
}
_JULIA_NAME = {np.dtype(np.int8): "Int8", np.dtype(np.int16): "Int16", np.dtype(np.int32): "Int32",
               np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8", np.dtype(np.uint16): "UInt16",
               np.dtype(np.uint32): "UInt32", np.dtype(np.uint64): "UInt64"}

_SEP = re.compile(r"s*,s*")   # the reference's header pattern, as written (lg.jl:65,67)


class LGHeader:
    """LGHeader (lg.jl:17-33)."""

    def __init__(self, nv: int, ne: int, is_directed: bool, name: str, ver: int = 1, dtype=np.int64,
                 code: str = "simplegraph"):
        self.nv, self.ne, self.is_directed, self.name = int(nv), int(ne), bool(is_directed), name
        self.ver, self.dtype, self.code = int(ver), np.dtype(dtype), code

    def __str__(self):
        d = "d" if self.is_directed else "u"
        return f"{self.nv},{self.ne},{d},{self.name},{self.ver},{_JULIA_NAME[self.dtype]},{self.code}"


def _parse_int(s: str) -> int:
    """parse(Int, s): blanks around the number allowed."""
    t = s.strip()
    try:
        if re.fullmatch(r"[+-]?(0x[0-9a-fA-F]+|0o[0-7]+|0b[01]+|[0-9]+)", t) is None:
            raise ValueError
        return int(t, 0) if re.match(r"[+-]?0[xob]", t) else int(t, 10)
    except ValueError:
        raise ValueError(f"invalid base 10 digit in {s!r}") from None


def _parse_header(s: str) -> LGHeader:
    """_parse_header (lg.jl:63-83)."""
    parts = _SEP.split(s, maxsplit=3)
    if len(parts) < 4:
        raise ValueError(f"malformed LG header {s!r}")
    nvstr, nestr, dirundir, graphname = parts[:4]
    ver, dtype, code = 1, np.int64, "simplegraph"
    if "," in graphname:
        extra = _SEP.split(graphname)
        if len(extra) < 4:
            raise ValueError(f"malformed LG header {s!r}")
        graphname, sver, sdtype, code = extra[:4]
        ver = _parse_int(sver)
        if sdtype not in _DTYPES:
            raise ValueError(f"unknown element type {sdtype!r} in LG header")
        dtype = _DTYPES[sdtype]
    return LGHeader(_parse_int(nvstr), _parse_int(nestr), dirundir.strip() != "u", graphname.strip(), ver,
                    dtype, code)


class _Reader:
    """The decompressed bytes of a file or stream, read line by line."""

    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0
        self._nl = None
        self._view = np.frombuffer(data, dtype=np.uint8) if data else np.zeros(1, dtype=np.uint8)

    def address(self) -> int:
        return int(self._view.ctypes.data)

    def eof(self) -> bool:
        return self.pos >= len(self.data)

    def readline(self) -> str:
        e = self.data.find(b"\n", self.pos)
        e = len(self.data) if e < 0 else e
        line = self.data[self.pos:e]
        self.pos = min(e + 1, len(self.data))
        if line.endswith(b"\r"):
            line = line[:-1]
        return line.decode()

    def lines_left(self) -> int:
        """Lines from the current position to the end (the last one may
        lack its line break)."""
        if self._nl is None:
            self._nl = np.flatnonzero(self._view[:len(self.data)] == 10)
        return int(len(self._nl) - np.searchsorted(self._nl, self.pos)) + 1

    def skip_lines(self, n: int):
        """_lg_skip_one_graph (lg.jl:54-58): n readlines, not parsed."""
        if n <= 0:
            return
        if self._nl is None:
            self._nl = np.flatnonzero(self._view[:len(self.data)] == 10)
        i = int(np.searchsorted(self._nl, self.pos)) + n - 1
        self.pos = int(self._nl[i]) + 1 if i < len(self._nl) else len(self.data)


def _auto_decompress(data: bytes) -> bytes:
    """auto_decompress (common.jl:54-72)."""
    if len(data) >= 2 and data[0] == 0x1F and data[1] == 0x8B:
        return gzip.decompress(data)
    return data


def _read_all(src) -> bytes:
    if isinstance(src, (bytes, bytearray)):
        return bytes(src)
    if hasattr(src, "read"):
        d = src.read()
        return d.encode() if isinstance(d, str) else d
    with open(src, "rb") as f:
        return f.read()


def _parse_edges(r: _Reader, h: LGHeader):
    """The header's ne edge lines (lg.jl:39-50) as 1-based (src, dst)
    arrays, the pairs add_edge! would drop (an endpoint outside 1:nv)
    removed; advances the reader past the lines."""
    # a header may promise more lines than the text holds; those read as ""
    nl = max(0, min(h.ne, r.lines_left()))
    cap = max(nl, 1)
    s = np.empty(cap, dtype=np.int64)
    d = np.empty(cap, dtype=np.int64)
    npairs = np.zeros(1, dtype=np.int64)
    consumed = np.zeros(1, dtype=np.int64)
    bad = np.zeros(1, dtype=np.int64)
    rc = _lib.load().lgmi_lg_parse_edges(r.address() + r.pos, len(r.data) - r.pos, nl, _lib.i64p(s),
                                         _lib.i64p(d), _lib.i64p(npairs), _lib.i64p(consumed), _lib.i64p(bad))
    if rc == -1:
        raise IndexError(f"graph {h.name!r}, edge line {int(bad[0])}: expected 'src,dst'")
    if rc == -2:
        raise ValueError(f"graph {h.name!r}, edge line {int(bad[0])}: not an integer")
    if rc == -3:
        raise OverflowError(f"graph {h.name!r}, edge line {int(bad[0])}: overflows Int64")
    if rc != 0:
        raise ValueError("lgmi_lg_parse_edges: bad arguments")
    r.pos += int(consumed[0])
    k = int(npairs[0])
    s, d = s[:k], d[:k]
    if k == 0:
        return s, d
    info = np.iinfo(h.dtype)
    lo, hi = min(s.min(), d.min()), max(s.max(), d.max())
    if lo < info.min or hi > info.max:
        raise OverflowError(f"graph {h.name!r}: an endpoint does not fit {_JULIA_NAME[h.dtype]}")
    if lo >= 1 and hi <= h.nv:
        return s, d
    keep = (s >= 1) & (s <= h.nv) & (d >= 1) & (d <= h.nv)
    return s[keep], d[keep]


def _read_one_graph(r: _Reader, h: LGHeader) -> _GraphBase:
    """_lg_read_one_graph (lg.jl:35-52): the header's ne lines, as edges
    added with add_edge! to a graph of nv vertices, built on the device."""
    cls = SimpleDiGraph if h.is_directed else SimpleGraph
    s, d = _parse_edges(r, h)
    g, _ = cls.from_edge_list(s, d, T=h.dtype, one_based=True, nv_min=h.nv)
    return g


def loadlg_mult(src) -> dict:
    """loadlg_mult(io) (lg.jl:87-98): name => graph for every graph."""
    r = _Reader(_auto_decompress(_read_all(src)))
    graphs = {}
    while not r.eof():
        line = r.readline().strip()
        if not (line.startswith("#") or line == ""):
            h = _parse_header(line)
            graphs[h.name] = _read_one_graph(r, h)
    return graphs


def loadlg(src, gname: str) -> _GraphBase:
    """loadlg(io, gname) (lg.jl:100-112); ValueError (the reference's
    ArgumentError) when no graph has that name."""
    r = _Reader(_auto_decompress(_read_all(src)))
    while not r.eof():
        line = r.readline().strip()
        if line.startswith("#") or line == "":
            continue
        h = _parse_header(line)
        if h.name == gname:
            return _read_one_graph(r, h)
        r.skip_lines(h.ne)
    raise ValueError(f"graph {gname} not found")


def loadgraph(fn, gname: str = "graph") -> _GraphBase:
    """loadgraph(file[, gname]) (common.jl:3-19): the graph named gname
    ("graph" by default)."""
    return loadlg(fn, gname)


def loadgraphs(fn) -> dict:
    """loadgraphs(file) (common.jl:22-39)."""
    return loadlg_mult(fn)


def _edge_pairs(g: _GraphBase):
    rp, col = g.csr()
    rp = np.asarray(rp, dtype=np.int64)
    u = np.repeat(np.arange(g.nv(), dtype=np.int64), np.diff(rp))
    v = np.asarray(col, dtype=np.int64)
    if not g.is_directed():
        keep = v >= u
        u, v = u[keep], v[keep]
    return u + 1, v + 1


def savelg(out, g: _GraphBase, gname: str) -> int:
    """savelg(io, g, gname) (lg.jl:120-130): returns 1."""
    h = LGHeader(g.nv(), g.ne(), g.is_directed(), gname, 2, g.T, "simplegraph")
    u, v = _edge_pairs(g)
    body = "".join(f"{a},{b}\n" for a, b in zip(u.tolist(), v.tolist()))
    out.write(f"{h}\n" + body)
    return 1


def savelg_mult(out, graphs: dict) -> int:
    """savelg_mult(io, graphs) (lg.jl:137-143)."""
    return sum(savelg(out, g, name) for name, g in graphs.items())


def savegraph(fn, g, gname: str = "graph") -> int:
    """savegraph(file, g[, gname]) / savegraph(file, dict) (common.jl:75-125),
    uncompressed; fn may also be a text stream."""
    if hasattr(fn, "write"):
        return savelg_mult(fn, g) if isinstance(g, dict) else savelg(fn, g, gname)
    with open(fn, "w") as f:
        return savelg_mult(f, g) if isinstance(g, dict) else savelg(f, g, gname)


def sprint_savegraph(g: _GraphBase) -> str:
    """sprint(savegraph, g, LGFormat()): the text savegraph writes for g
    under the default name."""
    s = _io.StringIO()
    savelg(s, g, "graph")
    return s.getvalue()
