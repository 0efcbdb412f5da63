"""Graph files, restated in plain Python: the reference's load_graph
(runtime/infra_gpu/graph.h:236-303) with the GAPBS reader and builder it
calls (runtime/infra_gapbs/reader.h:51-375, builder.h:62-246).

TEST INFRASTRUCTURE ONLY (the oracle): imported by tests, never by the
library. Pinned against the reference's own reader and builder compiled from
its sources into oracle/_ref/ (tests/test_graph_files.py).

The text is scanned as a C++ stream reads int32 values (`in >> x`):
whitespace skipped, an optional sign, decimal digits; a read that finds no
digits, or a value outside int32, fails.
"""
from __future__ import annotations

import struct

import numpy as np

_WS = b" \t\n\v\f\r"
I32_MIN, I32_MAX = -(1 << 31), (1 << 31) - 1


class _Cursor:
    def __init__(self, data: bytes, pos: int = 0, end: int | None = None):
        self.d, self.p = data, pos
        self.end = len(data) if end is None else end

    def at_end(self) -> bool:
        return self.p >= self.end

    def peek(self) -> int:
        return self.d[self.p] if self.p < self.end else -1

    def skip_space(self):
        while self.p < self.end and self.d[self.p] in _WS:
            self.p += 1

    def read_int(self, lo=I32_MIN, hi=I32_MAX):
        self.skip_space()
        q, neg = self.p, False
        if q < self.end and self.d[q] in b"+-":
            neg = self.d[q] == ord("-")
            q += 1
        s = q
        while q < self.end and 48 <= self.d[q] <= 57:
            q += 1
        if q == s:
            return None
        x = int(self.d[s:q])
        x = -x if neg else x
        if x < lo or x > hi:
            return None
        self.p = q
        return x

    def read_word(self) -> bytes:
        self.skip_space()
        q = self.p
        while self.p < self.end and self.d[self.p] not in _WS:
            self.p += 1
        return self.d[q:self.p]

    def ignore200(self):  # in.ignore(200, '\n')
        for _ in range(200):
            if self.p >= self.end:
                return
            c = self.d[self.p]
            self.p += 1
            if c == 10:
                return

    def getline(self) -> "_Cursor":
        q = self.p
        while self.p < self.end and self.d[self.p] != 10:
            self.p += 1
        line = _Cursor(self.d, q, self.p)
        if self.p < self.end:
            self.p += 1
        return line


def _el(c: _Cursor, weighted: bool):  # reader.h:51-68
    out = []
    while True:
        a = c.read_int()
        b = c.read_int() if a is not None else None
        w = 1
        if b is not None and weighted:
            w = c.read_int()
        if a is None or b is None or w is None:
            return out
        out.append((a, b, w))


def _gr(c: _Cursor):  # reader.h:70-85
    out = []
    while not c.at_end():
        if c.peek() == ord("a"):
            c.p += 1
            a = c.read_int()
            b = c.read_int() if a is not None else None
            w = c.read_int() if b is not None else None
            if w is None:
                return out
            out.append((a, b, w))
        else:
            c.ignore200()
    return out


def _metis(c: _Cursor):  # reader.h:88-141
    while True:
        if c.at_end():
            raise ValueError("METIS file without a header")
        if c.peek() == ord("%"):
            c.ignore200()
            continue
        h = c.getline()
        n, m = h.read_int(), h.read_int()
        if n is None or m is None:
            raise ValueError("bad METIS header")
        h.skip_space()
        weights = False
        if not h.at_end():
            fmt = h.read_int()
            if fmt == 1:
                weights = True
            elif fmt not in (0, 100):
                raise ValueError(f"METIS fmt {fmt}")
        break
    out, u = [], 0
    while u < n:
        if c.peek() == ord("%"):
            c.ignore200()
            continue
        line = c.getline()
        while True:
            v = line.read_int()
            w = 1
            if v is not None and weights:
                w = line.read_int()
            # `edge_stream >> v >> std::ws` fails once the digits ended the
            # line (libstdc++): the entry is taken only when more text follows
            if v is None or w is None or line.at_end():
                break
            out.append((u, v - 1, w))
        u += 1
    return out


def _mtx(c: _Cursor):  # reader.h:145-213
    start, obj, fmt, field, sym = (c.read_word() for _ in range(5))
    c.skip_space()
    if start != b"%%MatrixMarket" or obj != b"matrix" or fmt != b"coordinate":
        raise ValueError("not a MatrixMarket coordinate matrix")
    if field == b"pattern":
        weights = False
    elif field in (b"real", b"double", b"integer"):
        weights = True
    else:
        raise ValueError(f"field {field!r}")
    if sym == b"symmetric":
        undirected = True
    elif sym in (b"general", b"skew-symmetric"):
        undirected = False
    else:
        raise ValueError(f"symmetry {sym!r}")
    while c.peek() == ord("%"):
        c.ignore200()
    m, n = c.read_int(-(1 << 63), (1 << 63) - 1), c.read_int(-(1 << 63), (1 << 63) - 1)
    c.read_int(-(1 << 63), (1 << 63) - 1)
    c.skip_space()
    if m != n:
        raise ValueError("matrix must be square")
    out = []
    while not c.at_end():
        line = c.getline()
        a, b = line.read_int(), None
        if a is not None:
            b = line.read_int()
        if b is None:
            continue
        w = 1
        if weights:
            w = line.read_int()
            w = 0 if w is None else w
        out.append((a - 1, b - 1, w))
        if undirected:
            out.append((b - 1, a - 1, w))
    return out


def _build(edges, symmetrize: bool):
    """builder.h:62-71, 134-226: V = largest id + 1, lists sorted by
    (neighbour, weight), one entry per neighbour (least weight), no
    self-loops."""
    vmax = 0
    for a, b, _ in edges:
        if a < 0 or b < 0:
            raise ValueError("negative id")
        vmax = max(vmax, a, b)
    V = vmax + 1
    lists = [[] for _ in range(V)]
    for a, b, w in edges:
        lists[a].append((b, w))
        if symmetrize:
            lists[b].append((a, w))
    off, nb, wt = [0], [], []
    for u in range(V):
        last = None
        for v, w in sorted(lists[u]):
            if v == u or v == last:
                continue
            nb.append(v)
            wt.append(w)
            last = v
        off.append(len(nb))
    return (np.array(off, np.int32), np.array(nb, np.int32), np.array(wt, np.int32))


def _serialized(data: bytes, weighted: bool):  # reader.h:332-373
    directed = data[0] != 0
    E, V = struct.unpack_from("<qq", data, 1)
    at = 17
    off = np.frombuffer(data, np.int64, V + 1, at).astype(np.int32)
    at += 8 * (V + 1)
    if weighted:
        pairs = np.frombuffer(data, np.int32, 2 * E, at)
        return off, pairs[0::2].copy(), pairs[1::2].copy(), directed
    return off, np.frombuffer(data, np.int32, E, at).copy(), None, directed


def _graphit_bin(data: bytes):  # graph.h:252-268
    V, E = struct.unpack_from("<ii", data, 0)
    body = np.frombuffer(data, np.int32, 3 * E + V + 1, 8)
    return body[3 * E:].copy(), body[E:2 * E].copy(), body[2 * E:3 * E].copy(), True


def read_graph_file(path: str, symmetrize: bool = False):
    """(offsets, nbrs, weights or None, directed) as the reference loads the
    file (text formats always carry weights: 1 where the format has none)."""
    with open(path, "rb") as f:
        data = f.read()
    if path.endswith(".graphit_bin"):
        return _graphit_bin(data)
    suffix = path[path.rfind("."):]
    if suffix in (".sg", ".wsg"):
        return _serialized(data, suffix == ".wsg")
    c = _Cursor(data)
    if suffix == ".el":
        edges = _el(c, False)
    elif suffix == ".wel":
        edges = _el(c, True)
    elif suffix == ".gr":
        edges = _gr(c)
    elif suffix == ".graph":
        edges = _metis(c)
    elif suffix == ".mtx":
        edges = _mtx(c)
    else:
        raise ValueError(f"suffix {suffix}")
    off, nb, wt = _build(edges, symmetrize)
    return off, nb, wt, not symmetrize


def graphit_bin_bytes(offsets, nbrs, weights) -> bytes:
    """graph.h:286-293: int32 V, int32 E, sources, neighbours, weights,
    offsets."""
    off = np.asarray(offsets, np.int32)
    V, E = len(off) - 1, int(off[-1])
    src = np.repeat(np.arange(V, dtype=np.int32), np.diff(off))
    w = np.ones(E, np.int32) if weights is None else np.asarray(weights, np.int32)
    return (struct.pack("<ii", V, E) + src.tobytes() + np.asarray(nbrs, np.int32).tobytes()
            + w.tobytes() + off.tobytes())
