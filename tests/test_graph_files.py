"""Graph files (.el .wel .gr .graph .mtx .sg .wsg .graphit_bin) read on the
host by the library (gi_read_graph_file) and by the oracle's restatement,
against the reference's own GAPBS reader and builder: the committed
expectations in tests/golden/graphs/expected.npz were written by it
(tests/golden/make_graph_files.py), and when oracle/_ref is built the
randomized files below are also read by it directly. No device is used."""
import glob
import os
import zlib

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "graphs")


@pytest.fixture(scope="module")
def lib():
    import easy_graphit_amd as gi
    gi.lib()
    return gi


def fixtures():
    out = []
    for path in sorted(glob.glob(os.path.join(GOLD, "*"))):
        name = os.path.basename(path)
        if name.endswith(".npz"):
            continue
        for sym in ((False,) if name.endswith((".sg", ".wsg")) else (False, True)):
            out.append((name, sym))
    return out


def expected(name, sym):
    d = np.load(os.path.join(GOLD, "expected.npz"))
    k = f"{name}|{int(sym)}"
    directed, weighted = d[k + "|meta"]
    return d[k + "|offsets"], d[k + "|nbrs"], (d[k + "|weights"] if weighted else None), bool(directed)


def same(got, want):
    (o1, n1, w1, d1), (o2, n2, w2, d2) = got, want
    if (w1 is None) != (w2 is None):
        return False
    return (np.array_equal(o1, o2) and np.array_equal(n1, n2) and d1 == d2
            and (w1 is None or np.array_equal(w1, w2)))


@pytest.mark.parametrize("name,sym", fixtures())
def test_oracle_reads_fixture_as_reference(name, sym):
    assert same(O.read_graph_file(os.path.join(GOLD, name), sym), expected(name, sym))


@pytest.mark.parametrize("name,sym", fixtures())
def test_library_reads_fixture_as_reference(lib, name, sym):
    assert same(lib.read_graph_file(os.path.join(GOLD, name), sym), expected(name, sym))


def test_fixture_quirks_are_the_references():
    # a '#' line ends an .el read; DIMACS ids are kept (vertex 0 isolated);
    # a METIS line's last entry counts only when whitespace follows it;
    # real .mtx values are read as int32 text ("2.9" -> 2, "1e3" -> 1)
    off, nb, w, _ = expected("comment_stops.el", False)
    assert off.tolist() == [0, 1, 2, 2] and nb.tolist() == [1, 2]
    off, nb, w, _ = expected("road.gr", False)
    assert off[1] == 0 and nb.tolist() == [2, 3, 1, 2] and w.tolist() == [4, 5, 2, 8]
    off, nb, w, _ = expected("plain.graph", False)
    assert len(off) == 4 and nb.tolist() == [1, 0, 0, 1]
    off, nb, w, _ = expected("real_general.mtx", False)
    assert w.tolist() == [2, 0, 1]


def _random_edges(rng, n, m):
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    dup = rng.integers(0, m, m // 8)           # repeated pairs with other weights
    u = np.concatenate([u, u[dup], np.arange(0, n, 7)])
    v = np.concatenate([v, v[dup], np.arange(0, n, 7)])  # and self-loops
    w = rng.integers(-50, 1000, len(u))
    return u, v, w


def _write(path, fmt, u, v, w, n, rng, sym=False):
    eol = "\r\n" if rng.random() < 0.3 else "\n"
    lines = []
    if fmt == "el":
        lines = [f"{a} {b}" for a, b in zip(u, v)]
    elif fmt == "wel":
        lines = [f"{a}\t{b} {c}" for a, b, c in zip(u, v, w)]
    elif fmt == "gr":
        lines = ["c generated", f"p sp {n} {len(u)}"] + [f"a {a} {b} {c}" for a, b, c in zip(u, v, w)]
    elif fmt in ("graph", "wgraph"):
        adj = [[] for _ in range(n)]
        for a, b, c in zip(u, v, w):
            adj[a].append((b, c))
        lines = ["% generated", f"{n} {len(u)}" + (" 1" if fmt == "wgraph" else "")]
        for i in range(n):
            toks = [f"{b + 1} {c}" if fmt == "wgraph" else f"{b + 1}" for b, c in adj[i]]
            tail = " " if rng.random() < 0.5 else ""   # trailing space keeps the last entry
            lines.append(" ".join(toks) + tail)
    elif fmt == "mtx":
        field = "integer" if rng.random() < 0.5 else "pattern"
        symmetry = "symmetric" if sym else "general"
        lines = [f"%%MatrixMarket matrix coordinate {field} {symmetry}", "% generated",
                 f"{n} {n} {len(u)}"]
        lines += [f"{a + 1} {b + 1} {c}" if field == "integer" else f"{a + 1} {b + 1}"
                  for a, b, c in zip(u, v, w)]
    with open(path, "w", newline="") as f:
        f.write(eol.join(lines) + eol)


@pytest.mark.parametrize("fmt", ["el", "wel", "gr", "graph", "wgraph", "mtx", "mtx_sym"])
def test_random_files_against_reference_reader(lib, tmp_path, fmt):
    if O.ref() is None:
        pytest.skip("oracle/_ref (the reference's GAPBS reader) is not built")
    rng = np.random.default_rng(zlib.crc32(fmt.encode()))
    for trial in range(3):
        n = int(rng.integers(5, 300))
        u, v, w = _random_edges(rng, n, int(rng.integers(1, 4 * n)))
        suffix = {"wgraph": "graph", "mtx_sym": "mtx"}.get(fmt, fmt)
        path = str(tmp_path / f"g{trial}.{suffix}")
        _write(path, fmt.replace("_sym", ""), u, v, w, n, rng, sym=fmt == "mtx_sym")
        for sym in (False, True):
            want = O.ref_read_graph(path, sym)
            assert same(lib.read_graph_file(path, sym), want), (fmt, trial, sym)
            assert same(O.read_graph_file(path, sym), want), (fmt, trial, sym)


@pytest.mark.parametrize("weighted", [False, True])
def test_serialized_written_by_reference(lib, tmp_path, weighted):
    if O.ref() is None:
        pytest.skip("oracle/_ref (the reference's GAPBS reader) is not built")
    rng = np.random.default_rng(11 + weighted)
    n = 400
    u, v, w = _random_edges(rng, n, 2000)
    src = str(tmp_path / ("g.wel" if weighted else "g.el"))
    _write(src, "wel" if weighted else "el", u, v, w, n, rng)
    for sym in (False, True):
        out = str(tmp_path / f"g{int(sym)}.{'wsg' if weighted else 'sg'}")
        O.ref_write_serialized(src, sym, weighted, out)
        want = O.ref_read_graph(out)
        assert same(lib.read_graph_file(out), want)
        assert same(O.read_graph_file(out), want)
        assert want[3] == (not sym)


def test_graphit_bin_layout_and_round_trip(lib, tmp_path):
    off, nb, w, _ = expected("weighted.wel", True)
    path = str(tmp_path / "g.graphit_bin")
    lib.write_graphit_bin(path, off, nb, w)
    with open(path, "rb") as f:
        assert f.read() == O.graphit_bin_bytes(off, nb, w)  # graph.h:286-293 field order
    got = lib.read_graph_file(path)
    assert same(got, (off, nb, w, True))
    assert same(O.read_graph_file(path), (off, nb, w, True))
    # unweighted CSR: the cache stores weight 1, as load_graph's weighted graph has
    off2, nb2, _, _ = expected("dup_loops.sg", False)
    lib.write_graphit_bin(path, off2, nb2)
    assert np.array_equal(lib.read_graph_file(path)[2], np.ones(len(nb2), np.int32))


def test_errors_are_reported(lib, tmp_path):
    def bad(text, name):
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(lib.GIError):
            lib.read_graph_file(str(p))
    bad("0 1\n", "g.txt")                                      # unknown suffix
    bad("3 3 11\n2\n1\n1\n", "g.graph")                        # METIS fmt 11
    bad("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 2 1\n", "g.mtx")  # not square
    bad("%%MatrixMarket matrix array real general\n2 2\n", "g2.mtx")
    bad("0 1\n-1 2\n", "neg.el")                               # negative id
    with pytest.raises(lib.GIError):
        lib.read_graph_file(str(tmp_path / "missing.el"))


def test_empty_and_large_ids(lib, tmp_path):
    p = tmp_path / "e.el"
    p.write_text("")
    off, nb, w, d = lib.read_graph_file(str(p))
    assert off.tolist() == [0, 0] and len(nb) == 0  # one vertex, as GAPBS (largest id 0)
    p = tmp_path / "big.el"
    p.write_text("0 99999\n")
    off, nb, _, _ = lib.read_graph_file(str(p), True)
    assert len(off) == 100001 and nb.tolist() == [99999, 0]
    assert same(lib.read_graph_file(str(p), True), O.read_graph_file(str(p), True))


@pytest.mark.parametrize("suffix", ["sg", "wsg"])
def test_truncated_serialized_headers(lib, tmp_path, suffix):
    # a header whose V and E ask for far more bytes than the file holds is an
    # error before anything is allocated (not a multi-gigabyte allocation)
    import struct
    for V, E, body in ((1 << 30, 1 << 30, b""), (10, 1 << 31 - 1, b"\0" * 88), (3, 4, b"\0" * 20)):
        p = tmp_path / f"t{V}_{E}.{suffix}"
        p.write_bytes(struct.pack("<Bqq", 0, E, V) + body)
        with pytest.raises(lib.GIError):
            lib.read_graph_file(str(p))
