"""Graph files loaded onto the device (gi_graph_load): the device CSR equals
what the reference's GAPBS reader builds (the committed expectations), the
.graphit_bin cache next to a file is written and then read instead of the
file, and the apps run on a loaded graph."""
import os
import shutil

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "graphs")
TEXT = ["dup_loops.el", "comment_stops.el", "weighted.wel", "road.gr", "plain.graph",
        "weighted.graph", "pattern_sym.mtx", "int_general.mtx", "real_general.mtx"]


def expected(name, sym):
    d = np.load(os.path.join(GOLD, "expected.npz"))
    k = f"{name}|{int(sym)}"
    return d[k + "|offsets"], d[k + "|nbrs"], d[k + "|weights"]


@pytest.mark.parametrize("name", TEXT + ["weighted_sym.wsg"])
def test_load_matches_reference(gi, name):
    for sym in ((False,) if name.endswith("wsg") else (False, True)):
        g = gi.Graph.load(os.path.join(GOLD, name), symmetrize=sym)
        off, nb, w = g.read_csr(weights=True)
        eo, en, ew = expected(name, sym)
        assert np.array_equal(off, eo) and np.array_equal(nb, en) and np.array_equal(w, ew), (name, sym)
        g.destroy()


def test_load_serialized_unweighted(gi):
    g = gi.Graph.load(os.path.join(GOLD, "dup_loops.sg"))
    off, nb, _ = g.read_csr()
    eo, en, _ = expected("dup_loops.sg", False)
    assert np.array_equal(off, eo) and np.array_equal(nb, en)


def test_cache_written_then_used(gi, tmp_path):
    src = tmp_path / "w.wel"
    shutil.copy(os.path.join(GOLD, "weighted.wel"), src)
    g = gi.Graph.load(str(src), cache=True)
    cache = str(src) + ".i.graphit_bin"  # the reference's cache name (graph.h:239-248)
    assert os.path.exists(cache)
    eo, en, ew = expected("weighted.wel", False)
    with open(cache, "rb") as f:
        assert f.read() == O.graphit_bin_bytes(eo, en, ew)
    g.destroy()
    src.write_text("0 1 1\n")  # the cache, not the file, is read from now on
    g = gi.Graph.load(str(src), cache=True)
    off, nb, w = g.read_csr(weights=True)
    assert np.array_equal(off, eo) and np.array_equal(nb, en) and np.array_equal(w, ew)
    g.destroy()
    g = gi.Graph.load(cache)  # the cache file itself
    assert np.array_equal(g.read_csr()[1], en)
    g.destroy()
    gs = gi.Graph.load(str(src), symmetrize=True, cache=True)  # a cache of its own
    assert os.path.exists(str(src) + ".sym.i.graphit_bin")
    assert gs.read_csr()[1].tolist() == [1, 0]


def test_apps_on_a_loaded_graph(gi, tmp_path):
    # a DIMACS road-like file: grid plus random long edges, weights 1..1000
    rng = np.random.default_rng(5)
    w_, h_ = 60, 50
    lines = ["c grid", f"p sp {w_ * h_} 0"]
    for y in range(h_):
        for x in range(w_):
            v = y * w_ + x + 1
            if x + 1 < w_:
                lines.append(f"a {v} {v + 1} {rng.integers(1, 1001)}")
            if y + 1 < h_:
                lines.append(f"a {v} {v + w_} {rng.integers(1, 1001)}")
    for _ in range(40):
        a, b = rng.integers(1, w_ * h_ + 1, 2)
        lines.append(f"a {a} {b} {rng.integers(1, 1001)}")
    path = tmp_path / "road.gr"
    path.write_text("\n".join(lines) + "\n")
    g = gi.Graph.load(str(path), symmetrize=True)
    off, nb, w = g.read_csr(weights=True)
    host = O.CSR(off, nb, w)
    for src in (1, 1234):
        depth, _, _ = gi.bfs(g, src, None)
        assert np.array_equal(depth, O.bfs_depth(host, src))
        fused = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWC, kernel_fusion=1)
        dist, _ = gi.sssp(g, src, 64, fused)
        assert np.array_equal(dist, O.sssp(host, src))
    lab, _ = gi.cc(g, gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_EDGE_ONLY, kernel_fusion=1))
    assert np.array_equal(lab, O.cc(host))
