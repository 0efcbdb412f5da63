"""Degree-ordered layout: fused BFS (single- and multi-source, parents, the
learned selector), CC, PageRank and SSSP run on the renumbered copy and must
give the oracle's results in the original vertex ids."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095


def fused(gi, push_lb=1, t=0.005, selector=0):
    s1 = gi.Schedule(direction=0, load_balancing=push_lb, kernel_fusion=1)
    s2 = gi.Schedule(direction=1, load_balancing=0, kernel_fusion=1,
                     pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP)
    return gi.HybridSchedule(s1, s2, t, selector)


@pytest.fixture(scope="module")
def rmat(gi):
    g = gi.Graph.rmat(15, 16, SEED, True)
    off, nb, _ = g.read_csr()
    host = O.CSR(off, nb)
    g.set_layout(gi.LAYOUT_DEGREE)
    return g, host


def test_layout_flag(gi, rmat):
    g, _ = rmat
    assert g.layout == gi.LAYOUT_DEGREE


def test_layout_bfs_depths_and_parents(gi, rmat):
    g, host = rmat
    for src in O.pick_sources(host, 3, 5):
        for push_lb in (1, 2):
            depth, parent, st = gi.bfs(g, int(src), fused(gi, push_lb), want_parent=True)
            assert np.array_equal(depth, O.bfs_depth(host, int(src)))
            assert O.bfs_parent_violations(host, int(src), depth, parent) == 0


def test_layout_bfs_multi(gi, rmat):
    g, host = rmat
    srcs = O.pick_sources(host, 8, 11)
    depth, levels = gi.bfs_multi(g, srcs, fused(gi))
    want = O.bfs_multi_depth(host, srcs)
    assert np.array_equal(depth, want) and levels == want.max() + 1


def test_layout_cc(gi):
    n = 40000
    rng = np.random.default_rng(3)
    a = rng.integers(0, n, 25000)
    b = rng.integers(0, n, 25000)
    e = np.stack([a, b]).astype(np.int32)
    g = gi.Graph.from_edges(n, e[0], e[1], True)
    host = O.build_csr(n, e[0], e[1], True)
    g.set_layout(gi.LAYOUT_DEGREE)
    for lb in (6, 0):
        sched = gi.Schedule(direction=0, load_balancing=lb, kernel_fusion=1)
        for _ in range(2):
            lab, _ = gi.cc(g, schedule=sched)
            assert np.array_equal(lab, O.cc(host)), lb


def test_layout_directed_bfs(gi):
    g = gi.Graph.rmat(13, 8, 4, False)
    off, nb, _ = g.read_csr()
    host = O.CSR(off, nb)
    g.set_layout(gi.LAYOUT_DEGREE)
    for src in O.pick_sources(host, 2, 3):
        depth, _, _ = gi.bfs(g, int(src), fused(gi, t=0.02))
        assert np.array_equal(depth, O.bfs_depth(host, int(src)))


def test_layout_back_to_native(gi):
    g = gi.Graph.rmat(12, 8, 9, True)
    off, nb, _ = g.read_csr()
    host = O.CSR(off, nb)
    g.set_layout(gi.LAYOUT_DEGREE)
    g.set_layout(gi.LAYOUT_NATIVE)
    assert g.layout == gi.LAYOUT_NATIVE
    depth, _, _ = gi.bfs(g, 0, fused(gi))
    assert np.array_equal(depth, O.bfs_depth(host, 0))


def test_layout_sweep_records(gi, rmat):
    g, host = rmat
    src = int(O.pick_sources(host, 4, 1)[0])
    feats, times = gi.bfs_sweep(g, src, reps=1)
    assert feats.shape[0] == O.bfs_depth(host, src).max() + 1
    assert (times > 0).all()


def test_layout_pagerank(gi):
    g = gi.Graph.rmat(14, 16, SEED, False)
    off, nb, _ = g.read_csr()
    want = O.pagerank(O.CSR(off, nb), 20)
    sched = gi.Schedule(direction=1, load_balancing=0, kernel_fusion=1)
    native, _ = gi.pagerank(g, 20, schedule=sched)
    g.set_layout(gi.LAYOUT_DEGREE)
    r, _ = gi.pagerank(g, 20, schedule=sched)
    rel = np.abs(r - want).sum() / np.abs(want).sum()
    assert rel <= 1e-5
    assert np.abs(r - native).sum() / np.abs(native).sum() <= 1e-6  # only the summation order differs


def test_layout_sssp(gi):
    g = gi.Graph.rmat(14, 16, SEED, True)
    g.set_layout(gi.LAYOUT_DEGREE)
    g.set_weights(SEED, 1, 255)  # after the layout: the renumbered copy takes the weights
    sched = gi.Schedule(direction=0, load_balancing=1, kernel_fusion=1)
    off, nb, w = g.read_csr(weights=True)
    host = O.CSR(off, nb, w)
    srcs = [int(x) for x in O.pick_sources(host, 2, 7)]
    for src in srcs:
        for delta in (8, 64):
            d, _ = gi.sssp(g, src, delta, sched)
            assert np.array_equal(d, O.sssp(host, src)), (src, delta)
    g.set_weights(SEED + 1, 1, 9)  # new weights reach the renumbered copy too
    off, nb, w = g.read_csr(weights=True)
    host = O.CSR(off, nb, w)
    want = O.sssp(host, srcs[0])
    d, _ = gi.sssp(g, srcs[0], 4, sched)
    assert np.array_equal(d, want)
    d, _, _ = gi.sssp_trace(g, srcs[0], 4)
    assert np.array_equal(d, want)


def test_layout_bins_reach_the_view(gi, rmat):
    # bins set on the graph apply where the fused engine runs (the view);
    # a tiny grid-bin threshold sends most frontier vertices to the grid bin
    g, host = rmat
    srcs = O.pick_sources(host, 3, 13)
    for bins in [(16, 64, 4), (64, 512, 8)]:
        g.set_bins(*bins)
        for src in srcs:
            depth, parent, _ = gi.bfs(g, int(src), fused(gi, 4), want_parent=True)
            assert np.array_equal(depth, O.bfs_depth(host, int(src))), bins
            assert O.bfs_parent_violations(host, int(src), depth, parent) == 0


def test_layout_bfs_isolated_tail(gi, rmat):
    # edgeless vertices sit at the end of the view; their depths come back
    # as -1 without being read, except for a source among them
    g, host = rmat
    iso = np.flatnonzero(np.diff(host.offsets) == 0)
    assert len(iso) > 0
    for src in (int(iso[0]), int(iso[-1]), int(O.pick_sources(host, 1, 17)[0])):
        depth, _, _ = gi.bfs(g, src, fused(gi))
        assert np.array_equal(depth, O.bfs_depth(host, src)), src


def test_layout_bfs_deep_and_shallow(gi):
    # a 700-vertex path hung off a random graph, with isolated vertices: deep
    # runs (more than 255 levels) from the path's far end and shallow ones
    # from the random part map back alike
    n_path, n_rand = 700, 3000
    rng = np.random.default_rng(8)
    a = rng.integers(0, n_rand, 12000)
    b = rng.integers(0, n_rand, 12000)
    pa = np.arange(n_rand - 1, n_rand + n_path - 1)
    src = np.concatenate([a, pa]).astype(np.int32)
    dst = np.concatenate([b, pa + 1]).astype(np.int32)
    n = n_rand + n_path + 50  # 50 isolated vertices
    g = gi.Graph.from_edges(n, src, dst, True)
    host = O.build_csr(n, src, dst, True)
    g.set_layout(gi.LAYOUT_DEGREE)
    for s0 in (n_rand + n_path - 1, 0, n - 1):
        for _ in range(2):
            depth, _, _ = gi.bfs(g, int(s0), fused(gi))
            want = O.bfs_depth(host, int(s0))
            assert np.array_equal(depth, want), s0
    assert O.bfs_depth(host, n_rand + n_path - 1).max() > 255


def test_layout_bfs_runs_leave_no_stale_depths(gi):
    # two components of different sizes: the view's depths are not cleared
    # between runs, so a run from one component must report the other as
    # unreached although an earlier run wrote depths there
    rng = np.random.default_rng(21)
    na, nb = 5000, 3000
    a = rng.integers(0, na, 20000)
    b = rng.integers(0, na, 20000)
    c = na + rng.integers(0, nb, 12000)
    d = na + rng.integers(0, nb, 12000)
    src = np.concatenate([a, c]).astype(np.int32)
    dst = np.concatenate([b, d]).astype(np.int32)
    n = na + nb + 50  # and 50 isolated vertices
    g = gi.Graph.from_edges(n, src, dst, True)
    host = O.build_csr(n, src, dst, True)
    g.set_layout(gi.LAYOUT_DEGREE)
    sa, sb = int(a[0]), int(c[0])
    for s in (sa, sb, sa, n - 1, sb):
        depth, _, _ = gi.bfs(g, s, fused(gi))
        want = O.bfs_depth(host, s)
        assert np.array_equal(depth, want), s
        assert (depth[:na] >= 0).any() == (s < na) and (depth[na:na + nb] >= 0).any() == (na <= s < na + nb)
