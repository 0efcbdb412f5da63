"""SSSP, CC, PageRank and BC on the device vs the oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095


def host_csr(g, weights=False):
    off, nb, w = g.read_csr(weights=weights)
    return O.CSR(off, nb, w)


@pytest.mark.parametrize("lb", [0, 1, 2, 3, 4, 5])
def test_sssp_grid_bit_exact(gi, lb):
    g = gi.Graph.grid(96, 92, 3).set_weights(5, 1, 1000)
    host = host_csr(g, True)
    want = O.sssp(host, 0)
    s = gi.Schedule(direction=gi.PUSH, load_balancing=lb)
    dist, st = gi.sssp(g, 0, 64, s)
    assert np.array_equal(dist, want)


def test_sssp_rmat_default_schedule(gi):
    g = gi.Graph.rmat(14, 16, SEED, True).set_weights(SEED, 1, 255)
    host = host_csr(g, True)
    for src in O.pick_sources(host, 1, 3):
        dist, _ = gi.sssp(g, int(src), 64)
        assert np.array_equal(dist, O.sssp(host, int(src)))


def _fused_sssp(gi):
    return gi.Schedule(direction=gi.PUSH, load_balancing=1, kernel_fusion=1)


@pytest.mark.parametrize("delta", [1, 16, 64, 5000])
def test_sssp_fused_rmat(gi, delta):
    g = gi.Graph.rmat(14, 16, SEED, True).set_weights(SEED, 1, 255)
    host = host_csr(g, True)
    for src in O.pick_sources(host, 1, 3):
        for _ in range(2):  # the second run replays the captured phases
            dist, st = gi.sssp(g, int(src), delta, _fused_sssp(gi))
            assert np.array_equal(dist, O.sssp(host, int(src))), (delta, src)


def test_sssp_fused_grid_and_unreachable(gi):
    g = gi.Graph.grid(96, 92, 3).set_weights(5, 1, 1000)
    host = host_csr(g, True)
    dist, _ = gi.sssp(g, 0, 64, _fused_sssp(gi))
    assert np.array_equal(dist, O.sssp(host, 0))
    # two components and an isolated vertex: unreached stay INT_MAX
    e = np.array([[0, 1], [1, 2], [3, 4]], np.int32).T
    g2 = gi.Graph.from_edges(6, e[0], e[1], True).set_weights(2, 1, 9)
    h2 = host_csr(g2, True)
    d2, _ = gi.sssp(g2, 0, 4, _fused_sssp(gi))
    assert np.array_equal(d2, O.sssp(h2, 0)) and d2[3] == np.iinfo(np.int32).max


def test_sssp_fused_chain_unit_bounds(gi):
    # a weighted chain with long edges: each window holds a few consecutive
    # vertices, so the window scan skips most 1024-vertex units; long edges
    # leave far vertices pending at distances that later improve (stale unit
    # bounds), and some long edges are shortcuts
    n = 12000
    rng = np.random.default_rng(5)
    a = np.arange(n - 1)
    la = rng.integers(0, n, 200)
    lb = rng.integers(0, n, 200)
    src = np.concatenate([a, la]).astype(np.int32)
    dst = np.concatenate([a + 1, lb]).astype(np.int32)
    w = np.concatenate([rng.integers(1, 11, n - 1), rng.integers(1, 30000, 200)]).astype(np.int32)
    # symmetric, no self-loops, one entry per neighbour with its least weight
    s2, d2, w2 = np.concatenate([src, dst]), np.concatenate([dst, src]), np.concatenate([w, w])
    keep = s2 != d2
    s2, d2, w2 = s2[keep], d2[keep], w2[keep]
    order = np.lexsort((w2, d2, s2))
    s2, d2, w2 = s2[order], d2[order], w2[order]
    first = np.ones(len(s2), bool)
    first[1:] = (s2[1:] != s2[:-1]) | (d2[1:] != d2[:-1])
    s2, nb, wt = s2[first], d2[first].astype(np.int32), w2[first].astype(np.int32)
    off = np.concatenate([[0], np.cumsum(np.bincount(s2, minlength=n))]).astype(np.int32)
    g = gi.Graph.from_csr(off, nb, wt)
    want = O.sssp(O.CSR(off, nb, wt), 0)
    for delta in (8, 64):
        for _ in range(2):
            d, _ = gi.sssp(g, 0, delta, _fused_sssp(gi))
            assert np.array_equal(d, want), delta


def test_sssp_fused_skips_empty_windows(gi):
    # distances 0, 3, 503, 505 with delta 4: after [0, 4) the next window
    # must start at 503, the least pending distance
    off = np.array([0, 1, 3, 5, 6], np.int32)
    nb = np.array([1, 0, 2, 1, 3, 2], np.int32)
    w = np.array([3, 3, 500, 500, 2, 2], np.int32)
    g = gi.Graph.from_csr(off, nb, w)
    assert O.sssp(host_csr(g, True), 0).tolist() == [0, 3, 503, 505]
    for _ in range(2):
        d, _ = gi.sssp(g, 0, 4, _fused_sssp(gi))
        assert d.tolist() == [0, 3, 503, 505]
    d, _, info = gi.sssp_trace(g, 0, 4)
    assert d.tolist() == [0, 3, 503, 505]
    ends = set(info[:, 3].tolist())
    assert 507 in ends and not any(8 < h < 503 for h in ends)


def test_sssp_trace_matches_and_counts(gi):
    g = gi.Graph.rmat(13, 16, SEED, True).set_weights(SEED, 1, 255)
    host = host_csr(g, True)
    src = int(O.pick_sources(host, 1, 4)[0])
    want = O.sssp(host, src)
    _, st = gi.sssp(g, src, 32, _fused_sssp(gi))
    for pull_min in (-1, 0, 1 << 30):  # default rule, pull whenever F is non-empty, push only
        d, ms, info = gi.sssp_trace(g, src, 32, pull_min)
        assert np.array_equal(d, want), pull_min
        assert (ms >= 0).all() and len(ms) == len(info)
        relax = info[:, 0] == 0
        assert info[relax, 2].min() > 0  # a relax slot has a non-empty F
        if pull_min == 0:
            assert info[relax, 1][1:].all()
        if pull_min == 1 << 30:
            assert not info[relax, 1].any()
            # wide windows deferred their heavy edges to a heavy step (mode 2)
            assert (info[:, 0] == 2).any()
    assert st["iterations"] == st["push_iterations"] + st["pull_iterations"]


def test_sssp_heavy_step_pushed_and_pulled(gi):
    # The heavy step relaxes W's heavy edges once per window: pushed, or, on a
    # symmetric graph whose W holds an eighth of the edges or more, pulled
    # over the split in-lists. Relax steps push here (pull_min huge), so wide
    # frontiers defer their heavy edges.
    seen = {}
    for symmetric in (True, False):
        g = gi.Graph.rmat(13, 16, SEED, symmetric).set_weights(SEED, 1, 255)
        host = host_csr(g, True)
        for src in O.pick_sources(host, 1, 3):
            want = O.sssp(host, int(src))
            d, _, info = gi.sssp_trace(g, int(src), 32, 1 << 30)
            assert np.array_equal(d, want), (symmetric, int(src))
            heavy = info[info[:, 0] == 2]
            seen.setdefault(symmetric, set()).update(heavy[:, 1].tolist())
            d2, _ = gi.sssp(g, int(src), 32, _fused_sssp(gi))
            assert np.array_equal(d2, want), (symmetric, int(src))
    assert 1 in seen[True]           # a symmetric graph's large W is pulled
    assert seen[False] == {0}        # a directed graph's heavy steps push


def test_set_weights_after_fused_runs_directed(gi):
    # new weights drop the transposed graph; the fused engines must not keep
    # pointers into it
    g = gi.Graph.rmat(12, 8, 4, False).set_weights(1, 1, 100)
    fs = _fused_sssp(gi)
    fp = gi.Schedule(direction=gi.PULL, load_balancing=gi.LB_VERTEX_BASED, kernel_fusion=1)
    s1 = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWCE, kernel_fusion=1)
    s2 = gi.Schedule(direction=gi.PULL, load_balancing=gi.LB_VERTEX_BASED, kernel_fusion=1,
                     pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP)
    hy = gi.HybridSchedule(s1, s2, 0.02, gi.SELECT_HEURISTIC)
    host = host_csr(g, True)
    src = int(O.pick_sources(host, 1, 9)[0])
    for seed in (1, 2):
        if seed == 2:
            g.set_weights(seed, 1, 100)
            host = host_csr(g, True)
        d, _ = gi.sssp(g, src, 16, fs)
        assert np.array_equal(d, O.sssp(host, src)), seed
        r, _ = gi.pagerank(g, 10, schedule=fp)
        want = O.pagerank(host, 10)
        assert np.abs(r - want).sum() / np.abs(want).sum() <= 1e-5
        depth, _, _ = gi.bfs(g, src, hy)
        assert np.array_equal(depth, O.bfs_depth(host, src))


def test_autotune_delta_keeps_distances(gi):
    from easy_graphit_amd import autotune
    g = gi.Graph.rmat(12, 16, SEED, True).set_weights(SEED, 1, 255)
    host = host_csr(g, True)
    src = int(O.pick_sources(host, 1, 2)[0])
    best = autotune.tune_delta(g, src, deltas=(4, 32, 256), reps=1)
    assert len(best["table"]) == 3 and best["ms"] == min(r["ms"] for r in best["table"])
    d, _ = gi.sssp(g, src, best["delta"], _fused_sssp(gi))
    assert np.array_equal(d, O.sssp(host, src))


def test_sssp_fused_hub_bin(gi):
    # a centre wider than the grid-bin threshold, with a tail of long edges
    n = 9003
    e = [(0, i) for i in range(1, 9001)] + [(9000, 9001), (9001, 9002), (5, 9002)]
    e = np.array(e, np.int32).T
    g = gi.Graph.from_edges(n, e[0], e[1], True).set_weights(7, 1, 200)
    host = host_csr(g, True)
    for src in (0, 9002):
        dist, _ = gi.sssp(g, src, 32, _fused_sssp(gi))
        assert np.array_equal(dist, O.sssp(host, src))


def test_sssp_fused_wide_and_narrow_share_a_word(gi):
    # A pull step where a wide vertex (in-degree above the chunked-pull
    # threshold) and narrow vertices of the same 64-vertex word improve
    # together; each has a private successor that only it reaches.
    n = 5400
    e = [(100, leaf) for leaf in range(200, 5200)]          # frontier of 5000 leaves
    e += [(leaf, 1) for leaf in range(200, 5200)]            # vertex 1: in-degree 5000
    e += [(200 + 3 * k, k) for k in range(2, 64)]            # narrow vertices 2..63
    e += [(k, 5300 + k) for k in range(1, 64)]               # private successors
    e = np.array(e, np.int32).T
    g = gi.Graph.from_edges(n, e[0], e[1], True).set_weights(11, 1, 50)
    host = host_csr(g, True)
    for delta in (8, 64):
        dist, _ = gi.sssp(g, 100, delta, _fused_sssp(gi))
        assert np.array_equal(dist, O.sssp(host, 100)), delta


@pytest.mark.parametrize("lb", [6, 0, 2, 4])
def test_cc_bit_exact(gi, lb):
    g = gi.Graph.uniform(20000, 15000, 4, True)
    host = host_csr(g)
    lab, st = gi.cc(g, gi.Schedule(direction=gi.PUSH, load_balancing=lb))
    assert np.array_equal(lab, O.cc(host))


def test_cc_rmat(gi):
    g = gi.Graph.rmat(15, 16, SEED, True)
    lab, _ = gi.cc(g)
    assert np.array_equal(lab, O.cc(host_csr(g)))


def rel_l1(a, b):
    return float(np.abs(a.astype(np.float64) - b).sum() / np.abs(b).sum())


def test_pagerank_default_schedule(gi):
    g = gi.Graph.rmat(14, 16, SEED, False)
    r, _ = gi.pagerank(g, 20)
    want = O.pagerank(host_csr(g), 20)
    assert rel_l1(r, want) <= 1e-5  # FP32 vs FP64 oracle, relative L1


@pytest.mark.parametrize("direction,lb", [(1, 0), (1, 2), (0, 6), (1, 6), (0, 0)])
def test_pagerank_schedules(gi, direction, lb):
    g = gi.Graph.rmat(12, 16, 5, False)
    r, _ = gi.pagerank(g, 20, schedule=gi.Schedule(direction=direction, load_balancing=lb))
    assert rel_l1(r, O.pagerank(host_csr(g), 20)) <= 1e-5


@pytest.mark.parametrize("lb", [6, 0])  # EDGE_ONLY sweeps; VERTEX_BASED sampled + filtered rounds
@pytest.mark.parametrize("graph", ["rmat", "grid", "forest"])
def test_cc_fused_matches_oracle(gi, graph, lb):
    if graph == "rmat":
        g = gi.Graph.rmat(15, 8, SEED, True)
    elif graph == "grid":
        g = gi.Graph.grid(300, 40, 3)
    else:  # many small components, isolated vertices, long paths
        n = 50000
        rng = np.random.default_rng(9)
        a = rng.integers(0, n, 30000)
        b = np.minimum(a + rng.integers(1, 3, 30000), n - 1)
        path = np.arange(20000, 29999)
        e = np.concatenate([np.stack([a, b]), np.stack([path, path + 1])], axis=1).astype(np.int32)
        g = gi.Graph.from_edges(n, e[0], e[1], True)
    host = host_csr(g)
    want = O.cc(host)
    sched = gi.Schedule(direction=0, load_balancing=lb, kernel_fusion=1)
    for _ in range(2):  # second run replays the captured rounds
        lab, st = gi.cc(g, schedule=sched)
        assert np.array_equal(lab, want)
        assert st["iterations"] >= 1


def _fused_pr(gi):
    return gi.Schedule(direction=1, load_balancing=0, kernel_fusion=1)


@pytest.mark.parametrize("directed", [False, True])
def test_pagerank_fused_matches_oracle(gi, directed):
    g = gi.Graph.rmat(14, 16, SEED, not directed)
    want = O.pagerank(host_csr(g), 20)
    r1, _ = gi.pagerank(g, 20, schedule=_fused_pr(gi))
    assert rel_l1(r1, want) <= 1e-5
    r2, _ = gi.pagerank(g, 20, schedule=_fused_pr(gi))  # captured graph replayed
    assert np.array_equal(r1, r2)  # fixed summation order: bit-identical


def test_pagerank_fused_degree_boundaries(gi):
    # in-degrees 63, 64 (first wide), 512, 513 and 1100 (chunk edges), a
    # vertex with no out-edges and one with no in-edges
    n = 1200
    src, dst = [], []
    for hub, deg in ((0, 63), (1, 64), (2, 512), (3, 513), (4, 1100)):
        for k in range(deg):
            src.append(10 + (k * 11 + hub) % (n - 10))  # distinct for k < n - 10
            dst.append(hub)
    src += list(range(10, n - 1))
    dst += list(range(11, n))
    e = np.array([src, dst], np.int32)
    g = gi.Graph.from_edges(n, e[0], e[1], False)
    host = O.build_csr(n, e[0], e[1], False)
    for iters, d in ((0, 0.85), (1, 0.85), (20, 0.85), (7, 0.5)):
        r, _ = gi.pagerank(g, iters, damping=d, schedule=_fused_pr(gi))
        assert rel_l1(r, O.pagerank(host, iters, d)) <= 1e-5, (iters, d)


def test_bc_small(gi):
    g = gi.Graph.rmat(12, 8, SEED, True)
    host = host_csr(g)
    srcs = O.pick_sources(host, 7, 4)
    sc, _ = gi.bc(g, srcs)
    want = O.bc(host, srcs)
    assert rel_l1(sc, want) <= 1e-4


@pytest.mark.parametrize("batch", [2, 64, 256])
def test_bc_batched_matches_oracle(gi, batch):
    g = gi.Graph.rmat(12, 8, SEED, True)
    host = host_csr(g)
    srcs = O.pick_sources(host, 11, 40)
    srcs[5] = srcs[3]  # a repeated source counts twice, as two runs would
    sc, st = gi.bc(g, srcs, batch=batch)
    assert rel_l1(sc, O.bc(host, srcs)) <= 1e-4


def test_bc_batched_uneven_passes(gi):
    # 70 sources in passes of 64 and 6: the second pass records its own slots
    g = gi.Graph.rmat(12, 8, SEED, True)
    host = host_csr(g)
    srcs = O.pick_sources(host, 17, 70)
    sc, _ = gi.bc(g, srcs, batch=64)
    assert rel_l1(sc, O.bc(host, srcs)) <= 1e-4
    sc2, _ = gi.bc(g, srcs[:6], batch=64)  # a fresh call with the short pass alone
    assert rel_l1(sc2, O.bc(host, srcs[:6])) <= 1e-4


@pytest.mark.parametrize("graph", ["grid", "directed"])
def test_bc_batched_long_and_directed(gi, graph):
    # a grid: many levels, pushed and pulled in both passes; a directed graph:
    # push only (the pulled levels need the symmetric graph's in-neighbours)
    if graph == "grid":
        g = gi.Graph.grid(48, 30, 3)
    else:
        g = gi.Graph.rmat(11, 8, 5, False)
    host = host_csr(g)
    srcs = O.pick_sources(host, 13, 32)
    for batch in (1, 32):
        sc, _ = gi.bc(g, srcs, batch=batch)
        assert rel_l1(sc, O.bc(host, srcs)) <= 1e-4, (graph, batch)


@pytest.mark.parametrize("graph", ["rmat", "grid", "directed"])
def test_bc_fused_single_source(gi, graph):
    # the fused mode: one source at a time with the level loops on the device
    if graph == "rmat":
        g = gi.Graph.rmat(12, 8, SEED, True)
    elif graph == "grid":
        g = gi.Graph.grid(48, 30, 3)
    else:
        g = gi.Graph.rmat(11, 8, 5, False)
    host = host_csr(g)
    srcs = O.pick_sources(host, 9, 12)
    want = O.bc(host, srcs)
    fused = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWCE, kernel_fusion=1)
    sc, st = gi.bc(g, srcs, schedule=fused)
    assert rel_l1(sc, want) <= 1e-4, graph
    assert st["iterations"] == len(srcs)


@pytest.mark.parametrize("lb", [0, 1, 2])
def test_bc_schedule_unfused(gi, lb):
    # unfused, one source at a time with the schedule's push load balancing
    g = gi.Graph.rmat(11, 8, SEED, True)
    host = host_csr(g)
    srcs = O.pick_sources(host, 5, 8)
    sc, _ = gi.bc(g, srcs, schedule=gi.Schedule(direction=gi.PUSH, load_balancing=lb))
    assert rel_l1(sc, O.bc(host, srcs)) <= 1e-4, lb
    with pytest.raises(gi.GIError):  # pulled levels need the fused mode
        gi.bc(g, srcs, schedule=gi.Schedule(direction=gi.PULL, load_balancing=lb))


def test_bc_batched_equals_sequential_on_fixture(gi):
    e = np.array([[0, 1], [1, 2], [2, 3], [0, 4], [4, 3], [3, 5]], np.int32)
    g = gi.Graph.from_edges(7, e[:, 0], e[:, 1], True)  # vertex 6 isolated
    srcs = np.array([0, 3, 6, 5], np.int32)
    host = host_csr(g)
    want = O.bc(host, srcs)
    seq, _ = gi.bc(g, srcs, batch=1)
    bat, _ = gi.bc(g, srcs, batch=4)
    fus, _ = gi.bc(g, srcs, schedule=gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWCE, kernel_fusion=1))
    for got in (seq, bat, fus):
        assert np.allclose(got, want, rtol=1e-5)


def test_sssp_fused_tiny_graphs(gi):
    # no edges; one edge that is light (delta 2) or heavy (delta 1); a source
    # whose only edge is heavy and wide windows do not appear
    g0 = gi.Graph.from_csr(np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    d, _ = gi.sssp(g0, 0, 4, _fused_sssp(gi))
    assert d.tolist() == [0]
    off = np.array([0, 1, 2], np.int32)
    nb = np.array([1, 0], np.int32)
    w = np.array([1, 1], np.int32)
    g1 = gi.Graph.from_csr(off, nb, w)
    for delta in (1, 2, 1000):
        for src in (0, 1):
            d, _ = gi.sssp(g1, src, delta, _fused_sssp(gi))
            assert d.tolist() == ([0, 1] if src == 0 else [1, 0]), (delta, src)


def test_bc_tiny_and_empty(gi):
    fused = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWCE, kernel_fusion=1)
    g0 = gi.Graph.from_edges(1, np.zeros(0, np.int32), np.zeros(0, np.int32), True)
    for kw in ({"batch": 4}, {"schedule": fused}, {}):
        sc, _ = gi.bc(g0, [0, 0], **kw)
        assert sc.tolist() == [0.0], kw
        sc, st = gi.bc(g0, np.zeros(0, np.int32), **kw)
        assert sc.tolist() == [0.0] and st["iterations"] == 0, kw
    # a path 0-1-2: vertex 1 lies on the 0..2 paths
    e = np.array([[0, 1], [1, 2]], np.int32)
    g = gi.Graph.from_edges(3, e[:, 0], e[:, 1], True)
    want = O.bc(host_csr(g), np.array([0, 1, 2], np.int32))
    for kw in ({"batch": 3}, {"schedule": fused}, {}):
        sc, _ = gi.bc(g, [0, 1, 2], **kw)
        assert np.allclose(sc, want, rtol=1e-6), kw


def _tiny_graphs(gi):
    """No edges, one edge, a full last bitmap word (path of 64), a ragged
    one (star of 65), and two words apart joined by one edge."""
    z = np.zeros(0, np.int32)
    out = [gi.Graph.from_edges(1, z, z, True)]
    out.append(gi.Graph.from_edges(2, np.array([0], np.int32), np.array([1], np.int32), True))
    p = np.arange(63, dtype=np.int32)
    out.append(gi.Graph.from_edges(64, p, p + 1, True))
    out.append(gi.Graph.from_edges(65, np.full(64, 64, np.int32), np.arange(64, dtype=np.int32), True))
    out.append(gi.Graph.from_edges(130, np.array([0], np.int32), np.array([129], np.int32), True))
    return out


@pytest.mark.parametrize("mode_sched", [(0, 0), (0, 3), (0, 2), (0, 4), (0, 1), (1, 0), (1, 2)])
def test_fused_bfs_tiny_graphs(gi, mode_sched):
    d, lb = mode_sched  # one schedule per fused step mode
    s = gi.Schedule(direction=d, load_balancing=lb, kernel_fusion=1,
                    pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP)
    h = gi.HybridSchedule(s, s, selector=gi.SELECT_S1)
    for g in _tiny_graphs(gi):
        host = host_csr(g)
        for src in sorted({0, g.num_vertices - 1}):
            depth, _, _ = gi.bfs(g, src, h)
            assert np.array_equal(depth, O.bfs_depth(host, src)), (g.num_vertices, src)


def test_fused_cc_and_pagerank_tiny_graphs(gi):
    for g in _tiny_graphs(gi):
        host = host_csr(g)
        for lb in (gi.LB_EDGE_ONLY, gi.LB_VERTEX_BASED):
            lab, _ = gi.cc(g, schedule=gi.Schedule(direction=gi.PUSH, load_balancing=lb, kernel_fusion=1))
            assert np.array_equal(lab, O.cc(host)), (g.num_vertices, lb)
        r, _ = gi.pagerank(g, 20, schedule=_fused_pr(gi))
        assert rel_l1(r, O.pagerank(host, 20)) <= 1e-5, g.num_vertices
