"""Parity at the benchmark workloads' full sizes (the HIP path vs the oracle).

Each test builds its workload graph on the device, reads the CSR back and
runs the oracle on that same CSR (graph construction itself is pinned
bit-exact by test_gpu_graph.py at smaller scales):

* PageRank, RMAT-22 directed, 20 iterations, d = 0.85: FP32 on the device
  vs the FP64 oracle, relative L1 <= 1e-5 (default and fused schedules);
* delta-stepping SSSP, delta 64, on a 2048 x 2048 grid plus 1 % random long
  edges, weights U[1, 1000], source at the corner: bit-exact (fused, and the
  app's unfused TWCE loop over the bucketed priority queue);
* RMAT-24 symmetrized: direction-optimizing BFS under the learned selector
  and under GraphIt's fixed hybrid rule, bit-exact depths; CC with EDGE_ONLY
  and VERTEX_BASED fused schedules, bit-exact labels; SSSP with weights
  U[1, 255], bit-exact distances;
* batched Brandes BC from 256 sources on RMAT-20: relative L1 <= 1e-4 vs FP64;
  1,024 uniform graphs BFS'd in one launch with autotuned bins: bit-exact.
"""
import os
import sys
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def host_csr(g, weights=False):
    off, nb, w = g.read_csr(weights=weights)
    return O.CSR(off, nb, w)


def rel_l1(a, b):
    return float(np.abs(a.astype(np.float64) - b).sum() / np.abs(b).sum())


# ---------------------------------------------------------------- PageRank
@pytest.fixture(scope="module")
def pr22(gi):
    g = gi.Graph.rmat(22, 16, SEED, False)
    t0 = time.perf_counter()
    want = O.pagerank(host_csr(g), 20)
    say(f"PR RMAT-22 oracle {time.perf_counter() - t0:.1f}s E={g.num_edges}")
    yield g, want
    g.destroy()


@pytest.mark.parametrize("kind", ["default", "fused", "fused_degree_layout"])
def test_pagerank_rmat22_directed(gi, pr22, kind):
    g, want = pr22
    if kind == "default":  # the app's PULL EDGE_ONLY BLOCKED 0x42000 (apps/pr.cpp:78-81)
        sched = None
    else:
        sched = gi.Schedule(direction=gi.PULL, load_balancing=gi.LB_VERTEX_BASED, kernel_fusion=1)
    if kind == "fused_degree_layout":
        g.set_layout(gi.LAYOUT_DEGREE)
    try:
        r, st = gi.pagerank(g, 20, schedule=sched)
    finally:
        g.set_layout(gi.LAYOUT_NATIVE)
    err = rel_l1(r, want)
    say(f"PR {kind}: rel L1 {err:.3g}, {st['ms']:.2f} ms")
    assert err <= 1e-5


# ---------------------------------------------------------------- SSSP grid
@pytest.fixture(scope="module")
def grid2048(gi):
    g = gi.Graph.grid(2048, 2 * 2048 * 2047 // 100, SEED).set_weights(SEED, 1, 1000)
    t0 = time.perf_counter()
    want = O.sssp(host_csr(g, True), 0)
    say(f"SSSP grid oracle {time.perf_counter() - t0:.1f}s E={g.num_edges}")
    yield g, want
    g.destroy()


@pytest.mark.parametrize("fused", [True, False])
def test_sssp_grid2048_delta64(gi, grid2048, fused):
    g, want = grid2048
    if fused:
        s = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWC, kernel_fusion=1)
    else:  # the app's schedule: TWCE push, sparse frontier, host loop over the queue
        s = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWCE)
    for rep in range(2 if fused else 1):  # the second fused run replays recorded slots
        dist, st = gi.sssp(g, 0, 64, s)
        assert np.array_equal(dist, want), (fused, rep)
    say(f"SSSP grid fused={fused}: {st['ms']:.2f} ms, {st['iterations']} steps")


# ---------------------------------------------------------------- RMAT-24
@pytest.fixture(scope="module")
def rmat24(gi):
    g = gi.Graph.rmat(24, 16, SEED, True)
    host = host_csr(g)
    srcs = O.pick_sources(host, SEED, 3)
    t0 = time.perf_counter()
    depths = {int(s): O.bfs_depth(host, int(s)) for s in srcs}
    say(f"RMAT-24 BFS oracle {time.perf_counter() - t0:.1f}s for {len(srcs)} sources")
    yield g, host, depths
    g.destroy()


def _bfs_schedule(gi, selector, threshold):
    s1 = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWCE, frontier_creation=gi.FC_SPARSE,
                     cta_size=256, kernel_fusion=1)
    s2 = gi.Schedule(direction=gi.PULL, load_balancing=gi.LB_VERTEX_BASED,
                     pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP,
                     cta_size=256, kernel_fusion=1)
    return gi.HybridSchedule(s1, s2, threshold, selector)


@pytest.mark.parametrize("layout", ["degree", "native"])
def test_bfs_rmat24_learned_and_heuristic(gi, rmat24, layout):
    from easy_graphit_amd import selector as sel
    g, host, depths = rmat24
    model = sel.load()
    sel.install(model)
    g.set_layout(gi.LAYOUT_DEGREE if layout == "degree" else gi.LAYOUT_NATIVE)
    try:
        t = float(model.get("threshold", 0.0005))
        for name, h in (("learned", _bfs_schedule(gi, gi.SELECT_LEARNED, t)),
                        ("heuristic", _bfs_schedule(gi, gi.SELECT_HEURISTIC, t))):
            for s, want in depths.items():
                depth, _, st = gi.bfs(g, s, h)
                assert np.array_equal(depth, want), (name, layout, s)
            say(f"BFS RMAT-24 {layout} {name}: {st['iterations']} levels "
                f"({st['pull_iterations']} pulled) bit-exact")
    finally:
        g.set_layout(gi.LAYOUT_NATIVE)
        gi.selector_clear()


def test_bfs_rmat24_parents_valid(gi, rmat24):
    g, host, depths = rmat24
    s, want = next(iter(depths.items()))
    g.set_layout(gi.LAYOUT_DEGREE)
    try:
        depth, parent, _ = gi.bfs(g, s, _bfs_schedule(gi, gi.SELECT_HEURISTIC, 0.0005), want_parent=True)
    finally:
        g.set_layout(gi.LAYOUT_NATIVE)
    assert np.array_equal(depth, want)
    assert O.bfs_parent_violations(host, s, depth, parent) == 0


def test_cc_rmat24_edge_only_and_vertex_based(gi, rmat24):
    g, host, _ = rmat24
    t0 = time.perf_counter()
    want = O.cc(host)
    say(f"RMAT-24 CC oracle {time.perf_counter() - t0:.1f}s")
    g.set_layout(gi.LAYOUT_DEGREE)
    try:
        for lb in (gi.LB_EDGE_ONLY, gi.LB_VERTEX_BASED):
            lab, st = gi.cc(g, gi.Schedule(direction=gi.PUSH, load_balancing=lb, kernel_fusion=1))
            assert np.array_equal(lab, want), lb
            say(f"CC RMAT-24 {gi.LB_NAMES[lb]}: {st['ms']:.2f} ms bit-exact")
    finally:
        g.set_layout(gi.LAYOUT_NATIVE)


def test_sssp_rmat24(gi, rmat24):
    g, _, depths = rmat24
    g.set_weights(SEED, 1, 255)
    host = host_csr(g, True)
    s = next(iter(depths))
    t0 = time.perf_counter()
    want = O.sssp(host, s)
    say(f"RMAT-24 SSSP oracle {time.perf_counter() - t0:.1f}s")
    fs = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWC, kernel_fusion=1)
    for delta in (16, 64):
        dist, st = gi.sssp(g, s, delta, fs)
        assert np.array_equal(dist, want), delta
        say(f"SSSP RMAT-24 delta {delta}: {st['ms']:.2f} ms bit-exact")


# ---------------------------------------------------------------- BC + many graphs
def test_bc_rmat20_256_sources_batched(gi):
    g = gi.Graph.rmat(20, 16, SEED, True)
    host = host_csr(g)
    srcs = g.pick_sources(SEED, 256)
    t0 = time.perf_counter()
    want = O.bc(host, srcs)
    say(f"BC oracle {time.perf_counter() - t0:.1f}s on {O.set_threads(0)} threads")
    sc, st = gi.bc(g, srcs, batch=256)
    err = rel_l1(sc, want)
    big = want > 1e-3 * want.max()
    worst = float(np.max(np.abs(sc[big] - want[big]) / want[big]))
    say(f"BC 256 sources: rel L1 {err:.3g}, worst relative error over scores above 1e-3 of the "
        f"largest {worst:.3g}, {st['ms']:.1f} ms")
    assert err <= 1e-4
    assert worst <= 1e-4
    g.destroy()


def test_many_graphs_1024_autotuned(gi):
    from easy_graphit_amd import autotune
    g, offs = gi.Graph.uniform_batch(1024, 1024, 65536, 8, SEED)
    host = host_csr(g)
    deg = np.diff(host.offsets)
    srcs = []
    for i in range(1024):  # the first vertex with edges in each graph
        nz = np.flatnonzero(deg[offs[i]:offs[i + 1]])
        srcs.append(int(offs[i] + (nz[0] if len(nz) else 0)))
    srcs = np.array(srcs, np.int32)
    tuned = autotune.tune_bins(g, srcs, lane_max=(32, 64), hub_min=(512, 2048), blocks_per_cu=(8,),
                               push_lb=(gi.LB_TWC, gi.LB_TWCE), reps=1)
    want = O.bfs_multi_depth(host, srcs)
    depth, levels = gi.bfs_multi(g, srcs, autotune.schedule(tuned["push_lb"]))
    assert np.array_equal(depth, want)
    assert levels == want.max() + 1
    # each graph's depths are its own BFS: no vertex reached across graphs
    for i in (0, 511, 1023):
        part = depth[offs[i]:offs[i + 1]]
        sub = O.CSR(host.offsets[offs[i]:offs[i + 1] + 1] - host.offsets[offs[i]],
                    host.nbrs[host.offsets[offs[i]]:host.offsets[offs[i + 1]]] - offs[i])
        assert np.array_equal(part, O.bfs_depth(sub, int(srcs[i] - offs[i]))), i
    say(f"1,024 graphs: bins {tuned['lane_max']}/{tuned['hub_min']}/{tuned['blocks_per_cu']} "
        f"{gi.LB_NAMES[tuned['push_lb']]}, {tuned['seconds'] * 1e3:.2f} ms, bit-exact")
    g.destroy()
