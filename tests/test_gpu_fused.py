"""Fused BFS engine (device-side schedule decisions) and the learned
selector: every step mode, the hybrid rule and the MLP give depths
bit-exact with the oracle; the selector's device gradients match the FP64
oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095

# fused mode -> (direction, load balancing) that selects it
MODE_SCHED = {0: (0, 0), 1: (0, 3), 2: (0, 2), 3: (0, 4), 4: (0, 1), 5: (1, 0), 6: (1, 2)}


def fused(gi, mode):
    d, lb = MODE_SCHED[mode]
    return gi.Schedule(direction=d, load_balancing=lb, kernel_fusion=1,
                       pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP)


@pytest.fixture(scope="module")
def c1(gi):
    g = gi.Graph.rmat(16, 16, SEED, True)
    off, nb, _ = g.read_csr()
    host = O.CSR(off, nb)
    return g, host, O.bfs_depth(host, 0)


@pytest.mark.parametrize("mode", range(7))
def test_every_fused_mode_bit_exact(gi, c1, mode):
    g, host, want = c1
    s = fused(gi, mode)
    depth, parent, st = gi.bfs(g, 0, gi.HybridSchedule(s, s, selector=gi.SELECT_S1),
                               want_parent=True)
    assert np.array_equal(depth, want)
    assert O.bfs_parent_violations(host, 0, depth, parent) == 0
    assert st["iterations"] == want.max() + 1


@pytest.mark.parametrize("t", [0.0, 0.01, 0.1, 2.0])
def test_fused_hybrid_rule(gi, c1, t):
    g, host, want = c1
    for src in [0] + list(O.pick_sources(host, 9, 3)):
        h = gi.HybridSchedule(fused(gi, 4), fused(gi, 5), t, gi.SELECT_HEURISTIC)
        depth, _, _ = gi.bfs(g, int(src), h)
        assert np.array_equal(depth, O.bfs_depth(host, int(src)))


def test_fused_hub_bin_on_wide_star(gi):
    # centre 0 with 20000 leaves and a tail: the centre exceeds the hub degree
    n = 20003
    e = [(0, i) for i in range(1, 20001)] + [(20000, 20001), (20001, 20002)]
    e = np.array(e, np.int32)
    g = gi.Graph.from_edges(n, e[:, 0], e[:, 1], True)
    host = O.build_csr(n, e[:, 0], e[:, 1], True)
    for mode in (0, 1, 2, 3, 4):
        s = fused(gi, mode)
        depth, _, _ = gi.bfs(g, 0, gi.HybridSchedule(s, s, selector=gi.SELECT_S1))
        assert np.array_equal(depth, O.bfs_depth(host, 0)), mode


def test_fused_grid_graph_many_levels(gi):
    g = gi.Graph.grid(200, 0, 1)
    off, nb, _ = g.read_csr()
    host = O.CSR(off, nb)
    h = gi.HybridSchedule(fused(gi, 4), fused(gi, 5), 0.05, gi.SELECT_HEURISTIC)
    depth, _, st = gi.bfs(g, 0, h)
    assert np.array_equal(depth, O.bfs_depth(host, 0)) and st["iterations"] == 399


@pytest.mark.parametrize("layout", [0, 1])
def test_fused_bfs_deeper_than_level_history(gi, layout):
    # no depth limit (the reference's fused loop has none, apps/bfs.cpp:53-57):
    # a path of 9,000 vertices with a side branch every 1,000 is 9,000 levels
    # deep, past the 4,096 levels whose per-level history is kept; the
    # statistics still count every level
    n = 9000
    a = np.arange(n - 1, dtype=np.int32)
    side = np.arange(0, n, 1000, dtype=np.int32)
    src = np.concatenate([a, side]).astype(np.int32)
    dst = np.concatenate([a + 1, n + np.arange(len(side))]).astype(np.int32)
    V = n + len(side)
    g = gi.Graph.from_edges(V, src, dst, True).set_layout(layout)
    host = O.build_csr(V, src, dst, True)
    for mode_pair in ((4, 5), (0, 6)):
        h = gi.HybridSchedule(fused(gi, mode_pair[0]), fused(gi, mode_pair[1]), 0.05, gi.SELECT_HEURISTIC)
        for s in (0, n - 1):
            want = O.bfs_depth(host, s)
            depth, _, st = gi.bfs(g, s, h)
            assert np.array_equal(depth, want), (mode_pair, s)
            assert st["iterations"] == want.max() + 1
            assert st["push_iterations"] + st["pull_iterations"] == st["iterations"]
    recs = gi.bfs_profile(g, 0, h, gi.DeviceArray(V, np.int32))
    assert len(recs) == 4096  # per-level records: the first 4,096 levels


def test_sweep_records_every_mode(gi, c1):
    g, host, want = c1
    feats, times = gi.bfs_sweep(g, 0, reps=1)
    assert feats.shape[0] == want.max() + 1
    assert (times > 0).all()
    assert np.allclose(feats[:, 15], 1.0)  # bias feature


def _init_params(rng, n_out):
    p = [rng.normal(size=(16, 64)) * 0.25, np.zeros(64), rng.normal(size=(64, 64)) * 0.125,
         np.zeros(64), rng.normal(size=(64, n_out)) * 0.125, np.zeros(n_out)]
    return p


def _flat(p):
    return np.concatenate([a.reshape(-1) for a in p]).astype(np.float32)


def _unflat(v, n_out):
    shapes = [(16, 64), (64,), (64, 64), (64,), (64, n_out), (n_out,)]
    out, i = [], 0
    for s in shapes:
        k = int(np.prod(s))
        out.append(v[i:i + k].reshape(s).astype(np.float64))
        i += k
    return out


def test_selector_gradients_match_fp64_oracle(gi):
    rng = np.random.default_rng(3)
    n_out = 7
    p = _init_params(rng, n_out)
    p[1] = rng.normal(size=64) * 0.1
    p[3] = rng.normal(size=64) * 0.1
    X = rng.normal(size=(300, 16))
    T = rng.normal(size=(300, n_out))
    M = (rng.random((300, n_out)) > 0.2) * rng.uniform(0.1, 4.0, size=(300, n_out))
    loss, g = gi.selector_loss_grad(X, T, M, _flat(p), n_out)
    want_loss, want = O.mlp_loss_backward(X, T, M, *p)
    assert abs(loss - want_loss) <= 1e-4 * abs(want_loss)
    got = _unflat(g, n_out)
    for a, b in zip(got, want):
        assert np.allclose(a, b, rtol=2e-3, atol=2e-5 * np.abs(b).max())


def test_selector_training_reduces_loss(gi):
    rng = np.random.default_rng(4)
    n_out = 3
    X = rng.normal(size=(256, 16))
    W = rng.normal(size=(16, n_out))
    T = np.tanh(X @ W)
    M = np.ones_like(T)
    p0 = _flat(_init_params(rng, n_out))
    p, hist = gi.selector_train(X, T, M, p0, n_out, 400, 3e-3)
    assert hist[-1] < 0.2 * hist[0]
    loss, _ = gi.selector_loss_grad(X, T, M, p, n_out)
    want, _ = O.mlp_loss_backward(X, T, M, *_unflat(p, n_out))
    assert abs(loss - want) <= 1e-3 * max(want, 1e-6) + 1e-6


def test_learned_selector_outputs_stay_exact(gi, c1):
    g, host, want = c1
    rng = np.random.default_rng(5)
    p = _flat(_init_params(rng, 7))
    gi.selector_load(p, 7, 0x7f, np.zeros(16), np.ones(16))
    try:
        h = gi.HybridSchedule(fused(gi, 4), fused(gi, 5), 0.05, gi.SELECT_LEARNED)
        for src in [0, 17, 999]:
            depth, _, _ = gi.bfs(g, src, h)
            assert np.array_equal(depth, O.bfs_depth(host, src))
    finally:
        gi.selector_clear()


def _bf16(a):
    """float32 -> BF16 (round to nearest even) -> float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _raw_features(n, fdeg, unv, unv_e, level, prev_pull, prev_n, V, E):
    """The 16 selector inputs (bfs_raw_feature, which the device evaluates in FP32)."""
    lg = np.log2
    return np.array([n / V, fdeg / E, unv / V, unv_e / E, lg(1 + n) / 32, lg(1 + fdeg) / 40,
                     lg(1 + fdeg / max(n, 1)) / 20, lg(1 + E / V) / 20, lg(V) / 32,
                     min(level / 64, 1.0), float(prev_pull), (lg(1 + n) - lg(1 + prev_n)) / 16,
                     fdeg / (1 + unv_e), n / (1 + unv), lg(1 + unv) / 32, 1.0])


def test_learned_selector_picks_match_host_inference(gi, c1):
    """Every level's mode picked by the MLP at the start of its step launch equals the
    argmin of the same BF16 network evaluated on the host (float64) from the
    oracle's level sets, wherever the two best outputs are not a near-tie."""
    g, host, _ = c1
    V, E = host.num_vertices, len(host.nbrs)
    deg = np.diff(host.offsets).astype(np.int64)
    rng = np.random.default_rng(23)
    n_out, allowed = 7, 0x7F
    p = _init_params(rng, n_out)
    p[1] = rng.normal(size=64) * 0.1
    p[5] = rng.normal(size=n_out) * 0.1
    mean, scale = rng.normal(size=16) * 0.05, rng.uniform(0.5, 3.0, 16)
    gi.selector_load(_flat(p), n_out, allowed, mean, scale)
    try:
        h = gi.HybridSchedule(fused(gi, 4), fused(gi, 5), 0.05, gi.SELECT_LEARNED)
        depth = gi.DeviceArray(V, np.int32)
        checked = 0
        for src in O.pick_sources(host, 6, 29):
            recs = gi.bfs_profile(g, int(src), h, depth)
            d = O.bfs_depth(host, int(src))
            w = [_bf16(x) for x in p]
            prev_mode, prev_n, seen, seen_e = 0, 0, 0, 0
            for r in recs:
                lvl = r["level"]
                F = np.flatnonzero(d == lvl)
                n, fdeg = len(F), int(deg[F].sum())
                seen += n
                seen_e += fdeg
                x = (_raw_features(n, fdeg, V - seen, E - seen_e, lvl, prev_mode >= 5, prev_n, V, E)
                     .astype(np.float32).astype(np.float64) - mean.astype(np.float32)) * scale.astype(np.float32)
                h1 = np.maximum(x @ w[0] + w[1], 0)
                h2 = np.maximum(h1 @ w[2] + w[3], 0)
                y = h2 @ w[4] + w[5]
                order = np.argsort(y)
                got = r["load_balancing"] - 100
                if y[order[1]] - y[order[0]] > 1e-3 * max(1.0, abs(y[order[0]])):
                    assert got == order[0], (src, lvl, got, y)
                    checked += 1
                prev_mode, prev_n = got, n
        assert checked >= 10
    finally:
        gi.selector_clear()


def test_multi_source_bfs_matches_oracle(gi, c1):
    g, host, _ = c1
    srcs = O.pick_sources(host, 21, 9)
    h = gi.HybridSchedule(fused(gi, 4), fused(gi, 5), 0.01, gi.SELECT_HEURISTIC)
    depth, levels = gi.bfs_multi(g, srcs, h)
    want = O.bfs_multi_depth(host, srcs)
    assert np.array_equal(depth, want) and levels == want.max() + 1


def test_packed_small_graphs_bfs(gi):
    g, offs = gi.Graph.uniform_batch(64, 1024, 65536, 8, 5)
    want_offs, s, d = O.uniform_batch_edges(64, 1024, 65536, 8, 5)
    assert np.array_equal(offs, want_offs)
    host = O.build_csr(int(offs[-1]), s, d, True)
    off, nb, _ = g.read_csr()
    assert np.array_equal(off, host.offsets) and np.array_equal(nb, host.nbrs)
    srcs = offs[:-1] + (np.arange(64) * 7919) % np.diff(offs)
    h = gi.HybridSchedule(fused(gi, 4), fused(gi, 5), 0.02, gi.SELECT_HEURISTIC)
    for bins in [(64, 512, 8), (64, 4096, 8), (16, 1024, 4), (128, 16384, 8), (16, 64, 8)]:
        g.set_bins(*bins)
        depth, _ = gi.bfs_multi(g, srcs, h)
        assert np.array_equal(depth, O.bfs_multi_depth(host, srcs)), bins


def test_graph_replay_many_runs_stay_exact(gi, c1):
    """The level loop is a captured hipGraph replayed per run; every replay,
    with eager work between them, must give the oracle's depths."""
    g, host, _ = c1
    h = gi.HybridSchedule(fused(gi, 4), fused(gi, 5), 0.005, gi.SELECT_HEURISTIC)
    srcs = O.pick_sources(host, 33, 6)
    depth = gi.DeviceArray(host.num_vertices, np.int32)
    for s in srcs:
        gi.bfs_device(g, int(s), h, depth)
        got = depth.numpy()
        _ = gi.bfs(g, int(s), None)  # unfused run in between
        assert np.array_equal(got, O.bfs_depth(host, int(s)))


def test_autotune_picks_bins_that_keep_depths(gi):
    from easy_graphit_amd import autotune
    g, offs = gi.Graph.uniform_batch(32, 256, 4096, 8, 7)
    want_offs, s, d = O.uniform_batch_edges(32, 256, 4096, 8, 7)
    host = O.build_csr(int(offs[-1]), s, d, True)
    srcs = offs[:-1] + (np.arange(32) * 131) % np.diff(offs)
    best = autotune.tune_bins(g, srcs, lane_max=(16, 64), hub_min=(1024,), blocks_per_cu=(4, 8),
                              push_lb=(gi.LB_TWC, gi.LB_CM), reps=1)
    assert best["seconds"] == min(r["seconds"] for r in best["table"])
    assert len(best["table"]) == 2 * 1 * 2 + 2
    depth, _ = gi.bfs_multi(g, srcs, autotune.schedule(best["push_lb"]))
    assert np.array_equal(depth, O.bfs_multi_depth(host, srcs))
