"""Batched per-source runs: gi_bfs_batched (64 sources per pass) against
the oracle's BFS from each source, bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095


def host_csr(g, weights=False):
    off, nb, w = g.read_csr(weights=weights)
    return O.CSR(off, nb, w)


def check_rows(depth, host, srcs):
    for i, s in enumerate(srcs):
        want = O.bfs_depth(host, int(s))
        assert np.array_equal(depth[i], want), (i, int(s))


@pytest.mark.parametrize("n", [1, 63, 64, 100, 130])
def test_bfs_batched_rmat(gi, n):
    # full and ragged groups of 64; repeated sources get rows of their own
    g = gi.Graph.rmat(14, 16, SEED, True)
    host = host_csr(g)
    srcs = O.pick_sources(host, 3, n)
    if n > 10:
        srcs[7] = srcs[3]
    depth, st = gi.bfs_batched(g, srcs)
    assert depth.shape == (n, g.num_vertices)
    check_rows(depth, host, srcs)
    assert st["iterations"] >= 1


@pytest.mark.parametrize("graph", ["grid", "directed", "forest"])
def test_bfs_batched_deep_directed_and_disconnected(gi, graph):
    if graph == "grid":       # hundreds of levels, all pushed
        g = gi.Graph.grid(160, 30, 3)
    elif graph == "directed":  # pull over the transpose
        g = gi.Graph.rmat(13, 16, 5, False)
    else:                      # components, isolated vertices and isolated sources
        n = 30000
        rng = np.random.default_rng(8)
        a = rng.integers(0, n, 20000)
        b = np.minimum(a + rng.integers(1, 4, 20000), n - 1)
        e = np.stack([a, b]).astype(np.int32)
        g = gi.Graph.from_edges(n, e[0], e[1], True)
    host = host_csr(g)
    rng = np.random.default_rng(1)
    srcs = rng.integers(0, g.num_vertices, 70).astype(np.int32)
    depth, _ = gi.bfs_batched(g, srcs)
    check_rows(depth, host, srcs)


def test_bfs_batched_tiny_and_empty(gi):
    z = np.zeros(0, np.int32)
    g1 = gi.Graph.from_edges(1, z, z, True)
    d, _ = gi.bfs_batched(g1, [0, 0])
    assert d.tolist() == [[0], [0]]
    d, st = gi.bfs_batched(g1, np.zeros(0, np.int32))
    assert d.shape == (0, 1) and st["iterations"] == 0
    g = gi.Graph.from_edges(130, np.array([0, 64], np.int32), np.array([129, 65], np.int32), True)
    host = host_csr(g)
    srcs = np.array([0, 129, 64, 5], np.int32)
    d, _ = gi.bfs_batched(g, srcs)
    check_rows(d, host, srcs)
    with pytest.raises(gi.GIError):
        gi.bfs_batched(g, [130])


def test_bfs_batched_rmat20_256_sources(gi):
    # the batched configuration's graph: 256 sources, four groups
    g = gi.Graph.rmat(20, 16, SEED, True)
    host = host_csr(g)
    srcs = g.pick_sources(SEED, 256)
    depth = gi.DeviceArray(256 * g.num_vertices, np.int32)
    _, st = gi.bfs_batched(g, srcs, depth)
    d = depth.numpy().reshape(256, g.num_vertices)
    for i in range(0, 256, 5):  # the oracle's BFS on every fifth source
        assert np.array_equal(d[i], O.bfs_depth(host, int(srcs[i]))), i
    # every row: the source at 0, and reached exactly over its component
    lab = O.cc(host)
    assert (d[np.arange(256), srcs] == 0).all()
    for i in range(256):
        assert np.array_equal(d[i] >= 0, lab == lab[srcs[i]]), i


@pytest.mark.parametrize("batch", [0, 7, 64])
def test_sssp_batched_rmat(gi, batch):
    g = gi.Graph.rmat(13, 16, SEED, True).set_weights(SEED, 1, 255)
    host = host_csr(g, True)
    srcs = O.pick_sources(host, 5, 70)
    srcs[9] = srcs[2]
    pulls = 0
    for delta in (1, 16, 200, 4096):  # 1: every edge heavy; 4096: none (no light/heavy split)
        dist, st = gi.sssp_batched(g, srcs, delta, batch)
        for i, s in enumerate(srcs):
            assert np.array_equal(dist[i], O.sssp(host, int(s))), (batch, delta, i)
        assert st["iterations"] >= 1 and st["iterations"] == st["push_iterations"] + st["pull_iterations"]
        pulls += st["pull_iterations"]
    assert pulls > 0  # the pull steps ran too


@pytest.mark.parametrize("graph", ["grid", "directed", "forest"])
def test_sssp_batched_grid_directed_disconnected(gi, graph):
    if graph == "grid":
        g = gi.Graph.grid(96, 40, 3).set_weights(5, 1, 1000)
    elif graph == "directed":
        g = gi.Graph.rmat(12, 8, 4, False).set_weights(2, 1, 100)
    else:
        n = 20000
        rng = np.random.default_rng(8)
        a = rng.integers(0, n, 15000)
        b = np.minimum(a + rng.integers(1, 4, 15000), n - 1)
        e = np.stack([a, b]).astype(np.int32)
        g = gi.Graph.from_edges(n, e[0], e[1], True).set_weights(3, 1, 50)
    host = host_csr(g, True)
    rng = np.random.default_rng(2)
    srcs = rng.integers(0, g.num_vertices, 40).astype(np.int32)  # isolated sources too
    dist, _ = gi.sssp_batched(g, srcs, 64)
    for i, s in enumerate(srcs):
        assert np.array_equal(dist[i], O.sssp(host, int(s))), (graph, i)


def test_sssp_batched_matches_single_source_engine(gi):
    # the batched distances equal the fused single-source engine's
    g = gi.Graph.rmat(15, 16, SEED, True).set_weights(SEED, 1, 255)
    srcs = g.pick_sources(SEED, 20)
    dist, _ = gi.sssp_batched(g, srcs, 32)
    fs = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWC, kernel_fusion=1)
    for i, s in enumerate(srcs):
        d1, _ = gi.sssp(g, int(s), 32, fs)
        assert np.array_equal(dist[i], d1), i
    with pytest.raises(gi.GIError):
        gi.sssp_batched(g, srcs, 0)
    z, st = gi.sssp_batched(g, np.zeros(0, np.int32), 8)
    assert z.shape == (0, g.num_vertices) and st["iterations"] == 0
