"""gi_bfs_static, the BFS entry point that allocates nothing, against
gi_bfs (the entry point whose loop and kernels it shares): bit-exact depths,
valid parents, on two calls in a row of which the second has the smaller
graph (the fixed array keeps the first call's contents)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095


def _push(gi, lb):
    return gi.Schedule(direction=gi.PUSH, load_balancing=lb)


@pytest.mark.parametrize("lb", [0, 2, 3, 4])  # VERTEX_BASED, TWCE, WM, CM
def test_static_bfs_matches_gi_bfs_twice(gi, lb):
    big = gi.Graph.rmat(16, 16, SEED, True)
    small = gi.Graph.rmat(11, 8, 5, True)
    for g, src_seed in ((big, 3), (small, 4)):  # second call: smaller input
        host = O.CSR(*g.read_csr()[:2])
        src = int(O.pick_sources(host, src_seed, 1)[0])
        want, _, _ = gi.bfs(g, src, gi.HybridSchedule(_push(gi, lb), _push(gi, lb), 0.0, gi.SELECT_S1))
        depth = gi.DeviceArray(g.num_vertices, np.int32)
        parent = gi.DeviceArray(g.num_vertices, np.int32)
        gi.bfs_static(g, src, _push(gi, lb), depth, parent)
        d, p = depth.numpy(), parent.numpy()
        assert np.array_equal(d, want), (g.num_vertices, lb)
        assert np.array_equal(d, O.bfs_depth(host, src))
        assert O.bfs_parent_violations(host, src, d, p) == 0


def test_static_bfs_default_schedule_and_isolated_source(gi):
    e = np.array([[0, 1], [1, 2], [4, 5]], np.int32)
    g = gi.Graph.from_edges(7, e[:, 0], e[:, 1], True)
    depth = gi.DeviceArray(7, np.int32)
    for src, want in ((0, [0, 1, 2, -1, -1, -1, -1]), (6, [-1] * 6 + [0]), (5, [-1] * 4 + [1, 0, -1])):
        gi.bfs_static(g, src, None, depth)
        assert depth.numpy().tolist() == want


def test_static_bfs_rejects_what_it_cannot_run(gi):
    g = gi.Graph.rmat(8, 8, 1, True)
    depth = gi.DeviceArray(g.num_vertices, np.int32)
    for bad in (gi.Schedule(direction=gi.PULL, load_balancing=0),
                gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWC),
                gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_EDGE_ONLY),
                gi.Schedule(direction=gi.PUSH, load_balancing=0, frontier_creation=gi.FC_BITMAP),
                gi.Schedule(direction=gi.PUSH, load_balancing=0, kernel_fusion=1)):
        with pytest.raises(gi.GIError):
            gi.bfs_static(g, 0, bad, depth)
    with pytest.raises(gi.GIError):
        gi.bfs_static(g, g.num_vertices, None, depth)
    # more vertices than the fixed array holds
    n = gi.BFS_STATIC_MAX_VERTICES + 1
    big = gi.Graph.from_edges(n, np.array([0], np.int32), np.array([n - 1], np.int32), True)
    with pytest.raises(gi.GIError, match="up to"):
        gi.bfs_static(big, 0, None, gi.DeviceArray(n, np.int32))
    # after the refusals the entry point still works on the last good input
    gi.bfs_static(g, 0, None, depth)
    assert np.array_equal(depth.numpy(), O.bfs_depth(O.CSR(*g.read_csr()[:2]), 0))
