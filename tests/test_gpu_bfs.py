"""BFS on the device through the C ABI vs the serial oracle: depths
bit-exact for every load-balancing schedule, direction and frontier
representation; parents checked for validity."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095


@pytest.fixture(scope="module")
def c1(gi):
    """RMAT scale 16, EF 16, symmetrized (the minimum slice workload)."""
    g = gi.Graph.rmat(16, 16, SEED, True)
    off, nb, _ = g.read_csr()
    host = O.CSR(off, nb)
    return g, host, O.bfs_depth(host, 0)


def push(gi, lb, **kw):
    return gi.Schedule(direction=gi.PUSH, load_balancing=lb, **kw)


def pull(gi, lb, rep, **kw):
    return gi.Schedule(direction=gi.PULL, load_balancing=lb, pull_frontier_rep=rep, **kw)


@pytest.mark.parametrize("lb", range(7))
def test_push_all_load_balancers_bit_exact(gi, c1, lb):
    g, host, want = c1
    h = gi.HybridSchedule(push(gi, lb), push(gi, lb), selector=gi.SELECT_S1)
    depth, parent, st = gi.bfs(g, 0, h, want_parent=True)
    assert np.array_equal(depth, want)
    assert O.bfs_parent_violations(host, 0, depth, parent) == 0
    assert st["reached"] == int((want >= 0).sum())


@pytest.mark.parametrize("lb", range(7))
@pytest.mark.parametrize("rep", [0, 1])
def test_pull_all_load_balancers_bit_exact(gi, c1, lb, rep):
    g, host, want = c1
    s = pull(gi, lb, rep, frontier_creation=gi.FC_BITMAP if rep == 0 else gi.FC_BOOLMAP)
    h = gi.HybridSchedule(s, s, selector=gi.SELECT_S1)
    depth, parent, _ = gi.bfs(g, 0, h, want_parent=True)
    assert np.array_equal(depth, want)
    assert O.bfs_parent_violations(host, 0, depth, parent) == 0


@pytest.mark.parametrize("fc", [0, 1, 2])
def test_push_frontier_creation_formats(gi, c1, fc):
    g, _, want = c1
    s = push(gi, 2, frontier_creation=fc)
    depth, _, _ = gi.bfs(g, 0, gi.HybridSchedule(s, s, selector=gi.SELECT_S1))
    assert np.array_equal(depth, want)


@pytest.mark.parametrize("t", [0.0, 0.001, 0.05, 1.0])
def test_hybrid_heuristic_bit_exact(gi, c1, t):
    g, _, want = c1
    s1 = push(gi, gi.LB_TWCE)
    s2 = pull(gi, gi.LB_VERTEX_BASED, gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP)
    depth, _, st = gi.bfs(g, 0, gi.HybridSchedule(s1, s2, t, gi.SELECT_HEURISTIC))
    assert np.array_equal(depth, want)
    assert st["push_iterations"] + st["pull_iterations"] == st["iterations"]


def test_default_schedule_and_many_sources(gi, c1):
    g, host, _ = c1
    for src in O.pick_sources(host, 3, 6):
        depth, _, _ = gi.bfs(g, int(src), None)
        assert np.array_equal(depth, O.bfs_depth(host, int(src)))


def test_cta_sizes(gi, c1):
    g, _, want = c1
    for cta in (64, 256, 1024):
        for lb in (gi.LB_TWCE, gi.LB_CM, gi.LB_TWC):
            s = push(gi, lb, cta_size=cta, max_cta=37)
            depth, _, _ = gi.bfs(g, 0, gi.HybridSchedule(s, s, selector=gi.SELECT_S1))
            assert np.array_equal(depth, want), (cta, lb)


def test_reference_bfs_loop_through_edgeset_apply(gi, c1):
    """apps/bfs.cpp written against the operator API: parent[dst] = src with
    the CAS filter, apply_modified into the same frontier until empty."""
    g, host, want = c1
    n = host.num_vertices
    parent = gi.DeviceArray(n, np.int32).fill(-1)
    parent_host = np.full(n, -1, np.int32)
    parent_host[0] = 0
    parent.copy_from(parent_host)
    f = gi.Frontier(n)
    f.add_vertex(0)
    s = push(gi, gi.LB_TWCE)
    rounds = 0
    while f.size() != 0:
        gi.edgeset_apply(g, s, f, f, gi.OP_BFS_PARENT, gi.op_args(parent))
        rounds += 1
    p = parent.numpy()
    assert rounds == want.max() + 1
    assert O.bfs_parent_violations(host, 0, want, p) == 0


def test_isolated_source_and_errors(gi):
    e = np.array([[0, 1], [1, 2]], np.int32)
    g = gi.Graph.from_edges(5, e[:, 0], e[:, 1], True)
    depth, _, st = gi.bfs(g, 4, None)
    assert list(depth) == [-1, -1, -1, -1, 0] and st["iterations"] == 1
    with pytest.raises(gi.GIError) as ex:
        gi.bfs(g, 5, None)
    assert ex.value.code == -5


def test_frontier_operations(gi):
    n = 1000
    f = gi.Frontier(n)
    for v in (5, 70, 999, 70):
        f.add_vertex(v)
    assert f.size() == 4
    f.dedup()
    assert sorted(f.read().tolist()) == [5, 70, 999]
    for fmt in (gi.FMT_BITMAP, gi.FMT_BOOLMAP, gi.FMT_SPARSE, gi.FMT_BOOLMAP, gi.FMT_BITMAP):
        f.convert(fmt)
        assert f.size() == 3
        assert sorted(f.read().tolist()) == [5, 70, 999]
    full = gi.Frontier(n, init_all=True)
    assert full.size() == n
    lst = gi.FrontierList(n + 3)
    lst.insert(f)
    lst.insert(full)
    with pytest.raises(gi.GIError) as ex:
        lst.insert(f)  # 1006 > 1003
    assert ex.value.code == -5
    out = gi.Frontier(n)
    lst.retrieve(out)
    assert out.size() == n
    lst.retrieve(out)
    assert sorted(out.read().tolist()) == [5, 70, 999]
    with pytest.raises(gi.GIError) as ex:
        lst.retrieve(out)
    assert ex.value.code == -7
