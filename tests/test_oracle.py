"""Pins the CPU oracle: published vectors, hand-worked fixtures, the
reference's own GAPBS builder (compiled from its sources into oracle/_ref),
textbook library routines (scipy / networkx) and brute force. CPU only."""
import itertools
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "small_graphs.json")))


def fixture_csr(name):
    f = FIX[name]
    e = np.array(f["edges"], np.int32).reshape(-1, 2)
    return O.build_csr(f["V"], e[:, 0], e[:, 1], f["symmetrize"])


def to_scipy(g):
    n = g.num_vertices
    data = np.ones(g.num_edges) if g.weights is None else g.weights.astype(np.float64)
    return sp.csr_matrix((data, g.nbrs, g.offsets), shape=(n, n))


# ---------------------------------------------------------------- generators
def test_splitmix64_published_vectors():
    # SplitMix64 reference outputs: seed 0 -> 0xE220A8397B1DCDAF; seed 1234567
    # -> 6457827717110365317, 3203168211198807973, 9817491932198370423,
    # 4593380528125082431, 16408922859458223821.
    assert O.hash64(0, 0, 0) == 0xE220A8397B1DCDAF
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    assert [O.hash64(1234567, 0, i) for i in range(5)] == want


def test_permutation_is_bijective():
    L = O.lib()
    for scale in range(1, 15):
        img = {L.or_permute_id(x, scale, 27491095) for x in range(1 << scale)}
        assert img == set(range(1 << scale))


def test_rmat_quadrant_frequencies():
    # Top-level quadrant of each edge must follow (a,b,c,d) = (.57,.19,.19,.05).
    s, d = O.rmat_edges(12, 16, 7, permute=False)
    top_s, top_d = s >> 11, d >> 11
    q = np.bincount(top_s * 2 + top_d, minlength=4) / len(s)
    assert np.allclose(q, [0.57, 0.19, 0.19, 0.05], atol=0.01)


def test_rmat_scale2_hand_draws():
    # Edge 0 of a scale-2 graph: one 64-bit word, low half = level 0, high half
    # = level 1, each compared with floor(p * 2^32).
    h = O.hash64(99, 1, 0)
    tA, tAB, tABC = 2448131358, 3264175144, 4080218931
    s = d = 0
    for r in (h & 0xFFFFFFFF, h >> 32):
        s <<= 1
        d <<= 1
        if r < tAB:
            d += r >= tA
        else:
            s += 1
            d += r >= tABC
    gs, gd = O.rmat_edges(2, 1, 99, permute=False)
    assert (gs[0], gd[0]) == (s, d)


def test_grid_counts_and_degrees():
    w = 17
    s, d = O.grid_edges(w, 0, 1)
    g = O.build_csr(w * w, s, d, True)
    assert g.num_edges == 2 * 2 * w * (w - 1)
    deg = g.degrees().reshape(w, w)
    assert deg[0, 0] == 2 and deg[0, 5] == 3 and deg[5, 5] == 4


def test_weights_symmetric_and_in_range():
    s, d = O.rmat_edges(8, 8, 3)
    g = O.fill_weights(O.build_csr(256, s, d, True), 11, 1, 1000)
    assert g.weights.min() >= 1 and g.weights.max() <= 1000
    m = to_scipy(g)
    assert (m != m.T).nnz == 0


# ---------------------------------------------------------------- CSR build
def test_csr_fixture_squish():
    g = fixture_csr("self_loops_dups")
    assert list(g.offsets) == FIX["self_loops_dups"]["csr_offsets"]
    assert list(g.nbrs) == FIX["self_loops_dups"]["csr_nbrs"]


@pytest.mark.parametrize("sym", [True, False])
def test_csr_matches_gapbs_reference_builder(sym):
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    s, d = O.rmat_edges(11, 16, 5)
    mine = O.build_csr(1 << 11, s, d, sym)
    theirs = O.ref_build_csr(1 << 11, s, d, sym)
    assert np.array_equal(mine.offsets, theirs.offsets)
    assert np.array_equal(mine.nbrs, theirs.nbrs)


def test_csr_matches_gapbs_on_its_own_rmat():
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    R = O.ref()
    scale = 10
    m = (1 << scale) * 16
    s = np.empty(m, np.int32)
    d = np.empty(m, np.int32)
    assert R.ref_rmat_edges(scale, 16, s, d) == m
    mine = O.build_csr(1 << scale, s, d, True)
    theirs = O.ref_build_csr(1 << scale, s, d, True)
    assert np.array_equal(mine.offsets, theirs.offsets) and np.array_equal(mine.nbrs, theirs.nbrs)


def test_csr_matches_scipy_canonical():
    s, d = O.uniform_edges(500, 3000, 8)
    g = O.build_csr(500, s, d, True)
    a = sp.coo_matrix((np.ones(len(s)), (s, d)), shape=(500, 500)).tocsr()
    a = a + a.T
    a.setdiag(0)
    a.eliminate_zeros()
    a.sum_duplicates()
    a.sort_indices()
    assert np.array_equal(g.offsets, a.indptr) and np.array_equal(g.nbrs, a.indices)


def test_transpose_matches_scipy_and_involutes():
    s, d = O.rmat_edges(9, 8, 4)
    g = O.build_csr(512, s, d, False)
    t = O.transpose(g)
    a = to_scipy(g).T.tocsr()
    a.sort_indices()
    assert np.array_equal(t.offsets, a.indptr) and np.array_equal(t.nbrs, a.indices)
    tt = O.transpose(t)
    assert np.array_equal(tt.offsets, g.offsets) and np.array_equal(tt.nbrs, g.nbrs)


# ---------------------------------------------------------------- BFS
@pytest.mark.parametrize("name,src", [("path5", 0), ("star4", 1), ("two_triangles", 5)])
def test_bfs_fixtures(name, src):
    assert list(O.bfs_depth(fixture_csr(name), src)) == FIX[name][f"bfs_from_{src}"]


def test_bfs_matches_scipy():
    s, d = O.rmat_edges(12, 16, 27491095)
    g = O.build_csr(1 << 12, s, d, True)
    dep = O.bfs_depth(g, 0)
    ref = csg.shortest_path(to_scipy(g), unweighted=True, indices=0)
    want = np.where(np.isinf(ref), -1, ref).astype(np.int32)
    assert np.array_equal(dep, want)


def test_bfs_parent_checker_catches_bad_parent():
    g = fixture_csr("path5")
    dep = O.bfs_depth(g, 0)
    good = np.array([0, 0, 1, 2, -1], np.int32)
    assert O.bfs_parent_violations(g, 0, dep, good) == 0
    bad = good.copy()
    bad[3] = 1  # not adjacent
    assert O.bfs_parent_violations(g, 0, dep, bad) == 1


def test_pull_examined_counts():
    g = fixture_csr("path5")
    dep = O.bfs_depth(g, 0)
    # building level 2 from level 1 = {1}: unvisited {2,3,4}: vertex 2 finds
    # 1 at its first neighbour (1 step); 3's only neighbour 2 is not in the
    # frontier (1 step); 4 has no neighbours (0).
    assert O.bfs_pull_examined(g, dep, 1) == 1 + 1 + 0


# ---------------------------------------------------------------- SSSP
def test_sssp_fixture():
    f = FIX["weighted_diamond"]
    g = fixture_csr("weighted_diamond")
    w = np.empty(g.num_edges, np.int32)
    for u in range(g.num_vertices):
        for e in range(g.offsets[u], g.offsets[u + 1]):
            v = int(g.nbrs[e])
            w[e] = f["weights"][f"{min(u, v)}-{max(u, v)}"]
    g.weights = w
    assert list(O.sssp(g, 0)) == f["sssp_from_0"]
    assert list(O.sssp_delta_stepping(g, 0, 2)[0]) == f["sssp_from_0"]


def test_sssp_matches_scipy_dijkstra_and_delta_stepping():
    s, d = O.grid_edges(40, 16, 3)
    g = O.fill_weights(O.build_csr(1600, s, d, True), 9, 1, 1000)
    ref = csg.dijkstra(to_scipy(g), indices=0)
    mine = O.sssp(g, 0)
    assert np.array_equal(mine, ref.astype(np.int64))
    ds, nb = O.sssp_delta_stepping(g, 0, 64)
    assert np.array_equal(ds, mine) and nb > 1


def test_sssp_unreached_is_int_max():
    g = fixture_csr("two_triangles")
    g.weights = np.ones(g.num_edges, np.int32)
    dist = O.sssp(g, 0)
    assert dist[3] == 2**31 - 1 and dist[4] == 2**31 - 1 and dist[1] == 1


# ---------------------------------------------------------------- CC
@pytest.mark.parametrize("name", ["path5", "star4", "two_triangles"])
def test_cc_fixtures(name):
    g = fixture_csr(name)
    assert list(O.cc(g)) == FIX[name]["cc"]
    assert list(O.cc_label_propagation(g)[0]) == FIX[name]["cc"]


def test_cc_matches_scipy_min_id():
    s, d = O.uniform_edges(3000, 2000, 12)
    g = O.build_csr(3000, s, d, True)
    ncomp, lab = csg.connected_components(to_scipy(g), directed=False)
    mins = np.full(ncomp, 1 << 30)
    np.minimum.at(mins, lab, np.arange(3000))
    assert np.array_equal(O.cc(g), mins[lab])
    lp, sweeps = O.cc_label_propagation(g)
    assert np.array_equal(lp, mins[lab]) and sweeps >= 1


# ---------------------------------------------------------------- PageRank
def test_pr_cycle_uniform():
    g = fixture_csr("cycle3_directed")
    assert np.allclose(O.pagerank(g, 20), FIX["cycle3_directed"]["pr20"], rtol=0, atol=1e-15)


def test_pr_star_closed_form():
    # directed star: leaves 1..k point at 0; 0 has no out-edges (dangling,
    # mass is lost as in apps/pr.cpp). After one iteration: r0 = (1-d)/N +
    # d*k/N, leaves (1-d)/N.
    k = 5
    e = np.array([[i, 0] for i in range(1, k + 1)], np.int32)
    g = O.build_csr(k + 1, e[:, 0], e[:, 1], False)
    d = O.PR_DAMP
    r = O.pagerank(g, 1, d)
    n = k + 1
    assert np.isclose(r[0], (1 - d) / n + d * k / n) and np.allclose(r[1:], (1 - d) / n)


def test_pr_matches_networkx_on_dangling_free_graph():
    nx = pytest.importorskip("networkx")
    s, d = O.uniform_edges(300, 3000, 21)
    ring = np.arange(300, dtype=np.int32)
    s = np.concatenate([s, ring])
    d = np.concatenate([d, (ring + 1) % 300])  # every vertex has out-degree >= 1
    g = O.build_csr(300, s, d, False)
    G = nx.DiGraph()
    G.add_nodes_from(range(300))
    for u in range(300):
        for v in g.nbrs[g.offsets[u]:g.offsets[u + 1]]:
            G.add_edge(u, int(v))
    ref = nx.pagerank(G, alpha=O.PR_DAMP, tol=1e-14, max_iter=1000)
    mine = O.pagerank(g, 300)
    assert np.allclose(mine, [ref[i] for i in range(300)], rtol=1e-9, atol=1e-12)


def test_pr_matches_dense_power_iteration():
    s, d = O.rmat_edges(8, 8, 2)
    g = O.build_csr(256, s, d, False)
    n = 256
    A = np.zeros((n, n))
    for u in range(n):
        for v in g.nbrs[g.offsets[u]:g.offsets[u + 1]]:
            A[v, u] += 1.0
    deg = np.diff(g.offsets).astype(np.float64)
    M = A / np.where(deg > 0, deg, 1.0)[None, :]
    r = np.full(n, 1.0 / n)
    for _ in range(20):
        r = (1 - O.PR_DAMP) / n + O.PR_DAMP * (M @ r)
    assert np.allclose(O.pagerank(g, 20), r, rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- BC
@pytest.mark.parametrize("name", ["path5", "star4"])
def test_bc_fixtures(name):
    g = fixture_csr(name)
    assert np.allclose(O.bc(g, FIX[name]["bc_sources"]), FIX[name]["bc"])


def _brute_bc(g, sources):
    """Σ_s Σ_t σ_st(v)/σ_st over t != s, plus for v == s the count of other
    reachable vertices (the reference keeps the source's own term)."""
    n = g.num_vertices
    adj = [list(g.nbrs[g.offsets[u]:g.offsets[u + 1]]) for u in range(n)]
    out = np.zeros(n)
    for s in sources:
        dist = O.bfs_depth(g, s)
        paths = {}

        def all_paths(t):
            if t in paths:
                return paths[t]
            if t == s:
                paths[t] = [[s]]
                return paths[t]
            res = []
            for u in adj[t]:
                if dist[u] == dist[t] - 1:
                    res += [p + [t] for p in all_paths(u)]
            paths[t] = res
            return res

        for t in range(n):
            if t == s or dist[t] < 0:
                continue
            ps = all_paths(t)
            for p in ps:
                for v in p[:-1]:
                    out[v] += 1.0 / len(ps)
    return out


def test_bc_brute_force_small_graphs():
    for seed in range(4):
        s, d = O.uniform_edges(9, 14, 100 + seed)
        g = O.build_csr(9, s, d, True)
        srcs = [0, 3, 7]
        assert np.allclose(O.bc(g, srcs), _brute_bc(g, srcs), rtol=1e-12)


def test_bc_matches_networkx_dependencies():
    nx = pytest.importorskip("networkx")
    s, d = O.rmat_edges(8, 8, 6)
    g = O.build_csr(256, s, d, True)
    G = nx.Graph()
    G.add_nodes_from(range(256))
    for u in range(256):
        for v in g.nbrs[g.offsets[u]:g.offsets[u + 1]]:
            G.add_edge(u, int(v))
    srcs = O.pick_sources(g, 5, 6)
    mine = O.bc(g, srcs)
    # networkx betweenness over a source subset, unnormalized, sums
    # dependencies of v != s; add the reference's own-source term back.
    ref = nx.betweenness_centrality_subset(G, sources=list(map(int, srcs)),
                                           targets=list(range(256)), normalized=False)
    want = np.array([ref[i] for i in range(256)]) * 2.0  # subset BC halves undirected
    for s0 in srcs:
        want[s0] += (O.bfs_depth(g, int(s0)) >= 0).sum() - 1
    assert np.allclose(mine, want, rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- sources
def test_pick_sources_skip_degree_zero():
    g = fixture_csr("two_triangles")  # vertex 3 is isolated
    srcs = O.pick_sources(g, 1, 200)
    assert 3 not in set(srcs.tolist()) and len(srcs) == 200


# ---------------------------------------------------------------- MLP
def _mlp_params(rng, nin=16, nh=8, nout=5):
    return [rng.normal(size=(nin, nh)) * 0.3, rng.normal(size=nh) * 0.1,
            rng.normal(size=(nh, nh)) * 0.3, rng.normal(size=nh) * 0.1,
            rng.normal(size=(nh, nout)) * 0.3, rng.normal(size=nout) * 0.1]


def test_mlp_forward_matches_numpy():
    rng = np.random.default_rng(0)
    p = _mlp_params(rng)
    x = rng.normal(size=(7, 16))
    h1 = np.maximum(x @ p[0] + p[1], 0)
    h2 = np.maximum(h1 @ p[2] + p[3], 0)
    assert np.allclose(O.mlp_forward(x, *p), h2 @ p[4] + p[5], rtol=1e-13)


def test_mlp_gradients_finite_differences():
    rng = np.random.default_rng(1)
    p = _mlp_params(rng)
    x = rng.normal(size=(6, 16))
    t = rng.normal(size=(6, 5))
    # weights: zeros mask entries out, the rest weigh them
    m = (rng.random((6, 5)) > 0.3) * rng.uniform(0.2, 3.0, size=(6, 5))
    loss, grads = O.mlp_loss_backward(x, t, m, *p)
    y = O.mlp_forward(x, *p)
    assert np.isclose(loss, (m * (y - t) ** 2).sum() / m.sum(), rtol=1e-12)
    eps = 1e-6
    for pi, (param, grad) in enumerate(zip(p, grads)):
        flat = param.reshape(-1)
        for k in itertools.islice(range(flat.size), 0, flat.size, max(1, flat.size // 9)):
            old = flat[k]
            flat[k] = old + eps
            lp, _ = O.mlp_loss_backward(x, t, m, *p)
            flat[k] = old - eps
            lm, _ = O.mlp_loss_backward(x, t, m, *p)
            flat[k] = old
            assert abs((lp - lm) / (2 * eps) - grad.reshape(-1)[k]) < 1e-6, (pi, k)


# ---------------------------------------------------------------- batches
def test_bfs_multi_depth_is_min_over_single_sources():
    s, d = O.rmat_edges(10, 8, 77)
    g = O.build_csr(1 << 10, s, d, True)
    srcs = O.pick_sources(g, 5, 6)
    srcs = np.concatenate([srcs, srcs[:2]])  # repeated sources count once
    singles = np.stack([O.bfs_depth(g, int(x)) for x in srcs])
    want = np.where((singles >= 0).any(0), np.where(singles >= 0, singles, 1 << 30).min(0), -1)
    assert np.array_equal(O.bfs_multi_depth(g, srcs), want)


def test_uniform_batch_sizes_ranges_and_degree():
    G, vmin, vmax, deg = 400, 1024, 65536, 8
    offs, s, d = O.uniform_batch_edges(G, vmin, vmax, deg, 5)
    sizes = np.diff(offs)
    assert offs[0] == 0 and (sizes >= vmin).all() and (sizes <= vmax).all()
    # log-uniform sizes: log2(size) spreads evenly over the 6 octaves
    oct_ = np.floor(np.log2(sizes / vmin)).astype(int)
    hist = np.bincount(np.minimum(oct_, 5), minlength=6)
    assert hist.min() > 0.5 * G / 6 and hist.max() < 1.5 * G / 6
    # every edge stays inside its graph, pairs = size * avg_degree / 2
    gid = np.searchsorted(offs, s, side="right") - 1
    assert np.array_equal(gid, np.searchsorted(offs, d, side="right") - 1)
    per_graph = np.bincount(gid, minlength=G)
    assert np.array_equal(per_graph, sizes * deg // 2)
    # same seed -> same batch; another seed -> another batch
    o2, s2, _ = O.uniform_batch_edges(G, vmin, vmax, deg, 5)
    assert np.array_equal(o2, offs) and np.array_equal(s2, s)
    assert not np.array_equal(O.uniform_batch_edges(G, vmin, vmax, deg, 6)[0], offs)


def test_openmp_bfs_matches_serial():
    # the bench's CPU baseline: direction-optimizing OpenMP BFS, equal depths
    # to the serial queue BFS for every thread count, top-down only on a
    # directed graph
    s, d = O.rmat_edges(13, 16, 5)
    sym = O.build_csr(1 << 13, s, d, True)
    dirg = O.build_csr(1 << 13, s, d, False)
    gs, gd = O.grid_edges(60, 40, 2)
    grid = O.build_csr(60 * 60, gs, gd, True)
    path = O.build_csr(3000, np.arange(2999), np.arange(1, 3000), True)
    for g in (sym, dirg, grid, path):
        for src in list(O.pick_sources(g, 3, 3)) + [0]:
            want = O.bfs_depth(g, int(src))
            for threads in (1, 2, 8):
                got, used = O.bfs_depth_omp(g, int(src), threads)
                assert np.array_equal(got, want), (g.directed, src, threads)
                assert used == threads


def test_pagerank_pull_equals_serial_push_bit_for_bit():
    # the oracle pulls over in-neighbours in ascending source order (in
    # parallel); the serial push of apps/pr.cpp adds the same terms in the
    # same order, restated here with np.add.at (unbuffered, in index order)
    s, d = O.rmat_edges(10, 8, 21)
    g = O.build_csr(1024, s, d, False)
    n = g.num_vertices
    deg = np.diff(g.offsets)
    src_of_edge = np.repeat(np.arange(n), deg)
    rank = np.full(n, 1.0 / n)
    beta = (1.0 - O.PR_DAMP) / n
    for _ in range(7):
        contrib = np.where(deg > 0, rank / np.maximum(deg, 1), 0.0)
        nw = np.zeros(n)
        np.add.at(nw, g.nbrs, contrib[src_of_edge])
        rank = beta + O.PR_DAMP * nw
    for threads in (1, 8):
        prev = O.set_threads(threads)
        got = O.pagerank(g, 7)
        O.set_threads(prev)
        assert np.array_equal(got, rank), threads


def test_bc_thread_counts_agree():
    # sources split over threads: per-thread sums in source order, added in
    # thread order; equal to the serial loop up to FP64 rounding
    s, d = O.rmat_edges(10, 8, 31)
    g = O.build_csr(1024, s, d, True)
    srcs = O.pick_sources(g, 3, 37)
    prev = O.set_threads(1)
    one = O.bc(g, srcs)
    O.set_threads(8)
    many = O.bc(g, srcs)
    O.set_threads(prev)
    assert np.allclose(one, many, rtol=1e-12, atol=1e-12)
    assert np.allclose(one, O.bc(g, srcs[:1]) + O.bc(g, srcs[1:]), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("threads", [1, 2, 8])
def test_openmp_sssp_and_cc_baselines_equal_the_serial_oracle(threads):
    # symmetric RMAT, a directed graph, a grid and a weighted chain with long
    # edges (pending distances far beyond the current bucket)
    graphs = []
    s, d = O.rmat_edges(12, 16, 7)
    graphs.append(O.fill_weights(O.build_csr(4096, s, d, True), 3, 1, 255))
    s, d = O.rmat_edges(11, 8, 8)
    graphs.append(O.fill_weights(O.build_csr(2048, s, d, False), 4, 1, 1000))
    s, d = O.grid_edges(40, 30, 5)
    graphs.append(O.fill_weights(O.build_csr(1600, s, d, True), 5, 1, 1000))
    n = 3000
    a = np.arange(n - 1, dtype=np.int32)
    g = O.build_csr(n, np.concatenate([a, [0, 5]]).astype(np.int32),
                    np.concatenate([a + 1, [2999, 2000]]).astype(np.int32), True)
    graphs.append(O.fill_weights(g, 6, 1, 30000))
    for gi_, g in enumerate(graphs):
        for src in O.pick_sources(g, 11, 2):
            want = O.sssp(g, int(src))
            for delta in (1, 16, 300):
                got, used = O.sssp_omp(g, int(src), delta, threads)
                assert used == threads and np.array_equal(got, want), (gi_, int(src), delta)
        if gi_ != 1:  # components of the symmetric graphs
            lab, _ = O.cc_omp(g, threads)
            assert np.array_equal(lab, O.cc(g)), gi_
