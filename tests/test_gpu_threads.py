"""Threading contract of the C ABI: one host thread per graph handle, and
different graphs on different threads at the same time. Two threads, each
with its own graph and stream, run the fused BFS, SSSP and CC engines
concurrently (their per-graph plans are created and looked up while the
other thread does the same); every result equals the oracle's."""
import ctypes
import threading

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095


def test_fused_engines_on_two_threads(gi):
    graphs, hosts = [], []
    for k in range(2):
        g = gi.Graph.rmat(13, 16, SEED + k, True).set_weights(SEED + k, 1, 255)
        off, nb, w = g.read_csr(weights=True)
        graphs.append(g)
        hosts.append(O.CSR(off, nb, w))
    srcs = [[int(s) for s in O.pick_sources(h, 3, 3)] for h in hosts]
    want_bfs = [[O.bfs_depth(h, s) for s in ss] for h, ss in zip(hosts, srcs)]
    want_sssp = [[O.sssp(h, s) for s in ss] for h, ss in zip(hosts, srcs)]
    want_cc = [O.cc(h) for h in hosts]
    fb = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWC, kernel_fusion=1,
                     pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP)
    fp = gi.Schedule(direction=gi.PULL, load_balancing=gi.LB_VERTEX_BASED, kernel_fusion=1,
                     pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP)
    hyb = gi.HybridSchedule(fb, fp, 0.05, gi.SELECT_HEURISTIC)
    fs = gi.Schedule(direction=gi.PUSH, load_balancing=1, kernel_fusion=1)
    fc = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_EDGE_ONLY, kernel_fusion=1)
    errors = []

    def work(k):
        stream = ctypes.c_void_p()
        try:
            gi.check(gi.lib().gi_stream_create(ctypes.byref(stream)))
            for _ in range(2):  # the second pass replays the captured graphs
                for s, wb, ws in zip(srcs[k], want_bfs[k], want_sssp[k]):
                    depth, _, _ = gi.bfs(graphs[k], s, hyb, stream=stream)
                    assert np.array_equal(depth, wb), f"thread {k} BFS from {s}"
                    dist, _ = gi.sssp(graphs[k], s, 16, fs, stream=stream)
                    assert np.array_equal(dist, ws), f"thread {k} SSSP from {s}"
                lab, _ = gi.cc(graphs[k], schedule=fc, stream=stream)
                assert np.array_equal(lab, want_cc[k]), f"thread {k} CC"
        except BaseException as e:  # reported by the main thread
            errors.append(e)
        finally:
            if stream.value:
                gi.lib().gi_stream_destroy(stream)

    threads = [threading.Thread(target=work, args=(k,), daemon=True) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=90)
    assert not any(t.is_alive() for t in threads), "a worker thread did not finish"
    assert not errors, errors[0]
    for g in graphs:
        g.destroy()
