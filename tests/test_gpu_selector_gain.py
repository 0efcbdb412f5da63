"""What the trained selector gains over the method without it.

The committed model (easy_graphit_amd/selector_model.npz, trained on the
calibration graphs of easy_graphit_amd/selector.py) against GraphIt's fixed
rule |F| <= t*V (operators.cpp:526) with the push/pull mode pair and
threshold tuned on the same calibration graphs (stored with the model), on
an RMAT-24 graph held out of calibration: depths exact under both, and the
learned selector faster, on the same layout and against the rule on its own
best layout."""
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def timed(gi, g, sched, srcs, depth, reps=3):
    gi.bfs_device(g, int(srcs[0]), sched, depth)
    best = float("inf")
    for _ in range(reps):
        gi.synchronize()
        t0 = time.perf_counter()
        for s in srcs:
            gi.bfs_device(g, int(s), sched, depth)
        gi.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def test_learned_selector_beats_the_tuned_fixed_rule(gi):
    from easy_graphit_amd import selector as S
    model = S.load()
    assert "rule" in model, "the model carries the calibration-tuned fixed rule"
    S.install(model)
    try:
        g = gi.Graph.rmat(24, 16, 31, True)  # seed 31: not a calibration graph
        host = O.CSR(*g.read_csr()[:2])
        srcs = g.pick_sources(31, 6)
        learned = S.fused_schedule(threshold=model["threshold"], selector=gi.SELECT_LEARNED)
        rule = S.rule_schedule(*model["rule"])
        depth = gi.DeviceArray(g.num_vertices, np.int32)
        times = {}
        for layout in (gi.LAYOUT_DEGREE, gi.LAYOUT_NATIVE):
            g.set_layout(layout)
            for s in srcs[:2]:
                want = O.bfs_depth(host, int(s))
                for sched in (learned, rule):
                    d, _, _ = gi.bfs(g, int(s), sched)
                    assert np.array_equal(d, want), (layout, int(s))
            times[layout] = (timed(gi, g, learned, srcs, depth), timed(gi, g, rule, srcs, depth))
        (l_deg, r_deg), (l_nat, r_nat) = times[gi.LAYOUT_DEGREE], times[gi.LAYOUT_NATIVE]
        print(f"learned {l_deg * 1e3:.2f} ms (degree layout), {l_nat * 1e3:.2f} ms (native); "
              f"tuned rule {r_deg * 1e3:.2f} ms (degree), {r_nat * 1e3:.2f} ms (native)")
        # same layout: the rule counts frontier vertices, not their edges, and
        # pushes levels whose few vertices hold millions of edges into hubs
        assert l_deg < 0.5 * r_deg
        # the selector on its layout against the rule on the rule's best layout
        assert l_deg < 0.8 * min(r_deg, r_nat)
    finally:
        gi.selector_clear()
