"""Autotuners for the fused engines.

Bins and launch shape of the fused BFS engine: it walks a push frontier
vertex in one lane when its degree is below `lane_max`, with a whole
wavefront up to `hub_min`, and spreads wider vertices over the whole grid;
its persistent step kernels run `blocks_per_cu` workgroups of 256 threads per
CU (gi_bfs_set_bins). The best values depend on the degree distribution of
the workload: many small uniform graphs packed in one CSR want different bins
than one RMAT graph with hubs. tune_bins() times the real workload (a
multi-source fused BFS over the packed graphs, one source per graph) under
every candidate setting, checks that every setting gives the same depths,
keeps the fastest and leaves it applied to the graph.

Delta of the fused SSSP engine: a small delta settles vertices in many thin
windows (more launches, fewer repeated relaxations), a large one in few wide
windows (each vertex relaxed more often). The delta is a run-time argument
of the reference app (apps/sssp.cpp:34); tune_delta() times the candidates
on one source, checks that all give the same distances and returns the
fastest.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

import easy_graphit_amd as gi

LANE_MAX = (16, 32, 64, 128)
HUB_MIN = (256, 512, 2048, 8192)
BLOCKS_PER_CU = (4, 8)
PUSH_LB = (gi.LB_TWC, gi.LB_TWCE, gi.LB_CM)


def schedule(push_lb: int = gi.LB_TWC, threshold: float = 0.02):
    s1 = gi.Schedule(direction=gi.PUSH, load_balancing=push_lb, kernel_fusion=1,
                     frontier_creation=gi.FC_BITMAP, cta_size=256)
    s2 = gi.Schedule(direction=gi.PULL, load_balancing=gi.LB_VERTEX_BASED, kernel_fusion=1,
                     pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP, cta_size=256)
    return gi.HybridSchedule(s1, s2, threshold, gi.SELECT_HEURISTIC)


def time_multi(g: gi.Graph, sources, sched, depth: gi.DeviceArray, reps: int) -> float:
    """Best wall time (s) of one multi-source fused BFS over `reps` runs."""
    gi.bfs_multi(g, sources, sched, depth)  # warm-up (records the level graph)
    gi.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        gi.bfs_multi(g, sources, sched, depth)
        gi.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune_bins(g: gi.Graph, sources, lane_max=LANE_MAX, hub_min=HUB_MIN,
              blocks_per_cu=BLOCKS_PER_CU, push_lb=PUSH_LB, threshold: float = 0.02,
              reps: int = 3, log=None) -> dict:
    """Search bins x launch shape x push schedule; apply and return the best.

    Returns {"lane_max", "hub_min", "blocks_per_cu", "push_lb", "seconds",
    "default_seconds", "table"}; "default_seconds" is the engine's default
    setting (64, 512, 8, TWC) on the same workload."""
    depth = gi.DeviceArray(g.num_vertices, np.int32)
    ref = None
    table = []
    cands = []
    for lb in push_lb:
        if lb == gi.LB_TWC:  # the lane / wave / grid bins are TWC's
            cands += [(lm, hm, bpc, lb) for lm, hm, bpc in itertools.product(lane_max, hub_min, blocks_per_cu)
                      if hm > lm]
        else:                # TWCE and CM bin inside the workgroup: launch shape only
            cands += [(64, 512, bpc, lb) for bpc in blocks_per_cu]
    for lm, hm, bpc, lb in cands:
        g.set_bins(lm, hm, bpc)
        sched = schedule(lb, threshold)
        t = time_multi(g, sources, sched, depth, reps)
        got = depth.numpy()
        if ref is None:
            ref = got
        elif not np.array_equal(got, ref):
            raise AssertionError(f"bins {(lm, hm, bpc, lb)} changed the depths")
        table.append({"lane_max": lm, "hub_min": hm, "blocks_per_cu": bpc, "push_lb": lb,
                      "seconds": t})
        if log:
            log(f"  bins lane<{lm} hub>={hm} {bpc}/CU {gi.LB_NAMES[lb]}: {t * 1e3:.3f} ms")
    best = min(table, key=lambda r: r["seconds"])
    default = [r for r in table if (r["lane_max"], r["hub_min"], r["blocks_per_cu"], r["push_lb"])
               == (64, 512, 8, gi.LB_TWC)]
    g.set_bins(best["lane_max"], best["hub_min"], best["blocks_per_cu"])
    out = dict(best)
    out["default_seconds"] = default[0]["seconds"] if default else None
    out["table"] = table
    return out


DELTAS = (4, 8, 16, 32, 64, 128, 256)


def tune_delta(g: gi.Graph, source, deltas=DELTAS, reps: int = 2, log=None) -> dict:
    """Fastest delta of a fused SSSP from `source` (one vertex or a list,
    whose times add up): {"delta", "ms" (the sum), "table"}."""
    sched = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWC, kernel_fusion=1)
    sources = [int(source)] if np.isscalar(source) else [int(x) for x in source]
    refs = {}
    table = []
    for d in deltas:
        total = 0.0
        for src in sources:
            gi.sssp(g, src, d, sched)  # warm-up (builds the split, records the slots)
            best = float("inf")
            for _ in range(reps):
                dist, st = gi.sssp(g, src, d, sched)
                best = min(best, st["ms"])
            if src not in refs:
                refs[src] = dist
            elif not np.array_equal(dist, refs[src]):
                raise AssertionError(f"delta {d} changed the distances from {src}")
            total += best
        table.append({"delta": d, "ms": total})
        if log:
            log(f"  delta {d}: {total:.3f} ms")
    top = min(table, key=lambda r: r["ms"])
    return {"delta": top["delta"], "ms": top["ms"], "table": table}
