#!/usr/bin/env python3
"""Headline benchmark: BFS GTEPS on a symmetrized RMAT scale-24 graph (edge
factor 16, a/b/c = 0.57/0.19/0.19) on one MI355X, through the C ABI.

One step = one direction-optimizing BFS from the next of 64 sources picked
GAPBS-style (degree-0 vertices skipped), graph resident in HBM, depth array
preallocated. `value` = undirected edges of the reached components summed
over the K timed BFS runs / wall time of those runs (Graph500 convention).

Also reported on the same JSON line:
  roofline      the dominant traversal kernel's algorithmic bytes per launch
                over its average duration (hipEvents around that launch),
                against 8 TB/s HBM;
  cpu_baseline  the oracle's OpenMP direction-optimizing BFS on the host's
                cores, same graph and sources (a bounded sample); each app
                section carries its own OpenMP oracle baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 27491095


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def schedule(gi, threshold: float, push_lb: int, pull_lb: int, fused: bool,
             selector: int | None = None):
    """GraphIt's hybrid schedule: s1 = push (TWCE in the reference BFS app,
    apps/bfs.cpp:69-80), s2 = pull on a bitmap, s2 when |F| > t*V."""
    f = 1 if fused else 0
    s1 = gi.Schedule(direction=gi.PUSH, load_balancing=push_lb, frontier_creation=gi.FC_SPARSE,
                     cta_size=256, kernel_fusion=f)
    s2 = gi.Schedule(direction=gi.PULL, load_balancing=pull_lb, pull_frontier_rep=gi.PULL_BITMAP,
                     frontier_creation=gi.FC_BITMAP, cta_size=256, kernel_fusion=f)
    return gi.HybridSchedule(s1, s2, threshold,
                             gi.SELECT_HEURISTIC if selector is None else selector)


def time_bfs(gi, g, sched, sources, depth, edges_of, steps):
    """Wall time of `steps` BFS runs cycling over `sources`; GTEPS."""
    gi.synchronize()
    t0 = time.perf_counter()
    total_edges = 0
    for k in range(steps):
        src = int(sources[k % len(sources)])
        gi.bfs_device(g, src, sched, depth)
        total_edges += edges_of[src]
    gi.synchronize()
    dt = time.perf_counter() - t0
    return dt, total_edges / 2 / dt / 1e9


def roofline(gi, g, sched, sources, depth):
    """Dominant step kernel of the traced BFS runs from `sources`.

    Traced runs launch one step kernel per level between two HIP events on
    the BFS stream (bfs_push_step_traced or bfs_pull_step_traced, the same
    code as the recorded kernels under a symbol of their own, so a rocprof
    summary of this command lists exactly these launches). achieved =
    algorithmic bytes per launch / average launch duration."""
    kern, modes = {}, {}
    for i, src in enumerate(sources):
        recs = gi.bfs_profile(g, int(src), sched, depth)
        if i == 0:
            for r in recs:
                log(f"  level {r['level']}: dir={r['direction']} lb={r['load_balancing']} "
                    f"|F|={r['frontier_size']} |F'|={r['next_size']} {r['kernel_ms']:.3f} ms "
                    f"{r['alg_bytes'] / max(r['kernel_ms'], 1e-9) / 1e6:.1f} GB/s")
        for r in recs:
            pull = r["direction"] == gi.PULL
            name = "bfs_pull_step_traced" if pull else "bfs_push_step_traced"
            lb = r["load_balancing"]
            mode = gi.BFS_MODE_NAMES[lb - 100] if lb >= 100 else gi.LB_NAMES[lb]
            for d, key in ((kern, name), (modes, mode)):
                k = d.setdefault(key, {"ms": 0.0, "bytes": 0.0, "launches": 0})
                k["ms"] += r["kernel_ms"]
                k["bytes"] += r["alg_bytes"]
                k["launches"] += 1
    dom = max(kern, key=lambda k: kern[k]["ms"])
    d = kern[dom]
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9

    def summary(v):
        return {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2),
                "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else 0.0}

    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": dom, "avg_launch_us": round(1e3 * avg_ms, 2),
            "alg_bytes_per_launch": int(bytes_per_launch),
            "per_kernel": {k: summary(v) for k, v in kern.items()},
            "per_mode": {k: summary(v) for k, v in modes.items()}}


def pmc_traffic(kernel: str, scale: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE runs of this benchmark, scaled by a
    plain copy of known size); counters cannot be read from inside the run."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_rmat{scale}_pmc.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1])), os.path.relpath(path, ROOT)
    return None, None


def sssp_certificate_violations(g, source: int, dist) -> int:
    """Vertices at which `dist` fails the shortest-path certificate for
    positive integer weights: dist[source] == 0, dist[v] <= dist[u] + w for
    every edge (u, v, w), and every other finite dist[v] is met with equality
    by some in-edge (a tight edge leads back to the source, so dist is at
    least the true distance; the edge inequalities make it at most that).
    Checked in int64 over the CSR, in slices of vertices."""
    off, nb, wt = g.read_csr(weights=True)
    inf = np.iinfo(np.int32).max
    d = dist.astype(np.int64)
    tight = np.zeros(len(d), bool)
    bad = int(d[source] != 0)
    step = 1 << 20
    for v0 in range(0, len(d), step):
        v1 = min(v0 + step, len(d))
        e0, e1 = int(off[v0]), int(off[v1])
        src = np.repeat(np.arange(v0, v1, dtype=np.int64), np.diff(off[v0:v1 + 1]))
        du = d[src]
        reach = du < inf
        cand = du[reach] + wt[e0:e1][reach]
        tgt = nb[e0:e1][reach]
        dt = d[tgt]
        bad += int((cand < dt).sum())
        tight[tgt[cand == dt]] = True
    need = (d < inf) & ~tight
    need[source] = False
    return bad + int(need.sum())


def alg_frac(nbytes: float, seconds: float) -> float:
    return round(nbytes / seconds / (HBM_PEAK_GBS * 1e9), 4)


def cpu_leg(fn, edges: float, sample: str, unit_scale: float = 0.5):
    """One OpenMP oracle run on this host's cores as the app's CPU baseline:
    GTEPS over `edges` (edge entries; halved for undirected GTEPS)."""
    t0 = time.perf_counter()
    used = fn()
    t = time.perf_counter() - t0
    return {"value": round(edges * unit_scale / t / 1e9, 4), "unit": "GTEPS", "cores": used, "kind": "port",
            "seconds": round(t, 3), "sample": sample}


def bench_apps(gi, g, sources, edges_of, log, only: str | None = None, cpu_on: bool = True) -> dict:
    """The other apps on their configurations, each through its fused
    schedule, with GTEPS and the fraction of the HBM roofline of their
    algorithmic bytes (one definition per app, stated with each)."""
    from easy_graphit_amd import autotune
    out = {}
    V, E = g.num_vertices, g.num_edges
    fs = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWC, kernel_fusion=1)
    if only in (None, "sssp"):
        # SSSP: RMAT-24 with int32 weights U[1,255], delta-stepping; the delta is
        # the fastest of autotune.DELTAS summed over the timed sources
        g.set_weights(SEED, 1, 255)
        srcs = [int(s) for s in sources[:4]]
        tuned = autotune.tune_delta(g, srcs, log=log)
        delta = tuned["delta"]
        t, edges, vr, steps = 0.0, 0, 0, 0
        for s in srcs:
            gi.sssp(g, s, delta, fs)  # warm-up
            dist, st = gi.sssp(g, s, delta, fs)
            t += st["ms"] * 1e-3
            edges += edges_of[s]
            vr += int((dist < np.iinfo(np.int32).max).sum())
            steps += st["iterations"]
        cpu = None
        if cpu_on:
            import oracle
            off, nb, wt = g.read_csr(weights=True)
            host = oracle.CSR(off, nb, wt)
            cpu = cpu_leg(lambda: oracle.sssp_omp(host, srcs[0], delta)[1], edges_of[srcs[0]],
                          f"oracle OpenMP delta-stepping (delta {delta}) from the first of these sources")
            del host, off, nb, wt
        out["sssp"] = {"gteps": round(edges / 2 / t / 1e9, 3), "ms_per_run": round(t * 1e3 / len(srcs), 3),
                       "cpu_baseline": cpu,
                       "roofline_frac": alg_frac(16 * vr + 12 * edges, t), "relax_steps": steps,
                       "alg_bytes": "16 V_reached + 12 E_reached", "config": "RMAT-24 sym, w U[1,255], "
                       f"delta {delta} (fastest of {list(autotune.DELTAS)} summed over these sources), "
                       f"{len(srcs)} sources",
                       "delta_ms_summed": {r["delta"]: round(r["ms"], 3) for r in tuned["table"]}}
        log(f"sssp: {out['sssp']}")
    if only == "sssp_trace":
        # the SSSP slots launched one at a time with HIP events around each
        # launch (gi_sssp_trace, no recorded graphs): the form a kernel-trace
        # profile of the SSSP kernels is taken of
        g.set_weights(SEED, 1, 255)
        res = {}
        for name, gg, src, delta in (("rmat24", g, int(sources[0]), 16),
                                     ("grid2048", gi.Graph.grid(2048, 2 * 2048 * 2047 // 100, SEED)
                                      .set_weights(SEED, 1, 1000), 0, 64)):
            dist, ms, info = gi.sssp_trace(gg, src, delta, max_slots=8192)
            bad = sssp_certificate_violations(gg, src, dist)
            if bad:
                raise SystemExit(f"traced SSSP distances fail the shortest-path certificate ({bad} vertices)")
            cols = {0: "push", 1: "pull", 2: "hub", 4: "scan_and_relax"}  # (columns 3 and 5 are unused)
            tot = ms.sum(0)
            res[name] = {"steps": len(ms), "delta": delta,
                         "kernel_ms_summed": {c: round(float(tot[i]), 3) for i, c in cols.items()},
                         "heavy_steps": int((info[:, 0] == 2).sum())}
            log(f"sssp_trace {name}: {res[name]}")
        out["sssp_trace"] = res
    if only in (None, "sssp_grid"):
        # SSSP on the grid workload: 2048 x 2048 4-neighbour grid plus 1% random
        # long edges, w U[1,1000], source at the corner, delta 64
        gr = gi.Graph.grid(2048, 2 * 2048 * 2047 // 100, SEED).set_weights(SEED, 1, 1000)
        gi.sssp(gr, 0, 64, fs)
        ts = []
        for _ in range(3):
            dist, st = gi.sssp(gr, 0, 64, fs)
            ts.append(st["ms"] * 1e-3)
        t = float(np.median(ts))
        ge = gr.num_edges
        tuned_g = autotune.tune_delta(gr, 0, deltas=(64, 256, 1024), reps=1)
        cpu = None
        if cpu_on:
            import oracle
            off, nb, wt = gr.read_csr(weights=True)
            host = oracle.CSR(off, nb, wt)
            cpu = cpu_leg(lambda: oracle.sssp_omp(host, 0, 64)[1], ge,
                          "oracle OpenMP delta-stepping, delta 64, from the corner")
        out["sssp_grid"] = {"gteps": round(ge / 2 / t / 1e9, 3), "ms_per_run": round(t * 1e3, 3),
                            "cpu_baseline": cpu,
                            "relax_steps": st["iterations"],
                            "roofline_frac": alg_frac(16 * gr.num_vertices + 12 * ge, t),
                            "config": "2048x2048 grid + 1% random edges, w U[1,1000], source 0, delta 64",
                            "delta_ms": {r["delta"]: round(r["ms"], 3) for r in tuned_g["table"]}}
        log(f"sssp_grid: {out['sssp_grid']}")
        gr.destroy()
    if only in (None, "cc"):
        # CC on RMAT-24, fused. The line's figure is the VERTEX_BASED schedule:
        # two neighbour-sampling hook passes, then rounds in which only the
        # vertices outside the most frequent label hook their edges; the
        # EDGE_ONLY full sweeps (the app's own schedule) are beside it. Both
        # give the same labels.
        fc = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_EDGE_ONLY, kernel_fusion=1)
        fcs = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_VERTEX_BASED, kernel_fusion=1)
        want_lab, _ = gi.cc(g, fc)
        lab, _ = gi.cc(g, fcs)
        if not np.array_equal(lab, want_lab):
            raise SystemExit("sampled CC labels differ from the full sweeps'")
        ts, rounds_s = [], 0
        for _ in range(3):
            _, st = gi.cc(g, fcs)
            ts.append(st["ms"] * 1e-3)
            rounds_s = st["iterations"]
        t = float(np.median(ts))
        # the edges the filtered rounds hook: those of the vertices outside the
        # giant label, from the final labels and the degrees
        off, nb, _ = g.read_csr()
        deg_h = np.diff(off.astype(np.int64))
        vals, counts = np.unique(lab, return_counts=True)
        giant = vals[np.argmax(counts)]
        e_out = int(deg_h[lab != giant].sum())
        del deg_h
        cpu = None
        if cpu_on:
            import oracle
            host = oracle.CSR(off, nb)
            cpu = cpu_leg(lambda: oracle.cc_omp(host)[1], E, "oracle OpenMP label propagation with shortcutting")
            del host
        del off, nb
        out["cc"] = {"gteps": round(E / 2 / t / 1e9, 3), "ms_per_run": round(t * 1e3, 3),
                     "filtered_rounds": rounds_s, "cpu_baseline": cpu,
                     "roofline_frac": alg_frac(2 * (20 * V + 8 * V) + rounds_s * (8 * e_out + 12 * V + 8 * V), t),
                     "alg_bytes": "per hook sweep 8 E' + 12 V over the E' edges it hooks, 8 V per shortcut: "
                                  "2 sampling sweeps (E' = V, one neighbour per vertex) and the filtered rounds "
                                  f"(E' = {e_out:,}, the edges of the vertices outside the giant label), "
                                  "each with its shortcut",
                     "config": "RMAT-24 sym, VERTEX_BASED fused schedule (sampled hooking); labels equal the full "
                               "sweeps'"}
        log(f"cc: {out['cc']}")
        gi.cc(g, fc)
        ts, rounds = [], 0
        for _ in range(3):
            _, st = gi.cc(g, fc)
            ts.append(st["ms"] * 1e-3)
            rounds = st["iterations"]
        t = float(np.median(ts))
        out["cc_full_sweeps"] = {"gteps": round(E / 2 / t / 1e9, 3), "ms_per_run": round(t * 1e3, 3),
                                 "rounds": rounds,
                                 "roofline_frac": alg_frac(rounds * (8 * E + 12 * V) + rounds * 8 * V, t),
                                 "alg_bytes": "per round 8 E + 12 V (hook) + 8 V (shortcut)",
                                 "roofline_frac_bytes_read": alg_frac(rounds * (6 * E + 12 * V) + rounds * 8 * V, t),
                                 "bytes_read": "per round 6 E + 12 V (the sweep hooks one direction of each "
                                               "symmetric pair: every neighbour id, half the label gathers) + 8 V",
                                 "config": "RMAT-24 sym, EDGE_ONLY fused schedule (the app's own): full sweeps"}
        log(f"cc_full_sweeps: {out['cc_full_sweeps']}")
    if only in (None, "pr"):
        # PageRank: RMAT-22 directed, 20 iterations, FP32 pull
        gp = gi.Graph.rmat(22, 16, SEED, False).set_layout(g.layout)
        fp = gi.Schedule(direction=gi.PULL, load_balancing=gi.LB_VERTEX_BASED, kernel_fusion=1)
        gi.pagerank(gp, 20, schedule=fp)
        ts = []
        for _ in range(3):
            _, st = gi.pagerank(gp, 20, schedule=fp)
            ts.append(st["ms"] * 1e-3)
        t = float(np.median(ts))
        Vp, Ep = gp.num_vertices, gp.num_edges
        cpu = None
        if cpu_on:
            import oracle
            off, nb, _ = gp.read_csr()
            host = oracle.CSR(off, nb)
            cpu = cpu_leg(lambda: (oracle.pagerank(host, 20), oracle.set_threads(0))[1], 20 * Ep,
                          "oracle FP64 PageRank, 20 iterations (OpenMP pull over in-neighbours)", 1.0)
            del host, off, nb
        out["pr"] = {"gteps": round(20 * Ep / t / 1e9, 3), "ms_per_run": round(t * 1e3, 3),
                     "cpu_baseline": cpu,
                     "iterations_per_s": round(20 / t, 1),
                     "roofline_frac": alg_frac(20 * (8 * Ep + 20 * Vp), t),
                     "alg_bytes": "20 (8 E + 20 V)",
                     "config": f"RMAT-22 directed, 20 iterations, d 0.85, {'degree' if gp.layout else 'native'} layout"}
        log(f"pr: {out['pr']}")
        gp.destroy()
    if only == "batched":
        # the batched engines alone (the profiled command: the BC section also
        # runs the single-source engines it compares against): BC from 256
        # sources, per-source BFS from 256 and SSSP from 64 of them
        gb = gi.Graph.rmat(20, 16, SEED, True)
        bsrc = gb.pick_sources(SEED, 256)
        Vb = gb.num_vertices
        gi.bc(gb, bsrc[:8], 256)
        _, st_b = gi.bc(gb, bsrc, 256)
        dtmp = gi.DeviceArray(Vb, np.int32)
        be, bv = 0, 0
        for s in bsrc:
            gi.bfs_device(gb, int(s), None, dtmp)
            be += gb.reached_edges(dtmp)
            bv += int((dtmp.numpy() >= 0).sum())
        tb = st_b["ms"] * 1e-3
        out["bc_batched"] = {"sources": len(bsrc), "batched_ms": round(tb * 1e3, 2),
                             "gteps": round(be / 2 / tb / 1e9, 3),
                             "roofline_frac": alg_frac(20 * bv + 20 * be + 16 * Vb * len(bsrc), tb),
                             "config": "RMAT-20 sym, 256 sources, batch 256"}
        log(f"bc_batched: {out['bc_batched']}")
        dall = gi.DeviceArray(len(bsrc) * Vb, np.int32)
        gi.bfs_batched(gb, bsrc[:64], dall)
        _, st_f = gi.bfs_batched(gb, bsrc, dall)
        out["bfs_batched"] = {"sources": len(bsrc), "ms": round(st_f["ms"], 2), "config": "RMAT-20 sym, 256 sources"}
        log(f"bfs_batched: {out['bfs_batched']}")
        gb.set_weights(SEED, 1, 255)
        ss = [int(x) for x in bsrc[:64]]
        gi.sssp_batched(gb, ss[:8], 1024, 0, dall)
        _, st_s = gi.sssp_batched(gb, ss, 1024, 0, dall)
        out["sssp_batched"] = {"sources": len(ss), "ms": round(st_s["ms"], 2), "steps": st_s["iterations"],
                               "pull_steps": st_s["pull_iterations"],
                               "config": "RMAT-20 sym, w U[1,255], 64 sources, delta 1024"}
        log(f"sssp_batched: {out['sssp_batched']}")
    if only in (None, "bc"):
        # Batched Brandes BC: RMAT-20, 256 sources, batched vs one source at a time
        gb = gi.Graph.rmat(20, 16, SEED, True)
        bsrc = gb.pick_sources(SEED, 256)
        gi.bc(gb, bsrc[:8], 256)
        _, st_b = gi.bc(gb, bsrc, 256)
        gi.bc(gb, bsrc[:2], 1)
        _, st_s = gi.bc(gb, bsrc, 1)
        fused1 = gi.Schedule(direction=gi.PUSH, load_balancing=gi.LB_TWCE, kernel_fusion=1)
        gi.bc(gb, bsrc[:2], schedule=fused1)
        _, st_f = gi.bc(gb, bsrc, schedule=fused1)
        dtmp = gi.DeviceArray(gb.num_vertices, np.int32)
        be, bv = 0, 0
        be_of = {}
        for s in bsrc:
            gi.bfs_device(gb, int(s), None, dtmp)
            be_of[int(s)] = gb.reached_edges(dtmp)
            be += be_of[int(s)]
            bv += int((dtmp.numpy() >= 0).sum())
        tb, tsq, tf = st_b["ms"] * 1e-3, st_s["ms"] * 1e-3, st_f["ms"] * 1e-3
        Vb = gb.num_vertices
        # effective bytes: per source, a push BFS (20 V_r + 8 E_r), sigma (4 E_r + 4 V), backward (8 E_r + 12 V)
        bc_bytes = 20 * bv + 20 * be + 16 * Vb * len(bsrc)
        cpu = None
        if cpu_on:
            import oracle
            off, nb, _ = gb.read_csr()
            host = oracle.CSR(off, nb)
            sample = [int(x) for x in bsrc[:32]]
            cpu = cpu_leg(lambda: (oracle.bc(host, sample), oracle.set_threads(0))[1],
                          sum(be_of[x] for x in sample),
                          f"oracle FP64 Brandes (OpenMP over sources) from {len(sample)} of the 256 sources")
        out["bc"] = {"sources": len(bsrc), "batched_ms": round(tb * 1e3, 2), "sequential_ms": round(tsq * 1e3, 2),
                     "speedup": round(tsq / tb, 2), "gteps": round(be / 2 / tb / 1e9, 3),
                     "roofline_frac": alg_frac(bc_bytes, tb),
                     "alg_bytes": "per source 20 V_reached + 20 E_reached + 16 V (forward push BFS, sigma, "
                                  "backward), effective",
                     "cpu_baseline": cpu,
                     "fused_single_source_ms": round(tf * 1e3, 2), "speedup_vs_fused_single": round(tf / tb, 2),
                     "config": "RMAT-20 sym, 256 sources: batch 256 vs one source at a time (unfused per-level "
                               "host loop, and the fused single-source mode)"}
        log(f"bc: {out['bc']}")
        # batched BFS and SSSP on the same graph and sources: one launch set per
        # step for 64 (BFS) / 256 (SSSP) sources against the fused single-source
        # engines one source at a time
        dall = gi.DeviceArray(len(bsrc) * Vb, np.int32)
        gi.bfs_batched(gb, bsrc[:64], dall)
        gi.synchronize()
        t0 = time.perf_counter()
        gi.bfs_batched(gb, bsrc, dall)
        gi.synchronize()
        t_mb = time.perf_counter() - t0
        d_b = dall.numpy().reshape(len(bsrc), Vb)
        one = autotune.schedule(gi.LB_TWC, 0.02)
        gi.bfs_device(gb, int(bsrc[0]), one, dtmp)
        gi.synchronize()
        t0 = time.perf_counter()
        for k, s in enumerate(bsrc):
            gi.bfs_device(gb, int(s), one, dtmp)
        gi.synchronize()
        t_sq = time.perf_counter() - t0
        for k in (0, 100, 255):  # rows checked against the single-source engine
            gi.bfs_device(gb, int(bsrc[k]), one, dtmp)
            if not np.array_equal(dtmp.numpy(), d_b[k]):
                raise SystemExit("batched BFS depths differ from the single-source engine's")
        out["bfs_batched"] = {"sources": len(bsrc), "ms": round(t_mb * 1e3, 2), "one_at_a_time_ms": round(t_sq * 1e3, 2),
                              "speedup": round(t_sq / t_mb, 2), "gteps": round(be / 2 / t_mb / 1e9, 3),
                              "config": "RMAT-20 sym, the 256 BC sources; depth[256][V] in 4 passes of 64 sources "
                                        "vs the fused single-source engine one source at a time"}
        log(f"bfs_batched: {out['bfs_batched']}")
        gb.set_weights(SEED, 1, 255)
        ss = [int(x) for x in bsrc[:64]]
        # the window width is a schedule choice (the distances do not depend
        # on it): each engine runs at the best of a few widths, and at 32 side
        # by side; a pass of the batched engine shares one window among its 64
        # sources, so it favours wider windows than one source does
        def seq_ms(delta):
            gi.sssp(gb, ss[0], delta, fs)
            return sum(gi.sssp(gb, s, delta, fs)[1]["ms"] for s in ss)

        seq = {d: seq_ms(d) for d in (16, 32)}
        bat = {}
        for d in (32, 256, 1024):
            gi.sssp_batched(gb, ss[:8], d, 0, dall)
            bat[d] = gi.sssp_batched(gb, ss, d, 0, dall)[1]["ms"]
        ds, db = min(seq, key=seq.get), min(bat, key=bat.get)
        _, st_sb = gi.sssp_batched(gb, ss, db, 0, dall)
        d_sb = dall.numpy()[: len(ss) * Vb].reshape(len(ss), Vb)
        for k in (0, 5, 63):  # rows checked against the single-source engine
            d1, _ = gi.sssp(gb, ss[k], ds, fs)
            if not np.array_equal(d1, d_sb[k]):
                raise SystemExit("batched SSSP distances differ from the single-source engine's")
        t_sb = bat[db] * 1e-3
        out["sssp_batched"] = {"sources": len(ss), "ms": round(bat[db], 2), "delta": db,
                               "one_at_a_time_ms": round(seq[ds], 2), "one_at_a_time_delta": ds,
                               "speedup": round(seq[ds] / bat[db], 2),
                               "at_delta_32": {"ms": round(bat[32], 2), "one_at_a_time_ms": round(seq[32], 2),
                                               "speedup": round(seq[32] / bat[32], 2)},
                               "steps": st_sb["iterations"], "pull_steps": st_sb["pull_iterations"],
                               "gteps": round(sum(be_of[x] for x in ss) / 2 / t_sb / 1e9, 3),
                               "config": "RMAT-20 sym, w U[1,255], 64 of the BC sources in one pass (delta: best of "
                                         "32/256/1024) vs the fused single-source engine one source at a time "
                                         "(delta: best of 16/32)"}
        log(f"sssp_batched: {out['sssp_batched']}")
        gb.destroy()
    if only in (None, "many"):
        # 1,024 uniform graphs (V log-uniform in [1k, 64k], average degree 8), one
        # fused multi-source BFS over the packed CSR with autotuned bins
        gm, offs = gi.Graph.uniform_batch(1024, 1024, 65536, 8, SEED)
        moff, _, _ = gm.read_csr()
        msrc = []
        for i in range(1024):  # first vertex with edges in each graph
            deg = np.diff(moff[offs[i]:offs[i + 1] + 1])
            nz = np.flatnonzero(deg)
            msrc.append(int(offs[i] + (nz[0] if len(nz) else 0)))
        msrc = np.array(msrc, np.int32)
        tuned = autotune.tune_bins(gm, msrc)
        sched = autotune.schedule(tuned["push_lb"])
        depth = gi.DeviceArray(gm.num_vertices, np.int32)
        t = autotune.time_multi(gm, msrc, sched, depth, 5)
        me = gm.reached_edges(depth)
        # one graph at a time: 32 of the graphs built on their own, scaled to 1,024
        rng = np.random.default_rng(SEED)
        pick = rng.choice(1024, 32, replace=False)
        mnb = None
        tone = 0.0
        for i in pick:
            v0, v1 = int(offs[i]), int(offs[i + 1])
            if mnb is None:
                _, mnb, _ = gm.read_csr()
            sub_off = moff[v0:v1 + 1] - moff[v0]
            sub_nb = mnb[moff[v0]:moff[v1]] - v0
            gs = gi.Graph.from_csr(sub_off, sub_nb)
            ds = gi.DeviceArray(v1 - v0, np.int32)
            sc1 = autotune.schedule(gi.LB_TWC)
            gi.bfs_device(gs, int(msrc[i] - v0), sc1, ds)
            gi.synchronize()
            t0 = time.perf_counter()
            gi.bfs_device(gs, int(msrc[i] - v0), sc1, ds)
            gi.synchronize()
            tone += time.perf_counter() - t0
            gs.destroy()
        tone *= 1024 / len(pick)
        out["many_graphs"] = {"graphs": 1024, "ms": round(t * 1e3, 3), "graphs_per_s": round(1024 / t, 1),
                              "gteps": round(me / 2 / t / 1e9, 3),
                              "one_at_a_time_ms": round(tone * 1e3, 2), "speedup": round(tone / t, 2),
                              "bins": {k: tuned[k] for k in ("lane_max", "hub_min", "blocks_per_cu", "push_lb")},
                              "default_bins_ms": round(tuned["default_seconds"] * 1e3, 3)
                              if tuned["default_seconds"] else None,
                              "config": "1,024 uniform graphs, V log-uniform [1k,64k], avg degree 8, packed CSR"}
        log(f"many graphs: {out['many_graphs']}")
        gm.destroy()
    return out


def cpu_baseline(gi, g, sources, edges_of, budget_s: float):
    """The oracle's direction-optimizing OpenMP BFS (Beamer's rule, alpha 15,
    beta 18) on all of OpenMP's threads of this host, over the same graph and
    sources, until the budget is spent (the reference has no CPU BFS)."""
    import oracle
    off, nb, _ = g.read_csr()
    host = oracle.CSR(off, nb)
    done, edges, t, used = 0, 0, 0.0, 1
    for src in sources:
        t0 = time.perf_counter()
        _, used = oracle.bfs_depth_omp(host, int(src))
        t += time.perf_counter() - t0
        edges += edges_of[int(src)]
        done += 1
        if t >= budget_s:
            break
    return {"value": round(edges / 2 / t / 1e9, 4), "unit": "GTEPS", "cores": used, "kind": "port",
            "sample": f"oracle direction-optimizing BFS (OpenMP, {used} threads) from {done} of the same "
                      f"sources on the same graph ({t:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--threshold", type=float, default=None,
                    help="hybrid threshold t (default: the one tuned with the selector)")
    ap.add_argument("--engine", choices=["fused", "unfused"], default="fused")
    ap.add_argument("--push-lb", type=int, default=2, help="push schedule (default TWCE)")
    ap.add_argument("--pull-lb", type=int, default=0, help="pull schedule (default VERTEX_BASED)")
    ap.add_argument("--selector", choices=["learned", "heuristic"], default="learned")
    ap.add_argument("--layout", choices=["degree", "native"], default="degree",
                    help="vertex layout the fused engines run on (results are in original ids)")
    ap.add_argument("--profile-sources", type=int, default=4)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-apps", action="store_true", help="BFS line only (no SSSP/CC/PR/BC/many-graph runs)")
    ap.add_argument("--no-selector", action="store_true",
                    help="skip the fixed-rule comparison (profiles of the timed BFS runs alone)")
    ap.add_argument("--sources", type=int, default=64, help="distinct BFS sources cycled over")
    ap.add_argument("--app", choices=["sssp", "sssp_grid", "sssp_trace", "cc", "pr", "bc", "batched", "many"],
                    default=None,
                    help="only this app's section (for profiling; prints its results, no bench line)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run three device-to-device copies of MIB MiB (HBM counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the multi-GPU "
              f"sharding is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    import easy_graphit_amd as gi
    if gi.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.calibrate_copy > 0:
        n = args.calibrate_copy << 18  # int32 elements
        a, b = gi.DeviceArray(n, np.int32), gi.DeviceArray(n, np.int32)
        for _ in range(3):
            gi.copy_device(b, a)
        gi.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 3}), flush=True)
        return

    t0 = time.perf_counter()
    g = gi.Graph.rmat(args.scale, args.edge_factor, SEED, True)
    V, E = g.num_vertices, g.num_edges
    log(f"graph RMAT-{args.scale} sym: V={V} E={E} built in {time.perf_counter() - t0:.2f}s")
    layout = gi.LAYOUT_DEGREE if args.layout == "degree" else gi.LAYOUT_NATIVE
    t0 = time.perf_counter()
    g.set_layout(layout)
    gi.synchronize()
    layout_s = time.perf_counter() - t0
    log(f"layout {args.layout}: {layout_s:.2f}s")
    sources = g.pick_sources(SEED, 64)[: max(1, min(64, args.sources))]
    ns = len(sources)
    depth = gi.DeviceArray(V, np.int32)

    # the learned selector (trained on calibration graphs, not this one) and
    # GraphIt's fixed hybrid rule at the threshold tuned on the same graphs
    from easy_graphit_amd import selector as sel_mod
    model = sel_mod.load() if os.path.exists(sel_mod.MODEL_PATH) else None
    threshold = args.threshold
    if threshold is None:
        threshold = float(model["threshold"]) if model is not None and "threshold" in model else 0.0005
    fused = args.engine == "fused"
    heur = schedule(gi, threshold, args.push_lb, args.pull_lb, fused)
    learned = None
    if model is not None and fused:
        sel_mod.install(model)
        learned = schedule(gi, threshold, args.push_lb, args.pull_lb, fused, gi.SELECT_LEARNED)
    use_learned = args.selector == "learned" and learned is not None
    sched = learned if use_learned else heur

    # untimed: reached edges per source; both policies must give equal depths
    edges_of = {}
    t0 = time.perf_counter()
    for i, src in enumerate(sources):
        gi.bfs_device(g, int(src), sched, depth)
        edges_of[int(src)] = g.reached_edges(depth)
        if i < 4 and learned is not None:
            want = depth.numpy()
            gi.bfs_device(g, int(src), heur if use_learned else learned, depth)
            if not np.array_equal(want, depth.numpy()):
                raise SystemExit(f"learned and heuristic BFS disagree from source {src}")
        if i % 16 == 15:
            log(f"reference runs {i + 1}/{ns} ({time.perf_counter() - t0:.1f}s)")
    if args.app:
        print(json.dumps({"app": args.app, "results": bench_apps(gi, g, sources, edges_of, log, args.app,
                                                                 cpu_on=False)}),
              flush=True)
        return
    for k in range(args.warmup):
        gi.bfs_device(g, int(sources[k % ns]), sched, depth)
    gi.synchronize()

    dt, gteps = time_bfs(gi, g, sched, sources, depth, edges_of, args.steps)
    log(f"{args.steps} BFS in {dt * 1e3:.1f} ms -> {gteps:.2f} GTEPS "
        f"({'learned selector' if use_learned else 'heuristic'})")
    selector_cmp = None
    if learned is not None and not args.no_selector:
        # GraphIt's fixed rule |F| <= t*V, timed on both layouts: the
        # strongest rule (the push/pull mode pair and t tuned on the
        # calibration graphs, stored with the model) and the BFS app's own rule
        # (PUSH TWCE / PULL VERTEX_BASED); the headline ratio is against the
        # strongest rule on whichever layout suits it best
        rules = {"strongest": sel_mod.rule_schedule(*model["rule"]) if "rule" in model else heur, "app": heur}
        by_layout = {}
        g_learned = {}
        for lay_name, lay in (("degree", gi.LAYOUT_DEGREE), ("native", gi.LAYOUT_NATIVE)):
            g.set_layout(lay)
            g_learned[lay_name] = gteps if (lay == layout and use_learned) else None
            if g_learned[lay_name] is None:
                for k in range(args.warmup):
                    gi.bfs_device(g, int(sources[k % ns]), learned, depth)
                g_learned[lay_name] = time_bfs(gi, g, learned, sources, depth, edges_of, args.steps)[1]
            for name, rs in rules.items():
                gi.bfs_device(g, int(sources[0]), rs, depth)
                _, gr = time_bfs(gi, g, rs, sources, depth, edges_of, min(args.steps, 16))
                by_layout.setdefault(name, {})[lay_name] = round(gr, 3)
        g.set_layout(layout)
        g_l = g_learned[args.layout]
        best_rule = max(by_layout["strongest"].values())
        rule_desc = (f"push {gi.BFS_MODE_NAMES[model['rule'][0]]}, pull {gi.BFS_MODE_NAMES[model['rule'][1]]}, "
                     f"t={model['rule'][2]:g}") if "rule" in model else "the app's rule"
        selector_cmp = {"learned_gteps": round(g_l, 3),
                        "learned_gteps_by_layout": {k: round(v, 3) for k, v in g_learned.items()},
                        "heuristic_gteps": best_rule,
                        "heuristic_gteps_by_layout": by_layout["strongest"],
                        "speedup": round(g_l / best_rule, 3),
                        "speedup_same_layout": round(g_l / by_layout["strongest"][args.layout], 3),
                        "app_rule_gteps_by_layout": by_layout["app"],
                        "speedup_vs_app_rule": round(g_l / max(by_layout["app"].values()), 3),
                        "heuristic": f"GraphIt hybrid |F| <= t*V with the mode pair and threshold tuned on the "
                                     f"calibration graphs ({rule_desc}), on its best layout",
                        "app_rule": f"PUSH {gi.LB_NAMES[args.push_lb]} else PULL {gi.LB_NAMES[args.pull_lb]} "
                                    f"bitmap, t={threshold:g} (apps/bfs.cpp:69-80)",
                        "learned": "MLP 16-64-64-7 (BF16 weights, FP32 accumulation) in the decide "
                                   "kernel, trained on the production step kernel's timings over calibration "
                                   "RMAT/uniform/grid graphs (none of them RMAT-24)"}
        log(f"selector: {selector_cmp}")

    roof = roofline(gi, g, sched, sources[: args.profile_sources], depth) if args.profile_sources > 0 else None
    if roof is not None:
        roof["traffic"], roof["traffic_source"] = pmc_traffic(roof["kernel"], args.scale)
    log(f"roofline: {json.dumps(roof)}")
    apps = None if args.no_apps else bench_apps(gi, g, sources, edges_of, log, cpu_on=not args.no_cpu)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(gi, g, sources, edges_of, args.cpu_budget)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "GTEPS per app on RMAT-24 (1 GPU); learned vs heuristic schedule speedup",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded RMAT generated on the device)",
        "config": {"workload": f"direction-optimizing BFS, symmetrized RMAT scale {args.scale} "
                               f"edge factor {args.edge_factor}, {ns} GAPBS-style sources",
                   "vertices": V, "edge_entries": E, "layout": args.layout,
                   "layout_build_s": round(layout_s, 3), "schedule": {
                       "s1": repr(sched.s1), "s2": repr(sched.s2), "threshold": threshold,
                       "selector": "learned" if use_learned else "heuristic",
                       "engine": args.engine}},
        "selector": selector_cmp,
        "apps": apps,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
