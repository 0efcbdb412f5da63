"""Training pipeline of the learned BFS schedule selector.

1. record(): run the dataset recorder (gi_bfs_sweep) over calibration
   graphs — synthetic RMAT, uniform and grid graphs that are not the test
   graph — timing every fused step mode at every level from the same state.
2. fit(): normalise the 16 features, regress log(step time) of every mode
   with the device MLP (gi_selector_train, FP32 forward/backward/Adam
   kernels), and return the model.
3. tune_threshold(): the fixed hybrid rule's threshold t, chosen by a sweep
   on the same calibration graphs (the baseline the selector is judged
   against).
4. save()/load(): an .npz of plain arrays (numpy, allow_pickle=False).

The model is loaded into the runtime with gi_selector_load; inference then
happens on the device at the start of every BFS step launch, with no host
synchronisation.
"""
from __future__ import annotations

import math
import os
import time

import numpy as np

import easy_graphit_amd as gi

MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_model.npz")
N_OUT = gi.BFS_MODES
ALL_MODES = (1 << N_OUT) - 1

# Calibration graphs: (kind, args). None of them is a test graph: not the
# RMAT-24 benchmark graph (seed 27491095) nor anything of its scale, and not
# HELD_OUT below.
CALIBRATION = [
    ("rmat", dict(scale=16, edge_factor=16, seed=11)),
    ("rmat", dict(scale=18, edge_factor=16, seed=12)),
    ("rmat", dict(scale=20, edge_factor=16, seed=13)),
    ("rmat", dict(scale=19, edge_factor=8, seed=14)),
    ("rmat", dict(scale=20, edge_factor=32, seed=15)),
    ("rmat", dict(scale=21, edge_factor=16, seed=16)),
    ("rmat", dict(scale=22, edge_factor=16, seed=20)),
    ("rmat", dict(scale=22, edge_factor=8, seed=21)),
    ("rmat", dict(scale=23, edge_factor=16, seed=24)),
    ("uniform", dict(num_vertices=1 << 20, num_pairs=8 << 20, seed=17)),
    ("uniform", dict(num_vertices=1 << 18, num_pairs=4 << 18, seed=18)),
    ("grid", dict(width=512, num_random=2600, seed=19)),
]

# The graph the gain test holds out of calibration (and of the rule's tuning).
HELD_OUT = ("rmat", dict(scale=22, edge_factor=24, seed=31))

# Fused step modes and the schedule that selects each (fused_mode_of).
MODE_SCHED = {0: (gi.PUSH, gi.LB_VERTEX_BASED), 1: (gi.PUSH, gi.LB_WM), 2: (gi.PUSH, gi.LB_TWCE),
              3: (gi.PUSH, gi.LB_CM), 4: (gi.PUSH, gi.LB_TWC), 5: (gi.PULL, gi.LB_VERTEX_BASED),
              6: (gi.PULL, gi.LB_TWCE)}
PUSH_MODES = (0, 1, 2, 3, 4)
PULL_MODES = (5, 6)


# Vertex layout the engine runs on; the selector is trained and judged on it.
LAYOUT = gi.LAYOUT_DEGREE


def build_graph(kind: str, args: dict, layout: int = LAYOUT) -> gi.Graph:
    if kind == "rmat":
        g = gi.Graph.rmat(args["scale"], args["edge_factor"], args["seed"], True)
    elif kind == "uniform":
        g = gi.Graph.uniform(args["num_vertices"], args["num_pairs"], args["seed"], True)
    elif kind == "grid":
        g = gi.Graph.grid(args["width"], args["num_random"], args["seed"])
    else:
        raise ValueError(kind)
    return g.set_layout(layout)


def record(specs=CALIBRATION, sources_per_graph: int = 4, reps: int = 2, log=print):
    """Features [N,16] and per-mode step times [N,7] (ms) over all levels."""
    feats, times = [], []
    for kind, args in specs:
        g = build_graph(kind, args)
        srcs = g.pick_sources(args["seed"] * 7 + 1, sources_per_graph)
        for i, s in enumerate(srcs):
            # follow the fastest mode, and once a fixed hybrid-like path, so the
            # dataset covers the states both policies visit
            f, t = gi.bfs_sweep(g, int(s), ALL_MODES, -1 if i % 2 == 0 else 4, reps)
            feats.append(f)
            times.append(t)
        log(f"recorded {kind} {args}: {sum(len(x) for x in feats)} levels so far")
        g.destroy()
    return np.concatenate(feats), np.concatenate(times)


def targets(times: np.ndarray):
    """log(ms) targets and per-entry weights: 0 for modes not measured, else
    the level's best time relative to the mean best (clipped to
    [0.05, 20]), so that the levels where the choice costs time dominate
    the fit."""
    measured = times > 0
    t = np.where(measured, np.log(np.maximum(times, 1e-4)), 0.0).astype(np.float32)
    best = np.where(measured, times, np.inf).min(axis=1)
    level_w = np.clip(best / best.mean(), 0.05, 20.0)
    w = (measured * level_w[:, None]).astype(np.float32)
    return t, w


def init_params(rng, n_out: int = N_OUT):
    return np.concatenate([
        (rng.normal(size=(16, 64)) * math.sqrt(2.0 / 16)).reshape(-1), np.zeros(64),
        (rng.normal(size=(64, 64)) * math.sqrt(2.0 / 64)).reshape(-1), np.zeros(64),
        (rng.normal(size=(64, n_out)) * math.sqrt(1.0 / 64)).reshape(-1), np.zeros(n_out),
    ]).astype(np.float32)


def fit(feats, times, epochs: int = 3000, lr: float = 3e-3, seed: int = 0):
    mean = feats.mean(axis=0)
    std = feats.std(axis=0)
    scale = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 0.0).astype(np.float32)
    scale[15] = 0.0  # the constant feature carries no information once centred
    X = ((feats - mean) * scale).astype(np.float32)
    T, M = targets(times)
    p0 = init_params(np.random.default_rng(seed))
    params, hist = gi.selector_train(X, T, M, p0, N_OUT, epochs, lr, 1e-5)
    return {"params": params, "mean": mean.astype(np.float32), "scale": scale,
            "loss": hist, "n_out": N_OUT, "allowed": ALL_MODES}


def predict(model, feats):
    """Host-side forward pass in FP32 (for offline choice accuracy only)."""
    p = model["params"]
    n_out = model["n_out"]
    i = 0

    def take(n, shape):
        nonlocal i
        a = p[i:i + n].reshape(shape)
        i += n
        return a

    w1, b1 = take(16 * 64, (16, 64)), take(64, (64,))
    w2, b2 = take(64 * 64, (64, 64)), take(64, (64,))
    w3, b3 = take(64 * n_out, (64, n_out)), take(n_out, (n_out,))
    x = (feats - model["mean"]) * model["scale"]
    h = np.maximum(x @ w1 + b1, 0)
    h = np.maximum(h @ w2 + b2, 0)
    return h @ w3 + b3


def save(model, path: str = MODEL_PATH):
    extra = {"threshold": np.float32(model["threshold"])} if "threshold" in model else {}
    if "rule" in model:
        extra["rule"] = np.array(model["rule"], np.float64)
    np.savez(path, params=model["params"], mean=model["mean"], scale=model["scale"],
             loss=model["loss"], n_out=np.int32(model["n_out"]),
             allowed=np.int32(model["allowed"]), **extra)


def load(path: str = MODEL_PATH):
    with np.load(path, allow_pickle=False) as z:
        m = {"params": z["params"], "mean": z["mean"], "scale": z["scale"],
             "loss": z["loss"], "n_out": int(z["n_out"]), "allowed": int(z["allowed"])}
        if "threshold" in z.files:
            m["threshold"] = float(z["threshold"])
        if "rule" in z.files:  # the strongest fixed rule: push mode, pull mode, t
            r = z["rule"]
            m["rule"] = (int(r[0]), int(r[1]), float(r[2]))
        return m


def install(model):
    gi.selector_load(model["params"], model["n_out"], model["allowed"], model["mean"],
                     model["scale"])


def rule_schedule(push_mode: int, pull_mode: int, threshold: float,
                  selector: int = gi.SELECT_HEURISTIC):
    """GraphIt's hybrid rule (|F| <= t*V: push, else pull) over two fused
    step modes."""
    return fused_schedule(MODE_SCHED[push_mode][1], MODE_SCHED[pull_mode][1], threshold, selector)


def fused_schedule(push_mode_lb: int = gi.LB_TWCE, pull_lb: int = gi.LB_VERTEX_BASED,
                   threshold: float = 0.05, selector: int = gi.SELECT_HEURISTIC):
    s1 = gi.Schedule(direction=gi.PUSH, load_balancing=push_mode_lb, kernel_fusion=1,
                     cta_size=256)
    s2 = gi.Schedule(direction=gi.PULL, load_balancing=pull_lb, kernel_fusion=1,
                     pull_frontier_rep=gi.PULL_BITMAP, frontier_creation=gi.FC_BITMAP,
                     cta_size=256)
    return gi.HybridSchedule(s1, s2, threshold, selector)


def time_policy(g: gi.Graph, sched, sources, depth, reps: int = 1) -> float:
    """Wall seconds for BFS from every source (after one warm-up run)."""
    gi.bfs_device(g, int(sources[0]), sched, depth)
    gi.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for s in sources:
            gi.bfs_device(g, int(s), sched, depth)
    gi.synchronize()
    return time.perf_counter() - t0


THRESHOLDS = (0.0002, 0.0005, 0.001, 0.003, 0.01, 0.02, 0.05, 0.1, 0.2)


def tune_threshold(specs=CALIBRATION, grid=THRESHOLDS, sources_per_graph: int = 4,
                   push_lb: int = gi.LB_TWCE, log=print):
    """GraphIt's rule |F| <= t*V: pick t by total calibration time."""
    total = {t: 0.0 for t in grid}
    for kind, args in specs:
        g = build_graph(kind, args)
        srcs = g.pick_sources(args["seed"] * 7 + 1, sources_per_graph)
        depth = gi.DeviceArray(g.num_vertices, np.int32)
        for t in grid:
            total[t] += time_policy(g, fused_schedule(push_lb, threshold=t), srcs, depth)
        g.destroy()
    best = min(total, key=total.get)
    log(f"threshold sweep: {total} -> {best}")
    return best, total


def tune_rule(specs=CALIBRATION, grid=THRESHOLDS, push_modes=PUSH_MODES, pull_modes=PULL_MODES,
              sources_per_graph: int = 4, log=print):
    """The strongest fixed rule: GraphIt's |F| <= t*V rule over every pair
    of a push and a pull step mode and every threshold of `grid`, the triple
    with the least total time over the calibration graphs.
    Returns ((push_mode, pull_mode, t), {triple: seconds})."""
    total = {}
    for kind, args in specs:
        g = build_graph(kind, args)
        srcs = g.pick_sources(args["seed"] * 7 + 1, sources_per_graph)
        depth = gi.DeviceArray(g.num_vertices, np.int32)
        for pm in push_modes:
            for qm in pull_modes:
                for t in grid:
                    key = (pm, qm, t)
                    total[key] = total.get(key, 0.0) + time_policy(g, rule_schedule(pm, qm, t), srcs, depth)
        g.destroy()
        log(f"rule sweep: {kind} {args} done")
    best = min(total, key=total.get)
    log(f"strongest fixed rule on the calibration graphs: push mode {best[0]}, pull mode {best[1]}, "
        f"t={best[2]} ({total[best]:.4f} s)")
    return best, total


def compare(model, g: gi.Graph, srcs, rule, app_threshold: float, grid=THRESHOLDS, reps: int = 3,
            test_tuned: bool = True, log=print):
    """Learned selector vs fixed rules on graph g (the caller's layout):
    the calibration-tuned strongest rule `rule` = (push mode, pull mode, t),
    the reference app's rule (PUSH TWCE / PULL VERTEX_BASED,
    apps/bfs.cpp:69-80) at `app_threshold`, and, when test_tuned, the best
    rule of every mode pair and threshold on g itself (tuned on the test
    graph: an upper bound on what any fixed rule reaches). Depths of every
    policy are checked equal to the learned run's before timing."""
    depth = gi.DeviceArray(g.num_vertices, np.int32)
    install(model)
    learned = fused_schedule(threshold=app_threshold, selector=gi.SELECT_LEARNED)
    strongest = rule_schedule(*rule)
    app = rule_schedule(2, 5, app_threshold)
    for s in srcs:
        d_l, _, _ = gi.bfs(g, int(s), learned)
        for other in (strongest, app):
            d_o, _, _ = gi.bfs(g, int(s), other)
            if not np.array_equal(d_o, d_l):
                raise AssertionError(f"schedules disagree from source {s}")
    times = {"learned": [], "strongest_rule": [], "app_rule": []}
    for _ in range(reps):  # interleaved, so drift on the device hits all alike
        times["learned"].append(time_policy(g, learned, srcs, depth))
        times["strongest_rule"].append(time_policy(g, strongest, srcs, depth))
        times["app_rule"].append(time_policy(g, app, srcs, depth))
    t = {k: min(v) for k, v in times.items()}
    res = {"learned_s": t["learned"], "strongest_rule_s": t["strongest_rule"], "app_rule_s": t["app_rule"],
           "speedup_vs_strongest_rule": t["strongest_rule"] / t["learned"],
           "speedup_vs_app_rule": t["app_rule"] / t["learned"],
           "strongest_rule": {"push_mode": gi.BFS_MODE_NAMES[rule[0]], "pull_mode": gi.BFS_MODE_NAMES[rule[1]],
                              "threshold": rule[2]}}
    if test_tuned:
        best, best_t = None, float("inf")
        for pm in PUSH_MODES:
            for qm in PULL_MODES:
                for th in grid:
                    tt = min(time_policy(g, rule_schedule(pm, qm, th), srcs, depth) for _ in range(2))
                    if tt < best_t:
                        best, best_t = (pm, qm, th), tt
                log(f"  test-tuned sweep: push {pm} pull {qm} done (best so far {best} {best_t:.5f} s)")
        res["test_tuned_rule"] = {"push_mode": gi.BFS_MODE_NAMES[best[0]],
                                  "pull_mode": gi.BFS_MODE_NAMES[best[1]], "threshold": best[2],
                                  "seconds": best_t}
        res["speedup_vs_test_tuned_rule"] = best_t / t["learned"]
    log(f"compare: {res}")
    return res


def evaluate(model, threshold: float, scale: int = 24, edge_factor: int = 16,
             seed: int = 27491095, num_sources: int = 16, grid=THRESHOLDS, log=print):
    """Learned selector vs GraphIt's fixed hybrid rule on a test graph that
    is not a calibration graph.

    The rule is the reference BFS app's "power" schedule (apps/bfs.cpp:69-80:
    PUSH TWCE, else PULL VERTEX_BASED on a bitmap when |F| > t*V), timed at
    the threshold tuned on the calibration graphs and at every threshold of
    `grid` (so the best fixed t for this very graph is reported too). The
    same rule with this runtime's grid-binned TWC push is timed as well, as
    the strongest fixed schedule. Depths of every policy are checked equal to
    the learned run's on every source before timing."""
    g = gi.Graph.rmat(scale, edge_factor, seed, True).set_layout(LAYOUT)
    srcs = g.pick_sources(seed, num_sources)
    depth = gi.DeviceArray(g.num_vertices, np.int32)
    learned = fused_schedule(threshold=threshold, selector=gi.SELECT_LEARNED)
    install(model)
    heur = fused_schedule(gi.LB_TWCE, threshold=threshold)
    levels = {}
    for s in srcs:
        d_l, _, st_l = gi.bfs(g, int(s), learned)
        d_h, _, st_h = gi.bfs(g, int(s), heur)
        d_c, _, _ = gi.bfs(g, int(s), fused_schedule(gi.LB_TWC, threshold=threshold))
        if not (np.array_equal(d_h, d_l) and np.array_equal(d_c, d_l)):
            raise AssertionError(f"schedules disagree from source {s}")
        levels[int(s)] = {"heuristic_push_pull": (st_h["push_iterations"], st_h["pull_iterations"]),
                          "learned_push_pull": (st_l["push_iterations"], st_l["pull_iterations"])}
    t_l = min(time_policy(g, learned, srcs, depth) for _ in range(3))
    t_h = min(time_policy(g, heur, srcs, depth) for _ in range(3))
    sweep = {"twce": {}, "twc": {}}
    for name, lb in (("twce", gi.LB_TWCE), ("twc", gi.LB_TWC)):
        for t in grid:
            sweep[name][t] = min(time_policy(g, fused_schedule(lb, threshold=t), srcs, depth)
                                 for _ in range(2))
    best_twce = min(sweep["twce"], key=sweep["twce"].get)
    best_twc = min(sweep["twc"], key=sweep["twc"].get)
    res = {"graph": f"RMAT-{scale} EF{edge_factor} seed {seed}", "sources": num_sources,
           "threshold": threshold, "heuristic_s": t_h, "learned_s": t_l,
           "speedup": t_h / t_l,
           "best_test_threshold": best_twce,
           "speedup_vs_best_threshold": sweep["twce"][best_twce] / t_l,
           "twc_best_test_threshold": best_twc,
           "speedup_vs_best_twc_schedule": sweep["twc"][best_twc] / t_l,
           "threshold_sweep_s": {k: {str(t): v for t, v in d.items()} for k, d in sweep.items()},
           "levels": {str(k): v for k, v in levels.items()}}
    log(f"evaluate: {res}")
    return res


def main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train / evaluate the BFS schedule selector")
    ap.add_argument("cmd", choices=["train", "eval", "compare"])
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--dataset", default="")
    ap.add_argument("--out", default=MODEL_PATH)
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.cmd == "train":
        feats, times = record(sources_per_graph=args.sources, log=log)
        if args.dataset:
            np.savez(args.dataset, feats=feats, times=times)
        model = fit(feats, times, epochs=args.epochs)
        t_best, sweep = tune_threshold(log=log)
        model["threshold"] = t_best
        rule, _ = tune_rule(log=log)
        model["rule"] = rule
        save(model, args.out)
        # offline choice quality on the training set
        pred = predict(model, feats)
        measured = np.where(times > 0, times, np.inf)
        pick = np.argmin(np.where(times > 0, pred, np.inf), axis=1)
        best = measured.min(axis=1)
        regret = float((measured[np.arange(len(pick)), pick] - best).sum() / best.sum())
        print(json.dumps({"levels": int(len(feats)), "final_loss": float(model["loss"][-1]),
                          "first_loss": float(model["loss"][0]), "threshold": t_best,
                          "threshold_sweep": {str(k): v for k, v in sweep.items()},
                          "train_regret": regret}))
    elif args.cmd == "compare":  # held-out graph and the benchmark graph
        model = load(args.out)
        out = {}
        for name, (kind, gargs) in (("held_out", HELD_OUT),
                                     ("rmat24", ("rmat", dict(scale=24, edge_factor=16, seed=27491095)))):
            g = build_graph(kind, gargs)
            srcs = g.pick_sources(gargs["seed"], 16)
            out[name] = compare(model, g, srcs, model["rule"], model.get("threshold", 0.0005), log=log)
            g.destroy()
        print(json.dumps(out))
    else:
        model = load(args.out)
        print(json.dumps(evaluate(model, model.get("threshold", 0.0005), scale=args.scale, log=log)))


if __name__ == "__main__":
    main()
