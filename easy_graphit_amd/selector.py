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
happens inside the BFS decide kernel with no host synchronisation.
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

# Calibration graphs: (kind, args). None of them is the RMAT-24 test graph.
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
    ("rmat", dict(scale=24, edge_factor=16, seed=25)),
    ("uniform", dict(num_vertices=1 << 20, num_pairs=8 << 20, seed=17)),
    ("uniform", dict(num_vertices=1 << 18, num_pairs=4 << 18, seed=18)),
    ("grid", dict(width=512, num_random=2600, seed=19)),
]


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
    np.savez(path, params=model["params"], mean=model["mean"], scale=model["scale"],
             loss=model["loss"], n_out=np.int32(model["n_out"]),
             allowed=np.int32(model["allowed"]), **extra)


def load(path: str = MODEL_PATH):
    with np.load(path, allow_pickle=False) as z:
        m = {"params": z["params"], "mean": z["mean"], "scale": z["scale"],
             "loss": z["loss"], "n_out": int(z["n_out"]), "allowed": int(z["allowed"])}
        if "threshold" in z.files:
            m["threshold"] = float(z["threshold"])
        return m


def install(model):
    gi.selector_load(model["params"], model["n_out"], model["allowed"], model["mean"],
                     model["scale"])


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
    ap.add_argument("cmd", choices=["train", "eval"])
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
    else:
        model = load(args.out)
        print(json.dumps(evaluate(model, model.get("threshold", 0.0005), scale=args.scale, log=log)))


if __name__ == "__main__":
    main()
