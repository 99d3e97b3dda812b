"""Training data, training and evaluation of the learned engine selector.

The selector (an 8->32->32->2 FP32 MLP, include/grazelle_selector.h) predicts
the log edge-phase time of the pull and push engines from per-iteration
frontier statistics.  Its training data comes from the engine itself: the
probe engine (GZ_ENGINE_PROBE) runs every iteration of a search twice from the
same state, once per engine, records both times with the features the
selector would see, and continues with the fixed-threshold choice.  A BFS
visits the same frontiers whichever engine runs each level, so one probed
search gives the per-level cost of both engines along the only trajectory
there is, and any policy's total can be read off it.

Training minimises the squared error of the predicted log times with the
library's own device kernels (gz_selector_train: forward, backward and Adam
on the GPU).  `fit --per-edge` regresses log ms - log E instead (the choice
depends only on the difference of the two outputs); which of the two is
committed is decided on held-out graphs of the training distribution.

    python selector_train.py collect --graphs 200 --out data.npz
    python selector_train.py fit --data data.npz --out selector_bfs.bin
    python selector_train.py evaluate --data held.npz --weights selector_bfs.bin
    python selector_train.py collect --graphs 20 --first-seed 90000 --wide --out wide.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import grazelle as gz  # noqa: E402

DEFAULT_WEIGHTS = os.path.join(HERE, "selector_bfs.bin")


def training_graphs(count: int, first_seed: int = 1000):
    """(scale, edge factor, seed) of the training graphs: scales 18..22 and
    edge factors 8 / 16 / 24 / 32 in turn, seeds first_seed, first_seed+1, ..."""
    efs = (8, 16, 24, 32)
    return [(18 + i % 5, efs[(i // 5) % 4], first_seed + i) for i in range(count)]


def pick_roots(count: int, seed: int) -> list[int]:
    """Seeded roots among the loaded graph's vertices with out-degree > 0
    (original ids)."""
    od = gz.outdegrees()
    cand_dev = np.nonzero(od > 0)[0]
    rng = np.random.default_rng(seed)
    picks = rng.choice(cand_dev.shape[0], size=min(count, cand_dev.shape[0]), replace=False)
    # outdegrees() is indexed by device vertex; map back to original ids
    orig = gz.original_ids()
    return [int(orig[cand_dev[p]]) for p in picks]


def probe_bfs(root: int) -> dict:
    """One probed BFS: per-level features, both engines' times, fixed choice,
    and with gz.probe_wide(True) the times of three more edge-phase variants
    (uncompacted pull, compacted push, uncompacted push)."""
    gz.run(gz.options("bfs", engine="probe", root=root))
    recs = gz.records()
    return {
        "features": np.array([list(r.features) for r in recs], dtype=np.float32).reshape(-1, 8),
        "pull_ms": np.array([r.pull_ms for r in recs]),
        "push_ms": np.array([r.push_ms for r in recs]),
        "variant_ms": np.array([list(r.variant_ms) for r in recs], dtype=np.float64).reshape(-1, 3),
        "fixed": np.array([r.engine for r in recs], dtype=np.int32),
    }


def collect(specs, roots_per_graph: int, log=print, wide: bool = False) -> dict:
    runs = []
    t0 = time.time()
    gz.probe_wide(wide)
    try:
        for k, (scale, ef, seed) in enumerate(specs):
            gz.load_rmat_device(scale, (1 << scale) * ef, seed)
            for root in pick_roots(roots_per_graph, seed):
                run = probe_bfs(root)
                run["graph"] = (scale, ef, seed)
                run["root"] = root
                run["edges"] = gz.num_edges()
                runs.append(run)
            log(f"[collect] graph {k + 1}/{len(specs)} s{scale} ef{ef} seed {seed}: "
                f"{sum(r['features'].shape[0] for r in runs)} samples, {time.time() - t0:.1f} s")
    finally:
        gz.probe_wide(False)
    return pack_runs(runs)


def collect_workload(name: str, roots: int, seed: int, wide: bool = False) -> dict:
    """Probed searches from `roots` seeded roots on a named workload graph."""
    gz.load_workload_device(name)
    gz.probe_wide(wide)
    try:
        runs = []
        for root in pick_roots(roots, seed):
            run = probe_bfs(root)
            run["graph"] = (0, 0, seed)
            run["root"] = root
            run["edges"] = gz.num_edges()
            runs.append(run)
    finally:
        gz.probe_wide(False)
    return pack_runs(runs)


def pack_runs(runs: list[dict]) -> dict:
    lens = np.array([r["features"].shape[0] for r in runs], dtype=np.int64)
    return {
        "features": np.concatenate([r["features"] for r in runs]) if runs else np.zeros((0, 8), np.float32),
        "pull_ms": np.concatenate([r["pull_ms"] for r in runs]) if runs else np.zeros(0),
        "push_ms": np.concatenate([r["push_ms"] for r in runs]) if runs else np.zeros(0),
        "variant_ms": np.concatenate([r["variant_ms"] for r in runs]) if runs else np.zeros((0, 3)),
        "fixed": np.concatenate([r["fixed"] for r in runs]) if runs else np.zeros(0, np.int32),
        "run_lengths": lens,
        "graphs": np.array([r["graph"] for r in runs], dtype=np.int64).reshape(-1, 3),
        "roots": np.array([r["root"] for r in runs], dtype=np.int64),
        "edges": np.concatenate([np.full(r["features"].shape[0], r.get("edges", 1), dtype=np.float64)
                                 for r in runs]) if runs else np.zeros(0),
    }


def targets(data: dict, per_edge: bool = False) -> np.ndarray:
    """Regression targets: log edge-phase ms of (pull, push), less log E when
    per_edge (the choice only depends on their difference, and the per-edge
    cost varies far less with the graph's size than the time does)."""
    t = np.log(np.stack([data["pull_ms"], data["push_ms"]], axis=1))
    if per_edge:
        t = t - np.log(data["edges"])[:, None]
    return t.astype(np.float32)


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases, in the W1 b1 W2 b2 W3 b3 layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((8, 32), (32, 32), (32, 2)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, size=fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def normalize(X: np.ndarray, mean: np.ndarray, std: np.ndarray) -> np.ndarray:
    """(X - mean) / std per feature, and 0 for a feature whose std is 0 (a
    feature the model does not use), as the engine normalises."""
    used = std > 0
    out = np.zeros_like(X, dtype=np.float32)
    out[..., used] = ((X[..., used] - mean[used]) / std[used]).astype(np.float32)
    return out


def fit(data: dict, steps: int = 6000, lr: float = 1e-3, seed: int = 7, per_edge: bool = False,
        drop: tuple = ()):
    """Fit the MLP; features in `drop` get mean 0 and std 0 in the weights,
    so the model sees 0 for them and the engine need not compute them (for
    feature 3 the BFS edge phases then skip their in-degree reads)."""
    X = data["features"].astype(np.float32)
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    std[std < 1e-6] = 1.0
    for j in drop:
        mean[j] = 0.0
        std[j] = 0.0
    Xn = normalize(X, mean, std)
    params, hist = gz.selector_train(init_params(seed), Xn, targets(data, per_edge), steps, lr)
    return params, mean.astype(np.float32), std.astype(np.float32), hist


TRUST = (gz.TRUST_ABOVE, gz.TRUST_BELOW)  # the engine's default band (gz_selector_set_trust)


def policy_times(data: dict, params=None, mean=None, std=None, trust: tuple = TRUST) -> dict:
    """Total edge-phase ms of the fixed threshold, the learned selector and
    the per-level best, summed over the probed runs.  The learned policy sees
    its own previous choice and time (features 5 and 6): where it chose what
    the fixed threshold chose, the recorded feature (the device-clock time of
    that run) is exact; otherwise the other engine's time plus the same
    clock overhead stands in for it.  As in the engine, the model decides
    only while the level's vote (vertices plus out-edges of the frontier,
    features 0 and 1) lies in [E/5 / below, above * E/5], with (above, below)
    = `trust`; a factor below 1 lifts that side."""
    X = data["features"].astype(np.float32)
    tp, ts = data["pull_ms"], data["push_ms"]
    fixed = np.where(data["fixed"] == 1, tp, ts)
    out = {"fixed": float(fixed.sum()), "best": float(np.minimum(tp, ts).sum())}
    var = data.get("variant_ms")
    if var is not None and var.shape[0] == tp.shape[0] and np.all(var > 0):
        # the per-level best over five choices: the two engines as they run
        # and the uncompacted pull, compacted push and uncompacted push
        out["best_of_five"] = float(np.minimum(np.minimum(tp, ts), var.min(axis=1)).sum())
        out["levels_won_by_variant"] = [int(np.sum(var[:, j] < np.minimum(tp, ts))) for j in range(3)]
    if params is not None:
        start = 0
        total = 0.0
        for n in data["run_lengths"]:
            prev_engine, prev_ms = 0.0, 0.0
            for i in range(start, start + int(n)):
                x = X[i].copy()
                x[5] = prev_engine
                x[6] = np.log1p(prev_ms)
                y = gz.selector_forward(params, normalize(x[None, :], mean, std))[0]
                pull = bool(y[0] < y[1])
                above, below = trust
                vote = float(np.expm1(X[i][0]) + np.expm1(X[i][1]))
                threshold = float(data["edges"][i] // 5)
                if above >= 1.0 and vote > above * threshold:
                    pull = True
                elif below >= 1.0 and vote * below < threshold:
                    pull = False
                t = tp[i] if pull else ts[i]
                total += t
                chose_fixed = (1 if pull else 2) == int(data["fixed"][i])
                if i + 1 < start + int(n):
                    recorded = float(np.expm1(X[i + 1][6]))
                    prev_ms = recorded if chose_fixed else max(0.0, recorded - float(fixed[i]) + t)
                prev_engine = 1.0 if pull else 2.0
            start += int(n)
        out["learned"] = total
    return out


def save_weights(path: str, params, mean, std) -> None:
    with open(path, "wb") as f:
        f.write(np.ascontiguousarray(params, dtype="<f4").tobytes())
        f.write(np.ascontiguousarray(mean, dtype="<f4").tobytes())
        f.write(np.ascontiguousarray(std, dtype="<f4").tobytes())


def load_weights(path: str):
    raw = np.fromfile(path, dtype="<f4")
    p = raw[:gz.SEL_PARAMS]
    return p, raw[gz.SEL_PARAMS:gz.SEL_PARAMS + 8], raw[gz.SEL_PARAMS + 8:gz.SEL_PARAMS + 16]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--graphs", type=int, default=200)
    c.add_argument("--first-seed", type=int, default=1000)
    c.add_argument("--roots", type=int, default=2)
    c.add_argument("--out", required=True)
    c.add_argument("--wide", action="store_true", help="also time three more edge-phase variants per level")
    c.add_argument("--workload", default=None, help="probe a named workload graph (e.g. bfs-rmat24) instead")
    c.add_argument("--root-seed", type=int, default=24, help="root seed with --workload")
    f = sub.add_parser("fit")
    f.add_argument("--data", required=True)
    f.add_argument("--out", default=DEFAULT_WEIGHTS)
    f.add_argument("--steps", type=int, default=6000)
    f.add_argument("--lr", type=float, default=1e-3)
    f.add_argument("--per-edge", action="store_true", help="regress log ms - log E")
    f.add_argument("--drop-features", default="", help="comma-separated features the model does not use "
                   "(3: unvisited in-edges, whose in-degree reads the engine then skips)")
    e = sub.add_parser("evaluate")
    e.add_argument("--data", required=True)
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    gz.init(0)
    if a.cmd == "collect":
        if a.workload:
            data = collect_workload(a.workload, a.roots, a.root_seed, a.wide)
        else:
            data = collect(training_graphs(a.graphs, a.first_seed), a.roots, wide=a.wide)
        np.savez(a.out, **data)
        print(json.dumps({"samples": int(data["features"].shape[0]), "runs": int(data["run_lengths"].shape[0]),
                          **policy_times(data)}))
    elif a.cmd == "fit":
        data = dict(np.load(a.data))
        drop = tuple(int(x) for x in a.drop_features.split(",") if x.strip())
        params, mean, std, hist = fit(data, a.steps, a.lr, per_edge=a.per_edge, drop=drop)
        save_weights(a.out, params, mean, std)
        print(json.dumps({"samples": int(data["features"].shape[0]), "loss_first": float(hist[0]),
                          "loss_last": float(hist[-1]), **policy_times(data, params, mean, std)}))
    else:
        data = dict(np.load(a.data))
        params, mean, std = load_weights(a.weights)
        print(json.dumps(policy_times(data, params, mean, std)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
