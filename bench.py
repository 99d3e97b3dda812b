#!/usr/bin/env python3
"""Benchmark of the hybrid push/pull engine on one MI355X.

Default workload (the connected-components configuration): symmetrised R-MAT
scale 22, edge factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, plus 64 planted
1000-vertex random trees, uint64 labels, hybrid push/pull with the reference's
fixed E/5 threshold.  One step = one complete CC run from initial labels to
convergence (label initialisation + every edge and vertex phase, with the
per-iteration vote read-back that drives the engine switch).  Inputs are
resident in HBM before timing starts.

value = GTEPS, E * iterations / time (the reference's "Effective Processing
Rate", main.c:127-129), i.e. the mean per-iteration GTEPS of the run.

The same graph also gets 20 PageRank iterations (scheduler-aware pull, FP64)
reported under "pagerank_pull", with its own roofline figure.

roofline: algorithmic bytes of the dominant kernel's launch divided by its
average launch time, measured live with HIP events on the library's stream.
cpu_baseline: the CPU oracle (a pthreads restatement of the reference
engine) on the same graph on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "grazelle-ppopp18_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS/iter + % HBM peak, 1 MI355X; learned-vs-fixed switch & batched-vs-seq"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 summary, if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", open(path).read())
    return int(m.group(1)) if m else None


def cpu_baseline_cc(g, iters_gpu, threads, runs):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    G = orc.Graph(g.num_vertices, g.pull, g.push)
    times, iters = [], []
    for _ in range(runs):
        _, _, r = orc.cc(G, orc.options(threads=threads), log_cap=0)
        times.append(r.elapsed_ms)
        iters.append(r.iterations)
    ms = statistics.median(times)
    it = iters[times.index(min(times, key=lambda x: abs(x - ms)))]
    return {"value": round(g.num_edges * it / (ms * 1e-3) / 1e9, 3), "unit": "GTEPS", "cores": threads,
            "kind": "port",
            "sample": f"full workload (V={g.num_vertices}, E={g.num_edges}), CC hybrid E/5 threshold, "
                      f"{it} iterations, median of {runs} runs of the pthreads oracle, {ms:.1f} ms per run"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cc-rmat22-sym-trees")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=512)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the device copy kernel on MIB mebibytes (rocprof counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import ctypes
        import grazelle as gz

        gz.init(0)
        nbytes = args.calibrate_copy << 20
        gz.lib.gz_device_malloc.restype = ctypes.c_void_p
        gz.lib.gz_device_malloc.argtypes = [ctypes.c_size_t]
        gz.lib.gz_device_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        a = gz.lib.gz_device_malloc(nbytes)
        b = gz.lib.gz_device_malloc(nbytes)
        for _ in range(max(1, args.steps)):
            gz.lib.gz_device_copy(b, a, nbytes)
        gz.lib.gz_synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": max(1, args.steps)}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "The per-GPU shard design is described in DESIGN.md.", file=sys.stderr)
        return 2

    import grazelle as gz

    if gz.device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    dev = gz.init(0)
    t0 = time.time()
    g = gz.workload_graph(args.workload, seed=1)
    log(f"graph {args.workload}: V={g.num_vertices} E={g.num_edges} built in {time.time() - t0:.1f}s")
    gz.set_chunk_vectors(args.chunk)
    gz.load(g)
    V, E = g.num_vertices, g.num_edges
    nvp = gz.pull_vector_count()
    log(f"device {dev}: pull vectors {nvp}, push vectors {gz.push_vector_count()}")

    opts = gz.options("cc", engine="hybrid")
    for _ in range(args.warmup):
        gz.run(opts)
    gz.lib.gz_synchronize()
    gz.kernel_timing(True)
    results = []
    gz.lib.gz_synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        results.append(gz.run(opts))
    gz.lib.gz_synchronize()
    t_total = time.perf_counter() - t_start
    kt = gz.kernel_times()
    gz.kernel_timing(False)
    ms_per_step = t_total * 1e3 / args.steps
    iters = results[-1].iterations
    value = E * iters / (ms_per_step * 1e-3) / 1e9
    log(f"CC: {ms_per_step:.3f} ms/step, {iters} iterations ({results[-1].pull_iterations} pull / "
        f"{results[-1].push_iterations} push), {value:.1f} GTEPS")
    for k, (ms, n) in kt.items():
        if n:
            log(f"  kernel {k:15s} launches {n:5d} avg {ms / n * 1e3:9.1f} us  total {ms:8.2f} ms")

    # dominant kernel of the CC step
    dom = max(("pull_cc", "push_cc"), key=lambda k: kt[k][0])
    dom_ms, dom_n = kt[dom]
    if dom == "pull_cc":
        bytes_launch = 32 * nvp + 8 * E + 8 * V
    else:
        bytes_launch = None
    roofline = None
    if bytes_launch and dom_n:
        avg_s = dom_ms / dom_n * 1e-3
        achieved = bytes_launch / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": "gz_k_edge_pull_cc", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": read_traffic("gz_k_edge_pull_cc", args.workload),
                    "bytes_per_launch": bytes_launch, "avg_launch_us": round(avg_s * 1e6, 2),
                    "launches": dom_n}

    # PageRank pull on the same graph
    gz.run(gz.options("pr", iterations=5))
    gz.kernel_timing(True)
    gz.lib.gz_synchronize()
    tp = time.perf_counter()
    rp = gz.run(gz.options("pr", iterations=20))
    tp = time.perf_counter() - tp
    ktp = gz.kernel_times()
    gz.kernel_timing(False)
    pms, pn = ktp["pull_pr"]
    pr_bytes = 32 * nvp + 8 * E + 8 * V
    pr_avg = pms / pn * 1e-3
    pr = {"iterations": 20, "ms_per_iteration": round(tp * 1e3 / 20, 4),
          "gteps_per_iteration": round(E * 20 / tp / 1e9, 2),
          "edge_kernel": "gz_k_edge_pull_pr", "avg_launch_us": round(pr_avg * 1e6, 2),
          "bytes_per_launch": pr_bytes, "achieved_GBs": round(pr_bytes / pr_avg / 1e9, 1),
          "frac_hbm_peak": round(pr_bytes / pr_avg / 1e9 / HBM_PEAK_GBS, 4),
          "kernel_gteps": round(E / pr_avg / 1e9, 2),
          "traffic": read_traffic("gz_k_edge_pull_pr", args.workload)}
    log(f"PR pull: {pr}")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        t0 = time.time()
        cpu = cpu_baseline_cc(g, iters, threads, args.cpu_runs)
        log(f"cpu baseline ({time.time() - t0:.1f}s): {cpu}")

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": args.workload, "algorithm": "cc", "engine": "hybrid (fixed E/5 threshold)",
                   "num_vertices": V, "num_edges": E, "pull_vectors": nvp, "iterations": iters,
                   "pull_iterations": results[-1].pull_iterations,
                   "push_iterations": results[-1].push_iterations, "chunk_vectors": args.chunk,
                   "device": dev},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pagerank_pull": pr,
    }
    print(json.dumps(line), flush=True)
    gz.release()
    return 0


if __name__ == "__main__":
    sys.exit(main())
