#!/usr/bin/env python3
"""Benchmark of the hybrid push/pull engine on one MI355X.

Headline (value): edge-pull PageRank on the full-scale power-law graph --
R-MAT scale 26 (a/b/c/d = 0.57/0.19/0.19/0.05) folded by modulo into 41.7M
vertices, 1.494B raw edges, 1.47B after self-loop and duplicate removal,
built in HBM by device ingress with vertices renumbered by out-degree (results
are reported in original ids).  FP64 ranks, damping 0.85.  One step = one
PageRank iteration (pull sweep, merge, sink-mass reduction, vertex phase);
value = E / (time per iteration), GTEPS per iteration (the reference's rate,
main.c:127-129, per iteration).  Inputs are resident in HBM before timing.
The pull runs over the source-sliced layout (two tiers of eight 2^20-vertex
source windows, one window per XCD and tier, the hottest 20480 sources of each
tier-0 window in LDS, chunks of 2048 vectors); --plain-pull runs the
scheduler-aware pull over the ordinary list.

roofline: the pull sweep kernel's algorithmic bytes per launch, 32*Nvp
(vectors of the ordinary pull list) + 8*E (one FP64 gather per edge) + 8*V
(one accumulator store per vertex), over its average launch time measured
live with HIP events on the library's stream.  The sliced layout streams more
vectors than Nvp; those extra bytes are not counted as useful work.  traffic:
HBM bytes per launch from the committed rocprofv3 counter summary
(profiles/).

Sub-objects, each one configuration of the workload set:
  cc_rmat22   CC, symmetrised R-MAT s22 + 64 planted 1000-vertex trees,
              hybrid with the fixed E/5 threshold, time to solution with the
              source-sliced pull on the renumbered graph (labels checked
              against the plain pull's, which is reported beside it); the
              rate is E / time to solution (the GPU needs more sweeps than
              the sequential reference, so E x iterations / t would flatter it)
  bfs_rmat24  BFS from 64 seeded roots on R-MAT s24: fixed threshold versus
              the learned selector (selector_bfs.bin), depths checked equal;
              the same 64 roots as one batch (gz_bfs_batch_run, one pass per
              level for all searches); roofline from the algorithmic bytes per
              level the kernels count
  batch_1024  1024 R-MAT graphs (scale 10..17 with P(s) ~ 1/s, edge factor
              4..32): batched PageRank (10 iterations, captured HIP graph)
              and CC versus the same graphs run one after another
  pr_rmat16   PageRank 20 iterations on R-MAT s16 (the reference's own size)
cpu_baseline: the CPU oracle (pthreads restatement of the reference engine)
on the headline graph itself, its lists copied back from the device, on the
job's CPU share (OMP_NUM_THREADS, else every core), median of 5 runs; the
same oracle is timed beside cc_rmat22, bfs_rmat24 and pr_rmat16.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "grazelle-ppopp18_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS/iter + % HBM peak, 1 MI355X; learned-vs-fixed switch & batched-vs-seq"
FULL = "full-rmat26-fold41.7m"
DAMPING = 0.85
SLICES_LOG2 = 20      # source windows of 2^20 vertices, one per XCD and tier
SLICE_TIERS = 2       # two tiers of eight windows: the top 16.8 M sources
SLICE_BLOCKS = 5      # LDS blocks of 20480 sources carved from each tier-0 window
HOT_CACHE = 20480     # hottest sources of each tier-0 window (and each block in turn) in LDS (160 KiB)
SLICED_CHUNK = 2048   # vectors per chunk of the sliced pull


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 summary, if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", open(path).read())
    return int(m.group(1)) if m else None


def pull_bytes(gz) -> int:
    return 32 * gz.pull_vector_count() + 8 * gz.num_edges() + 8 * gz.num_vertices()


def headline(gz, args) -> dict:
    t0 = time.time()
    sliced = not args.plain_pull
    gz.load_workload_device(FULL, relabel=not args.no_relabel, slices_log2=SLICES_LOG2 if sliced else 0,
                            slice_tiers=SLICE_TIERS, slice_blocks=SLICE_BLOCKS)
    gz.lib.gz_synchronize()
    V, E, nvp = gz.num_vertices(), gz.num_edges(), gz.pull_vector_count()
    log(f"[headline] {FULL}: V={V} E={E} Nvp={nvp} ingress {time.time() - t0:.2f}s")
    variant = "source_sliced" if sliced else "sched_aware"
    if sliced:
        gz.set_chunk_vectors(SLICED_CHUNK)
        gz.slices_set_hot_cache(HOT_CACHE)
    opts = gz.options("pr", iterations=args.steps, damping=DAMPING, pull_variant=variant)
    if args.warmup:
        gz.run(gz.options("pr", iterations=args.warmup, damping=DAMPING, pull_variant=variant))
    gz.lib.gz_synchronize()
    w0 = time.perf_counter()
    r = gz.run(opts)
    gz.lib.gz_synchronize()
    wall_ms = (time.perf_counter() - w0) * 1e3
    ranks = gz.pr_ranks()
    # per-kernel times from a second pass with HIP events around every launch
    gz.kernel_timing(True)
    gz.run(opts)
    kt = gz.kernel_times()
    gz.kernel_timing(False)
    sliced_vectors = int(gz.source_slices_info()[0].sum()) if sliced else nvp
    gz.set_chunk_vectors(512)
    gz.slices_set_hot_cache(0)
    indeg = gz.indegrees()
    max_in, max_out = int(indeg.max()), int(gz.outdegrees().max())
    del indeg
    full_other = None if args.only_headline else full_cc_bfs(gz)
    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = cpu_baseline_full(gz, cpu_threads(args), args.cpu_runs)
        log(f"[cpu_baseline] {time.time() - t0:.1f}s {cpu}")
    ms_step = r.elapsed_ms / args.steps
    pms, pn = kt["pull_pr"]
    avg = pms / pn
    B = pull_bytes(gz)
    achieved = B / (avg * 1e-3) / 1e9
    log(f"[headline] {args.steps} iterations {r.elapsed_ms:.2f} ms ({ms_step:.3f} ms/it), pull {avg:.3f} ms "
        f"{achieved:.0f} GB/s, rank sum {ranks.sum():.12f}")
    kernel = "gz_k_edge_pull_pr_sliced_hot" if sliced else "gz_k_edge_pull_pr"
    return {
        "V": V, "E": E, "nvp": nvp, "ms_step": ms_step, "wall_ms": wall_ms, "sliced": sliced,
        "value": E / (ms_step * 1e-3) / 1e9,
        "phases_ms_per_iteration": {k: round(ms / n, 4) for k, (ms, n) in kt.items() if n},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(kernel, FULL), "bytes_per_launch": B,
                     "streamed_vectors": sliced_vectors,
                     "avg_launch_us": round(avg * 1e3, 2), "launches": pn,
                     "kernel_gteps": round(E / (avg * 1e-3) / 1e9, 2)},
        "rank_sum": float(ranks.sum()),
        "max_in_degree": max_in, "max_out_degree": max_out,
        "full_cc_bfs": full_other,
        "cpu_baseline": cpu,
    }


def full_cc_bfs(gz) -> dict:
    """CC (plain and source-sliced pull) and BFS (16 roots, fixed threshold and
    learned selector) on the already loaded full-scale graph."""
    import selector_train as st

    E = gz.num_edges()
    out = {}
    try:
        gz.run(gz.options("cc"))
        r = gz.run(gz.options("cc"))
        labels = gz.cc_labels()
        out["cc"] = {"time_to_solution_ms": round(r.elapsed_ms, 2), "iterations": r.iterations,
                     "pull_iterations": r.pull_iterations, "push_iterations": r.push_iterations,
                     "edges_per_s_to_solution_G": round(E / (r.elapsed_ms * 1e-3) / 1e9, 2)}
        if gz.source_slices_info() is not None:
            # the pull sweeps over the source-sliced layout (labels must not change)
            gz.set_chunk_vectors(SLICED_CHUNK)
            gz.slices_set_hot_cache(HOT_CACHE)
            gz.run(gz.options("cc", pull_variant="source_sliced"))
            r = gz.run(gz.options("cc", pull_variant="source_sliced"))
            same = bool(np.array_equal(gz.cc_labels(), labels))
            gz.set_chunk_vectors(512)
            gz.slices_set_hot_cache(0)
            out["cc_source_sliced"] = {"time_to_solution_ms": round(r.elapsed_ms, 2), "iterations": r.iterations,
                                       "pull_iterations": r.pull_iterations,
                                       "push_iterations": r.push_iterations, "labels_equal_plain": same}
        roots = st.pick_roots(16, 26)
        gz.run(gz.options("bfs", root=roots[0]))
        fixed = sum(gz.run(gz.options("bfs", root=x)).elapsed_ms for x in roots)
        gz.selector_load(st.DEFAULT_WEIGHTS)
        gz.run(gz.options("bfs", engine="learned", root=roots[0]))
        learned = sum(gz.run(gz.options("bfs", engine="learned", root=x)).elapsed_ms for x in roots)
        gz.selector_clear()
        out["bfs_16_roots"] = {"fixed_threshold_ms": round(fixed, 3), "learned_ms": round(learned, 3),
                               "learned_speedup": round(fixed / learned, 4),
                               "fixed_gteps_per_search": round(E / (fixed / len(roots) * 1e-3) / 1e9, 2)}
    except Exception as exc:  # reported, not hidden
        out["error"] = f"{type(exc).__name__}: {exc}"
    log(f"[full_cc_bfs] {json.dumps(out)}")
    return out


CC_SLICES_LOG2 = 19   # the CC configuration's sliced pull: one tier of 2^19-vertex windows
CC_SLICE_TIERS = 1


def _cc_timed(gz, variant: str):
    """Median time to solution of 5 runs, then one more run with per-kernel
    events (which add their own overhead) for the pull kernel's average."""
    gz.run(gz.options("cc", pull_variant=variant))
    times, res = [], None
    for _ in range(5):
        res = gz.run(gz.options("cc", pull_variant=variant))
        times.append(res.elapsed_ms)
    gz.kernel_timing(True)
    gz.run(gz.options("cc", pull_variant=variant))
    kt = gz.kernel_times()
    gz.kernel_timing(False)
    return statistics.median(times), res, kt["pull_cc"]


def cc_rmat22(gz, cpu: int) -> dict:
    # plain pull on the loaded ids, then the source-sliced pull on the graph
    # renumbered by out-degree; labels (original ids) must be identical
    gz.load_workload_device("cc-rmat22-sym-trees")
    V, E = gz.num_vertices(), gz.num_edges()
    ms_plain, res_plain, (pms, pn) = _cc_timed(gz, "sched_aware")
    labels = gz.cc_labels()
    cpu_fig = None
    if cpu:
        orc = _oracle()
        G = orc.Graph.from_vectors(V, gz.device_vectors(True), gz.device_vectors(False))
        times = []
        for _ in range(5):
            lab, _, r = orc.cc(G, orc.options(threads=cpu))
            times.append(r.elapsed_ms)
        if not np.array_equal(lab, labels):
            raise AssertionError("cc_rmat22: oracle labels differ from the GPU's")
        del G
        cpu_fig = {"time_to_solution_ms": round(statistics.median(times), 2), "threads": cpu,
                   "iterations": r.iterations, "runs": 5}
    B = pull_bytes(gz)
    plain = {"time_to_solution_ms": round(ms_plain, 3), "iterations": res_plain.iterations,
             "pull_iterations": res_plain.pull_iterations, "push_iterations": res_plain.push_iterations}
    if pn:
        avg = pms / pn
        plain["pull_cc_kernel"] = {"avg_launch_us": round(avg * 1e3, 2), "bytes_per_launch": B,
                                   "achieved_GBs": round(B / (avg * 1e-3) / 1e9, 1),
                                   "frac_hbm_peak": round(B / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                   "traffic": read_traffic("gz_k_edge_pull_cc", "cc-rmat22-sym-trees")}
    gz.load_workload_device("cc-rmat22-sym-trees", relabel=True, slices_log2=CC_SLICES_LOG2,
                            slice_tiers=CC_SLICE_TIERS)
    gz.set_chunk_vectors(SLICED_CHUNK)
    gz.slices_set_hot_cache(HOT_CACHE)
    try:
        ms, res, (sms, sn) = _cc_timed(gz, "source_sliced")
        same = bool(np.array_equal(gz.cc_labels(), labels))
    finally:
        gz.set_chunk_vectors(512)
        gz.slices_set_hot_cache(0)
    if not same:
        raise AssertionError("cc_rmat22: source-sliced labels differ from the plain pull's")
    out = {"workload": "cc-rmat22-sym-trees", "V": V, "E": E,
           "pull": f"source-sliced (renumbered, S=2^{CC_SLICES_LOG2}, LDS cache {HOT_CACHE}, chunk {SLICED_CHUNK})",
           "time_to_solution_ms": round(ms, 3), "iterations": res.iterations,
           "pull_iterations": res.pull_iterations, "push_iterations": res.push_iterations,
           "edges_per_s_to_solution_G": round(E / (ms * 1e-3) / 1e9, 2), "labels_equal_plain": same,
           "plain_pull": plain}
    if sn:
        avg = sms / sn
        out["pull_cc_kernel"] = {"avg_launch_us": round(avg * 1e3, 2), "bytes_per_launch": B,
                                 "achieved_GBs": round(B / (avg * 1e-3) / 1e9, 1),
                                 "frac_hbm_peak": round(B / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if cpu_fig:
        out["cpu_oracle"] = cpu_fig
    return out


def bfs_bytes(gz, root: int) -> tuple[int, int]:
    """(algorithmic bytes, levels) of one fixed-threshold search: per level
    32 B per vector loaded, 8 B per valid lane examined (both counted by the
    kernels), two passes over the V/8-byte frontier bitmap, 8 B per vertex
    written."""
    V = gz.num_vertices()
    gz.run(gz.options("bfs", root=root, collect_stats=True))
    recs = gz.records()
    depths = gz.bfs_result()[1]
    counts = np.bincount(depths[depths != np.iinfo(np.uint64).max].astype(np.int64), minlength=len(recs) + 1)
    total = 0
    for i, r in enumerate(recs):
        total += 32 * r.vectors + 8 * r.edges + 2 * ((V + 7) // 8) + 8 * int(counts[i + 1])
    return total, len(recs)


BFS_VOTE_BANDS = ((4096, "<=4K"), (65536, "<=64K"), (1 << 20, "<=1M"), (8 << 20, "<=8M"), (None, ">8M"))


def bfs_level_split(gz, roots, engine: str) -> dict:
    """Where the searches' edge-phase time goes: per engine and band of the
    frontier statistic (vote) that chose it, the levels and their time, from
    the per-iteration profile (HIP events around each edge phase) of an
    untimed pass over the same roots."""
    acc = {}
    for root in roots:
        gz.run(gz.options("bfs", engine=engine, root=root, profile=True))
        for r in gz.records():
            band = next(name for lim, name in BFS_VOTE_BANDS if lim is None or r.frontier_stat <= lim)
            key = ("pull" if r.engine == 1 else "push") + " vote " + band
            e = acc.setdefault(key, [0, 0.0])
            e[0] += 1
            e[1] += r.edge_ms
    return {k: {"levels": n, "ms_per_search": round(ms / len(roots), 4), "us_per_level": round(1e3 * ms / n, 1)}
            for k, (n, ms) in sorted(acc.items())}


def bfs_rmat24(gz, cpu: int) -> dict:
    import selector_train as st

    gz.load_workload_device("bfs-rmat24")
    V, E = gz.num_vertices(), gz.num_edges()
    roots = st.pick_roots(64, 24)
    gz.run(gz.options("bfs", root=roots[0]))
    out = {"workload": "bfs-rmat24", "V": V, "E": E, "roots": len(roots)}
    totals = {}
    for engine in ("hybrid", "learned"):
        if engine == "learned":
            gz.selector_load(st.DEFAULT_WEIGHTS)
            gz.run(gz.options("bfs", engine=engine, root=roots[0]))  # builds the in-degree array once
        t = 0.0
        for root in roots:
            t += gz.run(gz.options("bfs", engine=engine, root=root)).elapsed_ms
        totals[engine] = t
    # depths agree between the engines (spot check on four roots)
    for root in roots[:4]:
        gz.run(gz.options("bfs", engine="learned", root=root))
        d_l = gz.bfs_result()[1]
        gz.selector_clear()
        gz.run(gz.options("bfs", engine="hybrid", root=root))
        d_f = gz.bfs_result()[1]
        gz.selector_load(st.DEFAULT_WEIGHTS)
        if not np.array_equal(d_l, d_f):
            raise AssertionError(f"BFS depths differ between engines at root {root}")
    split_learned = bfs_level_split(gz, roots, "learned")
    gz.selector_clear()
    out["edge_time_by_level"] = {"fixed_threshold": bfs_level_split(gz, roots, "hybrid"), "learned": split_learned}
    out.update({"fixed_threshold_ms": round(totals["hybrid"], 3), "learned_ms": round(totals["learned"], 3),
                "learned_speedup": round(totals["hybrid"] / totals["learned"], 4),
                "fixed_ms_per_search": round(totals["hybrid"] / len(roots), 4),
                "fixed_gteps_per_search": round(E / (totals["hybrid"] / len(roots) * 1e-3) / 1e9, 2)})
    # the same 64 roots in one batch: one pass per level for all searches
    gz.bfs_batch(roots)
    batch = [gz.bfs_batch(roots) for _ in range(3)]
    bms = min(r.elapsed_ms for r in batch)
    for i, root in enumerate(roots[:4]):
        gz.run(gz.options("bfs", root=root))
        if not np.array_equal(gz.bfs_batch_result(i, parents=False)[1], gz.bfs_result()[1]):
            raise AssertionError(f"batched BFS depths differ from the single search at root {root}")
    out["batched_64_roots"] = {"ms": round(bms, 3), "levels": int(batch[0].iterations),
                               "pull_levels": int(batch[0].pull_iterations),
                               "speedup_vs_fixed_sequential": round(totals["hybrid"] / bms, 2),
                               "gteps_per_search": round(E * len(roots) / (bms * 1e-3) / 1e9, 2)}
    # algorithmic bytes of the fixed-threshold searches (a counting pass
    # outside the timed runs) over their time
    nbytes = levels = 0
    for root in roots:
        b, n = bfs_bytes(gz, root)
        nbytes += b
        levels += n
    achieved = nbytes / (totals["hybrid"] * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_search": nbytes // len(roots),
                       "levels_per_search": round(levels / len(roots), 2)}
    if cpu:
        orc = _oracle()
        G = orc.Graph.from_vectors(V, gz.device_vectors(True), gz.device_vectors(False))
        t = []
        for root in roots[:8]:
            t.append(orc.bfs(G, orc.options(threads=cpu, root=root))[3].elapsed_ms)
        del G
        out["cpu_oracle"] = {"ms_per_search": round(statistics.median(t), 2), "threads": cpu,
                             "roots": 8, "statistic": "median"}
    return out


def batch_1024(gz, cpu: int) -> dict:
    import batch_tune

    specs = gz.batch_specs()
    gz.load_batch_device(specs)
    E = gz.num_edges()
    tune = batch_tune.tune(reps=2)
    gz.batch_pr(10, damping=DAMPING, use_graph=True)
    pr_b = min(gz.batch_pr(10, damping=DAMPING, use_graph=True) for _ in range(3))
    gz.run(gz.options("cc"))
    cc_runs = [gz.run(gz.options("cc")) for _ in range(2)]
    cc_b = min(r.elapsed_ms for r in cc_runs)
    host_labels = gz.cc_labels()
    # the same batch with the engine chosen on the device and the iterations
    # replayed from a captured graph (eager until its allocations settle,
    # captured the second time in a row, then replayed)
    gz.cc_graph_replay(True)
    gz.run(gz.options("cc"))
    gz.run(gz.options("cc"))
    c0, l0 = gz.graph_replay_counts()
    cap_runs = [gz.run(gz.options("cc")) for _ in range(3)]
    c1, l1 = gz.graph_replay_counts()
    cap_equal = bool(np.array_equal(gz.cc_labels(), host_labels))
    del host_labels
    gz.cc_graph_replay(False)
    cc_cap = min(r.elapsed_ms for r in cap_runs)
    pr_s = cc_s = 0.0
    for s, ef, seed in specs:
        gz.load_rmat_device(s, (1 << s) * ef, seed)
        pr_s += gz.run(gz.options("pr", iterations=10, damping=DAMPING)).elapsed_ms
        cc_s += gz.run(gz.options("cc")).elapsed_ms
    return {"graphs": len(specs), "E_total": E, "tuned_config": tune["config"],
            "pagerank_10it": {"batched_ms": round(pr_b, 3), "sequential_ms": round(pr_s, 3),
                              "speedup": round(pr_s / pr_b, 2),
                              "batched_gteps_per_iteration": round(E * 10 / (pr_b * 1e-3) / 1e9, 2)},
            "cc": {"batched_ms": round(cc_b, 3), "sequential_ms": round(cc_s, 3), "speedup": round(cc_s / cc_b, 2),
                   "iterations": int(cc_runs[-1].iterations),
                   "captured_device_loop": {"batched_ms": round(cc_cap, 3), "speedup": round(cc_s / cc_cap, 2),
                                            "iterations": int(cap_runs[-1].iterations),
                                            "graph_captures": int(c1 - c0), "graph_launches": int(l1 - l0),
                                            "labels_equal_host_loop": cap_equal}}}


def pr_rmat16(gz, cpu: int) -> dict:
    # 257,559 pull vectors: chunks of 256 give about 1000 wavefronts (512
    # leaves most CUs with two); the merge, the sum of S and the vertex phase
    # run as one launch after each pull (the fused tail)
    chunk = 256
    gz.load_workload_device("pr-rmat16")
    gz.set_chunk_vectors(chunk)
    try:
        opts = gz.options("pr", iterations=20, damping=DAMPING)
        gz.run(opts)
        ms = min(gz.run(opts).elapsed_ms for _ in range(5))
        fused = bool(gz.lib.gz_pr_fused_tail_applies())
        gz.pr_fused_tail(False)
        gz.run(opts)
        ms_four = min(gz.run(opts).elapsed_ms for _ in range(5))
        gz.pr_fused_tail(True)
    finally:
        gz.set_chunk_vectors(512)
    E = gz.num_edges()
    out = {"workload": "pr-rmat16", "V": gz.num_vertices(), "E": E, "iterations": 20, "chunk_vectors": chunk,
           "fused_tail": fused, "ms": round(ms, 4), "gteps_per_iteration": round(E * 20 / (ms * 1e-3) / 1e9, 2),
           "four_launches_per_iteration_ms": round(ms_four, 4)}
    if cpu:
        orc = _oracle()
        g = gz.workload_graph("pr-rmat16")
        G = orc.Graph(g.num_vertices, g.pull, g.push)
        for threads in sorted({1, cpu}):
            t = [orc.pagerank(G, orc.options(threads=threads, iterations=20, damping=DAMPING))[2].elapsed_ms
                 for _ in range(5)]
            m = statistics.median(t)
            out[f"cpu_oracle_{threads}t"] = {"ms": round(m, 3), "gteps_per_iteration": round(E * 20 / (m * 1e-3) / 1e9, 3),
                                             "runs": 5, "statistic": "median"}
    return out


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    return orc


def cpu_threads(args) -> int:
    """Worker threads for the CPU oracle: the job's CPU share (the
    OMP_NUM_THREADS the host sets for a one-GPU job), else every core."""
    if args.cpu_threads:
        return args.cpu_threads
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_full(gz, threads: int, runs: int, iterations: int = 3) -> dict:
    """The oracle's PageRank on the headline graph itself: its pull and push
    lists copied back from the device, the pthreads restatement of the
    reference engine (scheduler-aware pull, 32 units per thread), timed over
    thread start to join as main.c:114-120 times threads_spawn."""
    orc = _oracle()
    V, E = gz.num_vertices(), gz.num_edges()
    G = orc.Graph.from_vectors(V, gz.device_vectors(True), gz.device_vectors(False))
    times = []
    for _ in range(runs):
        _, _, r = orc.pagerank(G, orc.options(threads=threads, iterations=iterations, damping=DAMPING))
        times.append(r.elapsed_ms / iterations)
    del G
    ms = statistics.median(times)
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"value": round(E / (ms * 1e-3) / 1e9, 3), "unit": "GTEPS/iter", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity_cpus": affinity, "cpu_model": cpu_model(),
            "ms_per_iteration": round(ms, 2),
            "sample": f"the headline graph itself ({V} vertices, {E} edges, the device's renumbered lists "
                      f"copied back): PageRank pull, damping {DAMPING}, {iterations} iterations per run, median "
                      f"of {runs} runs of the pthreads oracle on {threads} threads"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--only-headline", action="store_true", help="skip the other configurations (profiling)")
    ap.add_argument("--no-relabel", action="store_true")
    ap.add_argument("--plain-pull", action="store_true", help="headline over the ordinary pull list")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", choices=("cc_rmat22", "bfs_rmat24", "batch_1024", "pr_rmat16"), default=None,
                    help="only run this configuration and print its result (profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the device copy kernel on MIB mebibytes (rocprof counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}); "
              "the value is one GPU's throughput.", file=sys.stderr)
        return 2
    import grazelle as gz

    if args.calibrate_copy:
        import ctypes

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
    if gz.device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    dev = gz.init(0)
    configs = (("cc_rmat22", cc_rmat22), ("bfs_rmat24", bfs_rmat24), ("batch_1024", batch_1024),
               ("pr_rmat16", pr_rmat16))
    cpu = 0 if args.no_cpu else cpu_threads(args)
    if args.config:
        out = dict(configs)[args.config](gz, cpu)
        print(json.dumps({args.config: out}), flush=True)
        gz.release()
        return 0
    h = headline(gz, args)
    extra = {}
    if not args.only_headline:
        for name, fn in configs:
            t0 = time.time()
            try:
                extra[name] = fn(gz, cpu)
            except Exception as exc:  # a failing configuration is reported, not hidden
                extra[name] = {"error": f"{type(exc).__name__}: {exc}"}
            log(f"[{name}] {time.time() - t0:.1f}s {json.dumps(extra[name])}")
    line = {
        "metric": METRIC, "value": round(h["value"], 3), "unit": "GTEPS/iter", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(h["ms_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": FULL + " PageRank pull (FP64, damping 0.85)", "num_vertices": h["V"],
                   "num_edges": h["E"], "pull_vectors": h["nvp"],
                   "ingress": "device, renumbered by out-degree" if not args.no_relabel else "device",
                   "pull": (f"source-sliced (S=2^{SLICES_LOG2}, {SLICE_TIERS} tiers, {SLICE_BLOCKS} LDS blocks, "
                            f"LDS cache {HOT_CACHE}, chunk {SLICED_CHUNK})"
                            if h["sliced"] else "scheduler-aware, ordinary pull list"),
                   "max_in_degree": h["max_in_degree"], "max_out_degree": h["max_out_degree"],
                   "device": dev},
        "roofline": h["roofline"],
        "cpu_baseline": h["cpu_baseline"],
        "phases_ms_per_iteration": h["phases_ms_per_iteration"],
        "full_scale_cc_bfs": h["full_cc_bfs"],
        **extra,
    }
    print(json.dumps(line), flush=True)
    gz.release()
    return 0


if __name__ == "__main__":
    sys.exit(main())
