"""Python binding of libgrazelle.so (ctypes over the C ABI in include/).

This is a thin host-side mirror for tests, benchmarks and scripting: graph
generation and loading, the algorithm drivers and result download.  All
computation happens in the library's HIP kernels; there is no Python or CPU
fallback.  Importing this module raises if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgrazelle.so")

ALG = {"pr": 0, "cc": 1, "bfs": 2}
ENGINE = {"hybrid": 0, "pull": 1, "push": 2, "learned": 3, "probe": 4}
STAT = {"both": 0, "outdeg": 1, "count": 2}
PULL_VARIANT = {"sched_aware": 0, "traditional": 1, "source_sliced": 2, "merge_buffer": 3}

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libgrazelle.so is not built ({LIB_PATH}); run __graft_entry__.build()")

lib = C.CDLL(LIB_PATH)

u64 = C.c_uint64
u32 = C.c_uint32
vp = C.c_void_p


class RunOptions(C.Structure):
    _fields_ = [
        ("algorithm", u32), ("engine", u32), ("iterations", u32), ("stat_mode", u32),
        ("pull_variant", u32), ("edge_only", u32), ("vertex_only", u32), ("profile", u32),
        ("collect_stats", u32), ("max_iterations", u32), ("damping", C.c_double), ("root", u64),
    ]


class RunResult(C.Structure):
    _fields_ = [("iterations", u64), ("pull_iterations", u64), ("push_iterations", u64),
                ("elapsed_ms", C.c_double)]


class IterationRecord(C.Structure):
    _fields_ = [("iteration", u64), ("engine", u32), ("pad", u32), ("frontier_stat", u64),
                ("vectors", u64), ("edges", u64), ("edge_ms", C.c_double), ("features", C.c_float * 8),
                ("predicted", C.c_float * 2), ("pull_ms", C.c_double), ("push_ms", C.c_double),
                ("variant_ms", C.c_double * 3)]


class RmatParams(C.Structure):
    _fields_ = [("scale", u32), ("num_vertices", u64), ("num_edges", u64), ("a", C.c_double),
                ("b", C.c_double), ("c", C.c_double), ("seed", u64), ("edge_offset", u64)]


class PhaseConfig(C.Structure):
    _fields_ = [("damping", C.c_double), ("stat_mode", u32), ("pull_variant", u32),
                ("chunk_vectors", u32), ("bfs_level", u32), ("collect_stats", u32),
                ("push_frontier_skip", u32), ("vectors_per_lane", u32)]


class PhaseGraph(C.Structure):
    _fields_ = [("num_vertices", u64), ("num_edges", u64), ("pull_vectors", vp), ("pull_vector_count", u64),
                ("pull_vertex_offsets", vp), ("push_vectors", vp), ("push_vector_count", u64),
                ("push_vertex_offsets", vp), ("vertex_outdegrees", vp), ("vertex_props", vp),
                ("vertex_accumulators", vp), ("frontier_has_info", vp), ("frontier_wants_info", vp)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


P = C.POINTER
_sig("gz_runtime_device_count", C.c_int)
_sig("gz_runtime_init", C.c_int, C.c_int)
_sig("gz_runtime_shutdown", None)
_sig("gz_runtime_device_name", C.c_char_p)
_sig("gz_synchronize", None)
_sig("gz_edge_list_write", C.c_int, C.c_char_p, u64, u64, vp)
_sig("gz_edge_list_read", C.c_int, C.c_char_p, P(u64), P(u64), P(P(u64)))
_sig("gz_edges_sort", u64, vp, u64, C.c_int, C.c_int)
_sig("gz_pack_vectors", u64, vp, u64, C.c_int, vp)
_sig("gz_count_vectors", u64, vp, u64, C.c_int)
_sig("gz_build_vertex_index", None, vp, u64, u64, vp, P(u64), P(u64))
_sig("gz_build_vertex_offsets", None, vp, u64, u64, vp)
_sig("gz_count_outdegrees", None, vp, u64, u64, vp)
_sig("gz_rmat_draw", None, P(RmatParams), vp)
_sig("gz_splitmix64", u64, u64)
_sig("gz_graph_load_pairs", C.c_int, u64, u64, vp, u64, vp)
_sig("gz_graph_build_pairs_device", C.c_int, u64, u64, vp, C.c_int, u32)
_sig("gz_graph_build_rmat_device", C.c_int, P(RmatParams), u64, C.c_int, u64, vp, u32)
_sig("gz_graph_device_vertex", u64, u64)
_sig("gz_graph_original_ids", None, vp)
_sig("gz_graph_indegrees", None, vp)
_sig("gz_push_queue_capacity", C.c_int, u32)
_sig("gz_graph_build_source_slices", C.c_int, u64, u32, u32, u64)
_sig("gz_pr_graph_replay", None, C.c_int)
_sig("gz_pr_persistent", C.c_int, C.c_int, u32)
_sig("gz_pr_fused_tail", None, C.c_int)
_sig("gz_probe_wide", None, C.c_int)
_sig("gz_pr_persistent_runs", u64)
_sig("gz_pr_persistent_spin_limit", None, u32)
_sig("gz_pr_persistent_gave_up", u64)
_sig("gz_cc_graph_replay", None, C.c_int)
_sig("gz_cc_push_tail", None, u64)
_sig("gz_cc_push_tail_rounds", u64)
_sig("gz_graph_replay_counts", None, C.POINTER(u64), C.POINTER(u64))
_sig("gz_frontier_compact_narrow", None, C.c_int)
_sig("gz_graph_source_slices_info", C.c_int, vp, P(u64), P(u32), P(u32), P(u64))
_sig("gz_slices_chunk_table", u64, vp, u32, u64, u32, vp, u64)
_sig("gz_slices_set_cold_spread", None, u32)
_sig("gz_slices_set_cold_stream", None, u32)
_sig("gz_slices_set_hot_cache", C.c_int, u32)
_sig("gz_graph_source_slices_download", C.c_int, vp, vp)
_sig("gz_graph_release", None)
_sig("gz_graph_num_vertices", u64)
_sig("gz_graph_num_edges", u64)
_sig("gz_graph_pull_vector_count", u64)
_sig("gz_graph_push_vector_count", u64)
_sig("gz_graph_download_vectors", None, C.c_int, vp)
_sig("gz_graph_download_offsets", None, C.c_int, vp)
_sig("gz_graph_outdegrees_host", P(C.c_double))
_sig("gz_run_options_default", None, P(RunOptions))
_sig("gz_execution_prepare", None, P(RunOptions))
_sig("gz_execution_run", None, P(RunOptions), P(RunResult))
_sig("gz_execution_records", u64, vp, u64)
_sig("gz_result_pr_ranks", None, vp)
_sig("gz_result_pr_props", None, vp)
_sig("gz_result_cc_labels", None, vp)
_sig("gz_result_bfs", None, vp, vp)
_sig("gz_pull_units_for", u64, u64)
_sig("gz_main", C.c_int, C.c_int, P(C.c_char_p))
_sig("gz_kernel_timing_enable", None, C.c_int)
_sig("gz_kernel_timing_reset", None)
_sig("gz_kernel_timing_read", None, C.c_int, P(C.c_double), P(u64))
_sig("gz_selector_set_weights", C.c_int, vp, vp, vp)
_sig("gz_selector_clear", None)
_sig("gz_selector_load_file", C.c_int, C.c_char_p)
_sig("gz_selector_save_file", C.c_int, C.c_char_p)
_sig("gz_selector_loaded", C.c_int)
_sig("gz_selector_feature_used", C.c_int, u32)
_sig("gz_selector_set_trust", None, C.c_double, C.c_double)
_sig("gz_selector_forward_batch", C.c_int, vp, vp, vp, u64)
_sig("gz_selector_gradient", C.c_float, vp, vp, vp, u64, vp)
_sig("gz_selector_train", C.c_int, vp, vp, vp, u64, u32, C.c_float, vp)

phase_config = PhaseConfig.in_dll(lib, "gz_phase_config")
phase_graph = PhaseGraph.in_dll(lib, "gz_graph")

SEL_PARAMS = 32 * 8 + 32 + 32 * 32 + 32 + 2 * 32 + 2


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(vp)


def device_count() -> int:
    return int(lib.gz_runtime_device_count())


# ---- graph construction (host C code) ---------------------------------------

def rmat_edges(scale: int, num_edges: int, seed: int, a=0.57, b=0.19, c=0.19, num_vertices: int = 0,
               edge_offset: int = 0) -> np.ndarray:
    """Raw seeded R-MAT edges as an (E, 2) uint64 array of (src, dst)."""
    p = RmatParams(scale, num_vertices, num_edges, a, b, c, seed, edge_offset)
    pairs = np.empty((num_edges, 2), dtype=np.uint64)
    lib.gz_rmat_draw(C.byref(p), _ptr(pairs))
    return pairs


def sort_edges(pairs: np.ndarray, by_dst: bool, dedupe: bool = True) -> np.ndarray:
    """Sorted copy grouped by dst (pull) or src (push), optionally deduplicated
    and without self-loops."""
    out = np.ascontiguousarray(pairs, dtype=np.uint64).copy()
    n = lib.gz_edges_sort(_ptr(out), out.shape[0], int(by_dst), int(dedupe))
    return out[:n]


@dataclass
class EdgeLists:
    num_vertices: int
    pull: np.ndarray  # (E, 2) grouped by dst
    push: np.ndarray  # (E, 2) grouped by src

    @property
    def num_edges(self) -> int:
        return int(self.pull.shape[0])


def make_graph(pairs: np.ndarray, num_vertices: int, symmetrize: bool = False) -> EdgeLists:
    if symmetrize:
        pairs = np.concatenate([pairs, pairs[:, ::-1]], axis=0)
    pull = sort_edges(pairs, by_dst=True)
    push = sort_edges(pull, by_dst=False, dedupe=False)
    return EdgeLists(num_vertices, pull, push)


def rmat_graph(scale: int, edge_factor: int, seed: int, symmetrize: bool = False) -> EdgeLists:
    nv = 1 << scale
    return make_graph(rmat_edges(scale, nv * edge_factor, seed), nv, symmetrize)


def planted_trees(first_id: int, trees: int, tree_size: int, seed: int) -> np.ndarray:
    """Edges (child, parent) of `trees` random recursive trees of `tree_size`
    vertices with ids from first_id on: vertex k of a tree attaches to a
    uniformly drawn earlier vertex of the same tree."""
    rng = np.random.default_rng(seed)
    k = np.arange(1, tree_size, dtype=np.uint64)
    out = []
    for t in range(trees):
        base = np.uint64(first_id + t * tree_size)
        parent = (rng.random(tree_size - 1) * k.astype(np.float64)).astype(np.uint64)
        out.append(np.stack([base + k, base + parent], axis=1))
    return np.concatenate(out) if out else np.zeros((0, 2), np.uint64)


WORKLOADS = {
    # name: (description, builder)
    "pr-rmat16": "PageRank graph: R-MAT scale 16, edge factor 16, a/b/c = 0.57/0.19/0.19, seeded",
    "cc-rmat22-sym-trees": "CC graph: symmetrised R-MAT scale 22, edge factor 16, plus 64 planted 1000-vertex trees",
    "bfs-rmat24": "BFS graph: R-MAT scale 24, edge factor 16",
    "pr-rmat20": "R-MAT scale 20, edge factor 16",
    "cc-rmat16-sym-trees": "symmetrised R-MAT scale 16 plus 64 planted 100-vertex trees (reduced CC shape)",
    "full-rmat26-fold41.7m": "R-MAT scale 26 with ids folded by modulo into 41.7M vertices, 1.494B raw edges "
                             "(about 1.47B after self-loop and duplicate removal); device ingress only",
}

# The full-scale power-law graph: scale-26 R-MAT draws folded into 41.7M ids.
# The raw count is chosen so that about 1.47B edges survive deduplication
# (kept fraction 0.949 / 0.958 / 0.965 measured at the same shape scaled down
# to scales 20 / 21 / 22, extrapolated to about 0.984 at scale 26).
FULL_VERTICES = 41_700_000
FULL_RAW_EDGES = 1_494_000_000


def workload_graph(name: str, seed: int = 1) -> EdgeLists:
    if name == "pr-rmat16":
        return rmat_graph(16, 16, seed)
    if name == "pr-rmat20":
        return rmat_graph(20, 16, seed)
    if name == "bfs-rmat24":
        return rmat_graph(24, 16, seed)
    if name in ("cc-rmat22-sym-trees", "cc-rmat16-sym-trees"):
        scale, tsize = (22, 1000) if name.startswith("cc-rmat22") else (16, 100)
        nv0 = 1 << scale
        rm = rmat_edges(scale, nv0 * 16, seed)
        tr = planted_trees(nv0, 64, tsize, seed + 1000)
        return make_graph(np.concatenate([rm, tr]), nv0 + 64 * tsize, symmetrize=True)
    raise KeyError(name)


def pack_vectors(pairs: np.ndarray, by_dst: bool) -> np.ndarray:
    pairs = np.ascontiguousarray(pairs, dtype=np.uint64)
    out = np.zeros((max(1, pairs.shape[0]), 4), dtype=np.uint64)
    n = lib.gz_pack_vectors(_ptr(pairs), pairs.shape[0], int(by_dst), _ptr(out))
    return out[:n]


def vertex_index(vectors: np.ndarray, num_vertices: int):
    idx = np.zeros(num_vertices + 64, dtype=np.uint64)
    s, e = u64(0), u64(0)
    v = np.ascontiguousarray(vectors, dtype=np.uint64)
    lib.gz_build_vertex_index(_ptr(v), v.shape[0], num_vertices, _ptr(idx), C.byref(s), C.byref(e))
    return idx, s.value, e.value


def vertex_offsets(vectors: np.ndarray, num_vertices: int) -> np.ndarray:
    off = np.zeros(num_vertices + 1, dtype=np.uint64)
    v = np.ascontiguousarray(vectors, dtype=np.uint64)
    lib.gz_build_vertex_offsets(_ptr(v), v.shape[0], num_vertices, _ptr(off))
    return off


def write_edge_list(path: str, num_vertices: int, pairs: np.ndarray) -> None:
    pairs = np.ascontiguousarray(pairs, dtype=np.uint64)
    rc = lib.gz_edge_list_write(path.encode(), num_vertices, pairs.shape[0], _ptr(pairs))
    if rc != 0:
        raise OSError(f"cannot write {path} ({rc})")


def write_graph_files(base: str, g: EdgeLists) -> None:
    write_edge_list(base + "-pull", g.num_vertices, g.pull)
    write_edge_list(base + "-push", g.num_vertices, g.push)


# ---- device side -------------------------------------------------------------

def init(device: int = 0) -> str:
    lib.gz_runtime_init(device)
    return lib.gz_runtime_device_name().decode()


def load(g: EdgeLists, pull: bool = True, push: bool = True) -> None:
    init()
    pl = np.ascontiguousarray(g.pull, dtype=np.uint64)
    ps = np.ascontiguousarray(g.push, dtype=np.uint64)
    lib.gz_graph_load_pairs(g.num_vertices, pl.shape[0], _ptr(pl) if pull else None, ps.shape[0],
                            _ptr(ps) if push else None)


RELABEL_BY_OUTDEGREE = 1
SLICE_CHUNK = np.dtype([("begin", np.uint64), ("count", np.uint32), ("record", np.uint32)])
SLICE_RECORD_BITS = 25
SLICE_BLOCK_VERTICES = 20480


def _ingress_flags(relabel: bool, slices_log2: int, slice_tiers: int = 1, slice_blocks: int = 0) -> int:
    return ((RELABEL_BY_OUTDEGREE if relabel else 0) | ((slices_log2 & 0x3F) << 8) | ((slice_tiers & 0x3) << 14)
            | ((slice_blocks & 0x7) << 16))


def load_pairs_device(num_vertices: int, pairs: np.ndarray, symmetrize: bool = False, relabel: bool = False,
                      slices_log2: int = 0, slice_tiers: int = 1, slice_blocks: int = 0) -> None:
    """Device ingress from raw (src, dst) pairs: sort, dedupe, pack in HBM.
    slices_log2 = k > 0 also builds the source-sliced pull layout with
    S = 2^k, slice_tiers tiers and slice_blocks LDS blocks of
    SLICE_BLOCK_VERTICES sources (and, with relabel, deals each tier's 8*S
    ranks of the degree order into its eight windows)."""
    init()
    p = np.ascontiguousarray(pairs, dtype=np.uint64).reshape(-1, 2)
    lib.gz_graph_build_pairs_device(num_vertices, p.shape[0], _ptr(p) if p.shape[0] else None, int(symmetrize),
                                    _ingress_flags(relabel, slices_log2, slice_tiers, slice_blocks))


def load_rmat_device(scale: int, num_edges: int, seed: int, symmetrize: bool = False, num_vertices: int = 0,
                     extra: np.ndarray | None = None, rmat_vertices: int = 0, a=0.57, b=0.19, c=0.19,
                     relabel: bool = False, slices_log2: int = 0, slice_tiers: int = 1,
                     slice_blocks: int = 0) -> None:
    """Device ingress of a seeded R-MAT graph (the stream of rmat_edges) plus
    optional extra host pairs; V = max(num_vertices, R-MAT vertex count).
    relabel renumbers vertices by out-degree on the device (results keep
    original ids); slices_log2, slice_tiers and slice_blocks as for
    load_pairs_device."""
    init()
    p = RmatParams(scale, rmat_vertices, num_edges, a, b, c, seed, 0)
    ex = np.ascontiguousarray(extra if extra is not None else np.zeros((0, 2), np.uint64), dtype=np.uint64)
    lib.gz_graph_build_rmat_device(C.byref(p), num_vertices, int(symmetrize), ex.shape[0],
                                   _ptr(ex) if ex.shape[0] else None,
                                   _ingress_flags(relabel, slices_log2, slice_tiers, slice_blocks))


def load_workload_device(name: str, seed: int = 1, relabel: bool = False, slices_log2: int = 0,
                         slice_tiers: int = 1, slice_blocks: int = 0) -> None:
    """Build a named workload with device ingress (same graph as workload_graph)."""
    kw = dict(relabel=relabel, slices_log2=slices_log2, slice_tiers=slice_tiers, slice_blocks=slice_blocks)
    if name in ("pr-rmat16", "pr-rmat20", "bfs-rmat24"):
        scale = {"pr-rmat16": 16, "pr-rmat20": 20, "bfs-rmat24": 24}[name]
        load_rmat_device(scale, (1 << scale) * 16, seed, **kw)
        return
    if name in ("cc-rmat22-sym-trees", "cc-rmat16-sym-trees"):
        scale, tsize = (22, 1000) if name.startswith("cc-rmat22") else (16, 100)
        nv0 = 1 << scale
        tr = planted_trees(nv0, 64, tsize, seed + 1000)
        load_rmat_device(scale, nv0 * 16, seed, symmetrize=True, num_vertices=nv0 + 64 * tsize, extra=tr, **kw)
        return
    if name == "full-rmat26-fold41.7m":
        load_rmat_device(26, FULL_RAW_EDGES, seed, rmat_vertices=FULL_VERTICES, **kw)
        return
    raise KeyError(name)


def build_source_slices(slice_vertices: int, tiers: int = 1, blocks: int = 0,
                        block_vertices: int = SLICE_BLOCK_VERTICES) -> None:
    """Source-sliced pull layout of the loaded graph (pull_variant
    "source_sliced" runs PageRank and CC over it): T = tiers windows of S
    vertices per XCD, and `blocks` LDS blocks of block_vertices sources carved
    from each tier-0 window (as many as fit)."""
    lib.gz_graph_build_source_slices(slice_vertices, tiers, blocks, block_vertices)


def source_slices_layout():
    """(S, T, G, H) of the current sliced layout, or None."""
    s, t, g, h = u64(0), u32(0), u32(0), u64(0)
    if not lib.gz_graph_source_slices_info(None, C.byref(s), C.byref(t), C.byref(g), C.byref(h)):
        return None
    return int(s.value), int(t.value), int(g.value), int(h.value)


def source_slices_info():
    """(vectors per slice [8 (G + T) + 1], S), or None without a sliced layout."""
    lay = source_slices_layout()
    if lay is None:
        return None
    S, T, G, _ = lay
    v = np.zeros(8 * (T + G) + 1, np.uint64)
    lib.gz_graph_source_slices_info(_ptr(v), None, None, None, None)
    return v, S


def source_slices_download():
    """(concatenated slice vectors as (N, 4) uint64, per-vertex slice mask)."""
    info = source_slices_info()
    if info is None:
        return None
    n = int(info[0].sum())
    vecs = np.zeros((max(n, 1), 4), np.uint64)
    mask = np.zeros(max(num_vertices(), 1), np.uint64)
    lib.gz_graph_source_slices_download(_ptr(vecs), _ptr(mask))
    return vecs[:n], mask[:num_vertices()]


def slices_chunk_table(vectors, chunk_len: int, cold_spread: bool = True) -> np.ndarray:
    """The chunk slots the source-sliced pull launches for slices of the
    given vector counts (8 L + 1 of them for L levels; host computation)."""
    v = np.ascontiguousarray(vectors, dtype=np.uint64)
    levels = (len(v) - 1) // 8
    if len(v) != 8 * levels + 1:
        raise ValueError("vector counts must have 8L + 1 entries")
    n = lib.gz_slices_chunk_table(_ptr(v), levels, chunk_len, int(cold_spread), None, 0)
    out = np.zeros(max(int(n), 1), SLICE_CHUNK)
    lib.gz_slices_chunk_table(_ptr(v), levels, chunk_len, int(cold_spread), _ptr(out), n)
    return out[:n]


def slices_set_cold_spread(on: bool) -> None:
    lib.gz_slices_set_cold_spread(int(on))


def slices_set_cold_stream(on: bool) -> None:
    lib.gz_slices_set_cold_stream(int(on))


def slices_set_hot_cache(vertices: int) -> None:
    """LDS cache of the hottest `vertices` sources of each slice (0 = off)."""
    if lib.gz_slices_set_hot_cache(vertices) != 0:
        raise ValueError(f"hot cache size {vertices} is not 0, 4096, 8192, 16384 or 20480")


def set_chunk_vectors(n: int) -> None:
    phase_config.chunk_vectors = n


def pr_graph_replay(on: bool) -> None:
    """PageRank iterations 2.. from a replayed HIP graph (default on)."""
    init()
    lib.gz_pr_graph_replay(int(on))


def probe_wide(on: bool) -> None:
    """The BFS probe engine also times the uncompacted pull, the compacted push and the uncompacted push."""
    init()
    lib.gz_probe_wide(int(on))


def pr_fused_tail(on: bool) -> None:
    """PageRank's merge, sum and vertex phase as one launch on small lists (default on)."""
    init()
    lib.gz_pr_fused_tail(int(on))


def pr_persistent(mode: int, workgroups_per_cu: int = 2) -> None:
    """PageRank in one cooperative launch: 0 never (default), 1 small graphs, 2 any graph."""
    init()
    if lib.gz_pr_persistent(int(mode), int(workgroups_per_cu)) != 0:
        raise ValueError(f"bad one-launch PageRank setting {mode}, {workgroups_per_cu}")


def pr_persistent_runs() -> int:
    """PageRank runs that took the one-launch path so far."""
    init()
    return int(lib.gz_pr_persistent_runs())


def pr_persistent_spin_limit(polls: int) -> None:
    """Polls a grid barrier of the one-launch PageRank makes before giving up."""
    init()
    lib.gz_pr_persistent_spin_limit(int(polls))


def pr_persistent_gave_up() -> int:
    """One-launch PageRank runs whose grid barrier gave up so far."""
    init()
    return int(lib.gz_pr_persistent_gave_up())


def cc_graph_replay(on: bool) -> None:
    """CC iterations with the engine chosen on the device, from a captured graph (default off)."""
    init()
    lib.gz_cc_graph_replay(int(on))


def cc_push_tail(max_vote: int = 4096) -> None:
    """CC push iterations over frontiers of statistic <= max_vote run back to
    back in one workgroup (default 4096, at most 16384; 0 turns it off)."""
    init()
    lib.gz_cc_push_tail(int(max_vote))


def cc_push_tail_rounds() -> int:
    """Push iterations the one-workgroup tail ran since the library was loaded."""
    init()
    return int(lib.gz_cc_push_tail_rounds())


def compact_narrow(on: bool) -> None:
    """Frontier compaction over 32-bit vector counts (default) or 64-bit offsets."""
    init()
    lib.gz_frontier_compact_narrow(int(on))


def graph_replay_counts() -> tuple[int, int]:
    """(graph captures, graph launches) so far."""
    init()
    c, n = u64(0), u64(0)
    lib.gz_graph_replay_counts(C.byref(c), C.byref(n))
    return int(c.value), int(n.value)


def set_vectors_per_lane(n: int) -> None:
    phase_config.vectors_per_lane = n


def set_frontier_skip(on: bool) -> None:
    phase_config.push_frontier_skip = int(on)


def options(algorithm="pr", engine="hybrid", iterations=1, stat="both", pull_variant="sched_aware",
            damping=0.875, root=0, profile=False, collect_stats=False, max_iterations=0,
            edge_only=False, vertex_only=False) -> RunOptions:
    o = RunOptions()
    lib.gz_run_options_default(C.byref(o))
    o.algorithm = ALG[algorithm]
    o.engine = ENGINE[engine]
    o.iterations = iterations
    o.stat_mode = STAT[stat]
    o.pull_variant = PULL_VARIANT[pull_variant]
    o.damping = damping
    o.root = root
    o.profile = int(profile)
    o.collect_stats = int(collect_stats)
    o.max_iterations = max_iterations
    o.edge_only = int(edge_only)
    o.vertex_only = int(vertex_only)
    return o


def run(o: RunOptions) -> RunResult:
    r = RunResult()
    lib.gz_execution_prepare(C.byref(o))
    lib.gz_execution_run(C.byref(o), C.byref(r))
    return r


def records() -> list[IterationRecord]:
    n = lib.gz_execution_records(None, 0)
    arr = (IterationRecord * max(1, n))()
    lib.gz_execution_records(C.cast(arr, vp), n)
    return list(arr[:n])


def num_vertices() -> int:
    return int(lib.gz_graph_num_vertices())


def pr_ranks() -> np.ndarray:
    out = np.empty(num_vertices(), dtype=np.float64)
    lib.gz_result_pr_ranks(_ptr(out))
    return out


def pr_props() -> np.ndarray:
    out = np.empty(num_vertices(), dtype=np.float64)
    lib.gz_result_pr_props(_ptr(out))
    return out


def cc_labels() -> np.ndarray:
    out = np.empty(num_vertices(), dtype=np.uint64)
    lib.gz_result_cc_labels(_ptr(out))
    return out


def bfs_result():
    n = num_vertices()
    parents = np.empty(n, dtype=np.int64)
    depths = np.empty(n, dtype=np.uint64)
    lib.gz_result_bfs(_ptr(parents), _ptr(depths))
    return parents, depths


def device_vectors(pull: bool) -> np.ndarray:
    n = lib.gz_graph_pull_vector_count() if pull else lib.gz_graph_push_vector_count()
    out = np.zeros((max(1, n), 4), dtype=np.uint64)
    lib.gz_graph_download_vectors(int(pull), _ptr(out))
    return out[:n]


def device_offsets(pull: bool) -> np.ndarray:
    out = np.zeros(num_vertices() + 1, dtype=np.uint64)
    lib.gz_graph_download_offsets(int(pull), _ptr(out))
    return out


def outdegrees() -> np.ndarray:
    """FP64 outdegrees of the loaded graph (host mirror, V entries)."""
    n = num_vertices()
    return np.ctypeslib.as_array(lib.gz_graph_outdegrees_host(), shape=(n,)).copy()


def original_ids() -> np.ndarray:
    """Original id of every device vertex (identity unless renumbered)."""
    out = np.empty(num_vertices(), dtype=np.uint64)
    lib.gz_graph_original_ids(_ptr(out))
    return out


def device_vertex(original: int) -> int:
    return int(lib.gz_graph_device_vertex(original))


def indegrees() -> np.ndarray:
    """In-degree of every device vertex (device numbering)."""
    out = np.empty(num_vertices(), dtype=np.uint32)
    lib.gz_graph_indegrees(_ptr(out))
    return out


def push_queue_capacity(pieces: int) -> None:
    """Overflow-queue capacity of the small-frontier push (default 131072)."""
    init()
    if lib.gz_push_queue_capacity(pieces) != 0:
        raise ValueError(f"queue capacity {pieces} outside 1..131072")


def num_edges() -> int:
    return int(lib.gz_graph_num_edges())


def release() -> None:
    lib.gz_graph_release()


KERNELS = ["pull_pr", "push_pr", "vertex_pr", "merge", "pull_cc", "push_cc", "pull_bfs_find",
           "pull_bfs_visit", "push_bfs", "compact", "reduce"]


def kernel_timing(on: bool) -> None:
    lib.gz_kernel_timing_enable(int(on))
    lib.gz_kernel_timing_reset()


def kernel_times() -> dict:
    """{kernel: (total_ms, launches)} recorded with HIP events since the last reset."""
    out = {}
    for i, k in enumerate(KERNELS):
        ms, n = C.c_double(), u64()
        lib.gz_kernel_timing_read(i, C.byref(ms), C.byref(n))
        out[k] = (ms.value, n.value)
    return out


def pull_vector_count() -> int:
    return int(lib.gz_graph_pull_vector_count())


def push_vector_count() -> int:
    return int(lib.gz_graph_push_vector_count())


def main(argv: list[str]) -> int:
    args = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    return int(lib.gz_main(len(argv), args))


# ---- breadth-first searches from many roots at once --------------------------

BFS_BATCH_MAX_ROOTS = 64
_sig("gz_bfs_batch_run", C.c_int, vp, u32, P(RunResult))
_sig("gz_bfs_batch_result", None, u32, vp, vp)
_sig("gz_bfs_batch_release", None)


def bfs_batch(roots) -> RunResult:
    """BFS from up to 64 roots (original ids) in one pass per level."""
    r = np.ascontiguousarray(roots, dtype=np.uint64)
    res = RunResult()
    if lib.gz_bfs_batch_run(_ptr(r), r.shape[0], C.byref(res)) != 0:
        raise ValueError(f"BFS batch of {r.shape[0]} roots failed (see stderr)")
    return res


def bfs_batch_result(search: int, parents: bool = True):
    """(parents or None, depths) of search `search` of the last batch."""
    n = num_vertices()
    depths = np.empty(n, dtype=np.uint64)
    par = np.empty(n, dtype=np.int64) if parents else None
    lib.gz_bfs_batch_result(search, _ptr(par) if parents else None, _ptr(depths))
    return par, depths


# ---- learned selector -------------------------------------------------------

def selector_forward(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.empty((X.shape[0], 2), dtype=np.float32)
    lib.gz_selector_forward_batch(_ptr(params), _ptr(X), _ptr(Y), X.shape[0])
    return Y


def selector_gradient(params: np.ndarray, X: np.ndarray, T: np.ndarray):
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    g = np.empty(SEL_PARAMS, dtype=np.float32)
    loss = lib.gz_selector_gradient(_ptr(params), _ptr(X), _ptr(T), X.shape[0], _ptr(g))
    return float(loss), g


def selector_train(params: np.ndarray, X: np.ndarray, T: np.ndarray, steps: int, lr: float):
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    hist = np.empty(steps, dtype=np.float32)
    lib.gz_selector_train(_ptr(p), _ptr(X), _ptr(T), X.shape[0], steps, lr, _ptr(hist))
    return p, hist


def selector_set(params: np.ndarray, mean: np.ndarray, std: np.ndarray) -> None:
    p = np.ascontiguousarray(params, dtype=np.float32)
    m = np.ascontiguousarray(mean, dtype=np.float32)
    s = np.ascontiguousarray(std, dtype=np.float32)
    lib.gz_selector_set_weights(_ptr(p), _ptr(m), _ptr(s))


def selector_clear() -> None:
    lib.gz_selector_clear()


def selector_load(path: str) -> None:
    rc = lib.gz_selector_load_file(path.encode())
    if rc != 0:
        raise OSError(f"cannot load selector weights from {path} ({rc})")


def selector_save(path: str) -> None:
    rc = lib.gz_selector_save_file(path.encode())
    if rc != 0:
        raise OSError(f"cannot write selector weights to {path} ({rc})")


TRUST_ABOVE, TRUST_BELOW = 1.0, 32.0  # the engine's default band


def selector_set_trust(above: float, below: float | None = None) -> None:
    """The learned engine follows the model only while threshold/below <=
    vote <= above*threshold (defaults 1 and 32; a factor < 1 lifts that side;
    one factor sets both)."""
    init()
    lib.gz_selector_set_trust(float(above), float(above if below is None else below))


def selector_feature_used(i: int) -> bool:
    """Whether the loaded selector model uses input feature i (std > 0 in its weights)."""
    return bool(lib.gz_selector_feature_used(int(i)))


def selector_loaded() -> bool:
    return bool(lib.gz_selector_loaded())


# ---- batches of independent graphs ------------------------------------------

BATCH_BUCKETS = 4


class BatchConfig(C.Structure):
    _fields_ = [("vectors_per_lane", u32), ("workgroup_size", u32), ("chunk", u32 * BATCH_BUCKETS)]

    def as_dict(self) -> dict:
        return {"vectors_per_lane": self.vectors_per_lane, "workgroup_size": self.workgroup_size,
                "chunk": list(self.chunk)}


_sig("gz_graph_build_rmat_union_device", C.c_int, P(RmatParams), u32, C.c_int, vp)
_sig("gz_batch_attach", C.c_int, vp, u32)
_sig("gz_batch_release", None)
_sig("gz_batch_bucket_of_graph", u32, u32)
_sig("gz_batch_set_config", C.c_int, P(BatchConfig))
_sig("gz_batch_get_config", None, P(BatchConfig))
_sig("gz_batch_pr_prepare", None, C.c_double)
_sig("gz_batch_pr_run", C.c_double, u32, C.c_int)
_sig("gz_batch_time_pull", C.c_double, P(BatchConfig), u32, u32)
_sig("gz_batch_pr_ranks", None, vp)


def batch_specs(count: int = 1024, seed: int = 4, smin: int = 10, smax: int = 17, efmin: int = 4, efmax: int = 32):
    """(scale, edge factor, seed) of a batch: scale s in [smin, smax] drawn
    with P(s) proportional to 1/s ("log-uniform" over the integers), edge
    factor uniform over the integers [efmin, efmax], graph seeds seed*100000+i."""
    rng = np.random.default_rng(seed)
    scales = np.arange(smin, smax + 1)
    p = 1.0 / scales
    p /= p.sum()
    s = rng.choice(scales, size=count, p=p)
    ef = rng.integers(efmin, efmax + 1, size=count)
    return [(int(a), int(b), seed * 100000 + i) for i, (a, b) in enumerate(zip(s, ef))]


def load_batch_device(specs, symmetrize: bool = False) -> np.ndarray:
    """Device ingress of the disjoint union of R-MAT graphs (scale, ef, seed);
    returns the vertex bases (len(specs)+1) and attaches the batch."""
    init()
    params = (RmatParams * len(specs))()
    for i, (scale, ef, seed) in enumerate(specs):
        params[i] = RmatParams(scale, 0, (1 << scale) * ef, 0.57, 0.19, 0.19, seed, 0)
    bases = np.zeros(len(specs) + 1, dtype=np.uint64)
    lib.gz_graph_build_rmat_union_device(params, len(specs), int(symmetrize), _ptr(bases))
    if lib.gz_batch_attach(_ptr(bases), len(specs)) != 0:
        raise RuntimeError("batch attach failed")
    return bases


def batch_set_config(vectors_per_lane=1, workgroup_size=256, chunk=(512, 512, 512, 512)) -> None:
    c = BatchConfig(vectors_per_lane, workgroup_size, (u32 * BATCH_BUCKETS)(*chunk))
    if lib.gz_batch_set_config(C.byref(c)) != 0:
        raise ValueError(f"invalid batch configuration {c.as_dict()}")


def batch_get_config() -> dict:
    c = BatchConfig()
    lib.gz_batch_get_config(C.byref(c))
    return c.as_dict()


def batch_pr(iterations: int = 10, damping: float = 0.875, use_graph: bool = True) -> float:
    """Batched PageRank from the initial ranks; returns the device ms of the iterations."""
    lib.gz_batch_pr_prepare(damping)
    return float(lib.gz_batch_pr_run(iterations, int(use_graph)))


def batch_pr_ranks() -> np.ndarray:
    out = np.empty(num_vertices(), dtype=np.float64)
    lib.gz_batch_pr_ranks(_ptr(out))
    return out


def batch_time_pull(vectors_per_lane: int, workgroup_size: int, chunk, bucket: int, reps: int = 3) -> float:
    c = BatchConfig(vectors_per_lane, workgroup_size, (u32 * BATCH_BUCKETS)(*chunk))
    return float(lib.gz_batch_time_pull(C.byref(c), bucket, reps))
