"""ctypes binding of the CPU oracle (oracle/_build/liborcl.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU reference.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborcl.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"oracle not built ({LIB_PATH}); run `make -C oracle`")
lib = C.CDLL(LIB_PATH)
u64, u32, vp = C.c_uint64, C.c_uint32, C.c_void_p


class Options(C.Structure):
    _fields_ = [("threads", u32), ("units_per_thread", u32), ("engine", u32), ("stat_mode", u32),
                ("traditional_pull", u32), ("iterations", u32), ("max_iterations", u32),
                ("damping", C.c_double), ("root", u64)]


class Iteration(C.Structure):
    _fields_ = [("engine", u32), ("pad", u32), ("vote", u64)]


class Result(C.Structure):
    _fields_ = [("iterations", u64), ("pull_iterations", u64), ("push_iterations", u64),
                ("elapsed_ms", C.c_double)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


P = C.POINTER
_sig("orc_pack", u64, vp, u64, C.c_int, vp)
_sig("orc_vertex_index", None, vp, u64, u64, vp, P(u64), P(u64))
_sig("orc_outdegrees", None, vp, u64, u64, vp)
_sig("orc_unit_range", None, u64, u64, u64, P(u64), P(u64))
_sig("orc_merge", None, vp, u64, vp)
_sig("orc_shared_vertex", u64, vp)
_sig("orc_graph_create", vp, u64, vp, u64, vp, u64)
_sig("orc_graph_from_vectors", vp, u64, vp, u64, vp, u64)
_sig("orc_graph_destroy", None, vp)
_sig("orc_graph_pull_count", u64, vp)
_sig("orc_graph_push_count", u64, vp)
_sig("orc_pagerank", C.c_int, vp, P(Options), vp, vp, P(Result))
_sig("orc_cc", C.c_int, vp, P(Options), vp, vp, u64, P(Result))
_sig("orc_cc_from", C.c_int, vp, P(Options), vp, vp, vp, u64, P(Result))
_sig("orc_bfs", C.c_int, vp, P(Options), vp, vp, vp, u64, P(Result))
_sig("orc_bfs_certificate_violations", u64, vp, u64, vp, vp, u32)
_sig("orc_mlp_forward", None, vp, vp, vp)
_sig("orc_mlp_gradient", C.c_float, vp, vp, vp, u64, vp)

ENGINE = {"hybrid": 0, "pull": 1, "push": 2}
STAT = {"both": 0, "outdeg": 1, "count": 2}
MERGE_DTYPE = np.dtype([("initial", "<u8"), ("final", "<u8"), ("partial", "<f8"), ("pad", "<u8")])


def _ptr(a):
    return a.ctypes.data_as(vp)


def pack(pairs: np.ndarray, by_dst: bool) -> np.ndarray:
    pairs = np.ascontiguousarray(pairs, dtype=np.uint64)
    out = np.zeros((max(1, pairs.shape[0]), 4), dtype=np.uint64)
    n = lib.orc_pack(_ptr(pairs), pairs.shape[0], int(by_dst), _ptr(out))
    return out[:n]


def vertex_index(vectors: np.ndarray, num_vertices: int):
    v = np.ascontiguousarray(vectors, dtype=np.uint64)
    idx = np.zeros(num_vertices + 64, dtype=np.uint64)
    s, e = u64(0), u64(0)
    lib.orc_vertex_index(_ptr(v), v.shape[0], num_vertices, _ptr(idx), C.byref(s), C.byref(e))
    return idx, s.value, e.value


def outdegrees(pairs: np.ndarray, num_vertices: int) -> np.ndarray:
    pairs = np.ascontiguousarray(pairs, dtype=np.uint64)
    out = np.zeros(num_vertices + 8, dtype=np.float64)
    lib.orc_outdegrees(_ptr(pairs), pairs.shape[0], num_vertices, _ptr(out))
    return out


def unit_range(nvec: int, units: int, unit: int):
    f, l = u64(0), u64(0)
    lib.orc_unit_range(nvec, units, unit, C.byref(f), C.byref(l))
    return f.value, l.value


def merge(records: np.ndarray, acc: np.ndarray) -> None:
    assert records.dtype == MERGE_DTYPE and acc.dtype == np.float64
    lib.orc_merge(_ptr(records), records.shape[0], _ptr(acc))


class Graph:
    def __init__(self, num_vertices: int, pull: np.ndarray | None, push: np.ndarray | None):
        self._pull = None if pull is None else np.ascontiguousarray(pull, dtype=np.uint64)
        self._push = None if push is None else np.ascontiguousarray(push, dtype=np.uint64)
        self.num_vertices = num_vertices
        self.h = lib.orc_graph_create(
            num_vertices,
            None if self._pull is None else _ptr(self._pull), 0 if self._pull is None else self._pull.shape[0],
            None if self._push is None else _ptr(self._push), 0 if self._push is None else self._push.shape[0])

    @classmethod
    def from_vectors(cls, num_vertices: int, pull: np.ndarray | None, push: np.ndarray | None) -> "Graph":
        """Graph over already packed (N, 4) uint64 Vector-Sparse lists, kept
        by reference (no copy), e.g. a device ingress's lists copied back."""
        self = cls.__new__(cls)
        self._pull = None if pull is None else np.ascontiguousarray(pull, dtype=np.uint64)
        self._push = None if push is None else np.ascontiguousarray(push, dtype=np.uint64)
        self.num_vertices = num_vertices
        self.h = lib.orc_graph_from_vectors(
            num_vertices,
            None if self._pull is None else _ptr(self._pull), 0 if self._pull is None else self._pull.shape[0],
            None if self._push is None else _ptr(self._push), 0 if self._push is None else self._push.shape[0])
        return self

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_graph_destroy(self.h)
            self.h = None

    @property
    def pull_count(self):
        return int(lib.orc_graph_pull_count(self.h))


def options(threads=1, engine="hybrid", stat="both", traditional=False, iterations=1, damping=0.875, root=0,
            max_iterations=0, units_per_thread=32) -> Options:
    return Options(threads, units_per_thread, ENGINE[engine], STAT[stat], int(traditional), iterations,
                   max_iterations, damping, root)


def pagerank(g: Graph, o: Options):
    props = np.empty(g.num_vertices, dtype=np.float64)
    ranks = np.empty(g.num_vertices, dtype=np.float64)
    r = Result()
    lib.orc_pagerank(g.h, C.byref(o), _ptr(props), _ptr(ranks), C.byref(r))
    return props, ranks, r


def cc(g: Graph, o: Options, log_cap: int = 4096, init: np.ndarray | None = None):
    """Labels from min-label propagation; init (V uint64, < 2^53) replaces the
    starting labels, which are the vertex ids otherwise."""
    labels = np.empty(g.num_vertices, dtype=np.uint64)
    log = (Iteration * log_cap)()
    r = Result()
    if init is None:
        lib.orc_cc(g.h, C.byref(o), _ptr(labels), C.cast(log, vp), log_cap, C.byref(r))
    else:
        init = np.ascontiguousarray(init, dtype=np.uint64)
        assert init.shape[0] >= g.num_vertices
        lib.orc_cc_from(g.h, C.byref(o), _ptr(init), _ptr(labels), C.cast(log, vp), log_cap, C.byref(r))
    return labels, [(x.engine, x.vote) for x in log[: min(r.iterations, log_cap)]], r


def bfs(g: Graph, o: Options, log_cap: int = 4096):
    parents = np.empty(g.num_vertices, dtype=np.int64)
    depths = np.empty(g.num_vertices, dtype=np.uint64)
    log = (Iteration * log_cap)()
    r = Result()
    rc = lib.orc_bfs(g.h, C.byref(o), _ptr(parents), _ptr(depths), C.cast(log, vp), log_cap, C.byref(r))
    if rc != 0:
        raise ValueError("invalid BFS root")
    return parents, depths, [(x.engine, x.vote) for x in log[: min(r.iterations, log_cap)]], r


def bfs_certificate_violations(g: Graph, root: int, parents: np.ndarray, depths: np.ndarray, threads: int = 1) -> int:
    """Vertices violating the BFS certificate over the pull list (0: depths are
    the BFS distances from root and the parents form a BFS tree)."""
    p = np.ascontiguousarray(parents, dtype=np.int64)
    d = np.ascontiguousarray(depths, dtype=np.uint64)
    assert p.shape[0] >= g.num_vertices and d.shape[0] >= g.num_vertices
    return int(lib.orc_bfs_certificate_violations(g.h, root, _ptr(p), _ptr(d), threads))


def mlp_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty(2, dtype=np.float32)
    lib.orc_mlp_forward(_ptr(p), _ptr(xx), _ptr(y))
    return y


def mlp_gradient(params: np.ndarray, X: np.ndarray, T: np.ndarray):
    p = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    g = np.empty(p.shape[0], dtype=np.float32)
    loss = lib.orc_mlp_gradient(_ptr(p), _ptr(X), _ptr(T), X.shape[0], _ptr(g))
    return float(loss), g
