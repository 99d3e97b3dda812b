"""The drop-in boundary driven from outside the library: the reference's
driver loops replayed through the exported per-phase symbols only (the names
and argument meaning of phases.h:53-86), as a maintainer's driver would call
them, and compared with the oracle.

  PageRank   execution_pr.c:99-158: reset, pull, merge, write S, vertex phase
  CC         execution_cc.c:94-268: engine by the E/5 rule on the vote,
             edge phase, vote from phase_op_combine_global_var_from_buf,
             HasInfo / HasInfo* swap, vertex phase
  BFS        execution_bfs.c:120-276: the same loop with the level set before
             each edge phase

The only library calls besides the phase symbols are the graph load, the
vertex-array initialisation (gz_execution_prepare, the analogue of
graph_helper_initialize_vertex_info), the reduce and merge buffers and the
result getters.  Bars: PageRank within 1e-12 of the oracle; CC labels, BFS
depths, parents and the per-iteration engine and vote sequence bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

u64, vp = C.c_uint64, C.c_void_p


def _bind(gz):
    lib = gz.lib
    for alg in ("pr", "cc", "bfs"):
        for kind in ("pull", "push"):
            f = getattr(lib, f"perform_edge_{kind}_phase_{alg}")
            f.restype, f.argtypes = None, [vp, u64]
        f = getattr(lib, f"perform_vertex_phase_{alg}")
        f.restype, f.argtypes = None, [u64, u64, vp]
    lib.phase_op_reset_global_accum.restype = None
    lib.phase_op_write_global_accum_to_buf.restype, lib.phase_op_write_global_accum_to_buf.argtypes = None, [vp]
    lib.phase_op_combine_global_var_from_buf.restype = u64
    lib.phase_op_combine_global_var_from_buf.argtypes = [vp]
    lib.edge_pull_op_merge_with_merge_buffer.restype = None
    lib.edge_pull_op_merge_with_merge_buffer.argtypes = [vp, u64, vp]
    lib.gz_reduce_buffer_create.restype = vp
    lib.gz_reduce_buffer_destroy.argtypes = [vp]
    lib.gz_merge_buffer_allocate.argtypes = [u64]
    return lib


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("engine", ["pull", "push"])
def test_pagerank_loop_through_phase_symbols(gz, orc, device, engine):
    lib = _bind(gz)
    g = gz.rmat_graph(13, 16, seed=8)
    gz.load(g)
    pg = gz.phase_graph
    V = g.num_vertices
    lib.gz_execution_prepare(C.byref(gz.options("pr", damping=0.85)))
    rb = lib.gz_reduce_buffer_create()
    units = lib.gz_pull_units_for(pg.pull_vector_count)
    lib.gz_merge_buffer_allocate(units)
    merge_buffer = vp.in_dll(lib, "gz_merge_buffer")
    try:
        for _ in range(20):
            lib.phase_op_reset_global_accum()
            if engine == "pull":
                lib.perform_edge_pull_phase_pr(pg.pull_vectors, pg.pull_vector_count)
                lib.edge_pull_op_merge_with_merge_buffer(merge_buffer, units, pg.vertex_accumulators)
            else:
                lib.perform_edge_push_phase_pr(pg.push_vectors, pg.push_vector_count)
            lib.phase_op_write_global_accum_to_buf(rb)
            lib.perform_vertex_phase_pr(0, V, rb)
        lib.gz_synchronize()
    finally:
        lib.gz_reduce_buffer_destroy(rb)
    ranks = gz.pr_ranks()
    _, ref, _ = orc.pagerank(orc.Graph(V, g.pull, g.push), orc.options(iterations=20, damping=0.85, engine=engine))
    err = _rel(ranks, ref)
    print(f"PageRank {engine} loop through the phase symbols: max relative error {err:.3e}")
    assert err <= 1e-12
    assert abs(ranks.sum() - 1.0) < 1e-10


def _hybrid_loop(gz, lib, bfs, V, E, root=0):
    """The CC / BFS driver loop over the phase symbols; returns the
    (engine, vote) sequence."""
    pg = gz.phase_graph
    od = gz.outdegrees()
    vote = (int(od[root]) + 1) if bfs else E + V
    threshold = E // 5
    rb = lib.gz_reduce_buffer_create()
    log = []
    alg = "bfs" if bfs else "cc"
    try:
        ctr = 0
        while True:
            ctr += 1
            pull = vote > threshold
            gz.phase_config.bfs_level = ctr
            lib.phase_op_reset_global_accum()
            if pull:
                getattr(lib, f"perform_edge_pull_phase_{alg}")(pg.pull_vectors, pg.pull_vector_count)
            else:
                getattr(lib, f"perform_edge_push_phase_{alg}")(pg.push_vectors, pg.push_vector_count)
            lib.phase_op_write_global_accum_to_buf(rb)
            log.append((1 if pull else 2, vote))
            vote = int(lib.phase_op_combine_global_var_from_buf(rb))
            if vote == 0:
                break
            # execution_cc.c:245-250: HasInfo and HasInfo* trade places
            pg.frontier_has_info, pg.vertex_accumulators = pg.vertex_accumulators, pg.frontier_has_info
            getattr(lib, f"perform_vertex_phase_{alg}")(0, V, None)
            assert ctr < 10_000
    finally:
        lib.gz_reduce_buffer_destroy(rb)
    return log


@pytest.mark.parametrize("case", ["rmat13", "planted_trees"])
def test_cc_loop_through_phase_symbols(gz, orc, device, case):
    lib = _bind(gz)
    g = gz.rmat_graph(13, 16, seed=8) if case == "rmat13" else graphs.planted_trees(gz, base_scale=12)
    gz.load(g)
    lib.gz_execution_prepare(C.byref(gz.options("cc")))
    log = _hybrid_loop(gz, lib, False, g.num_vertices, g.num_edges)
    want, _, _ = orc.cc(orc.Graph(g.num_vertices, g.pull, g.push), orc.options())
    assert np.array_equal(gz.cc_labels(), want)
    assert {e for e, _ in log} == {1, 2}


@pytest.mark.parametrize("root", [0, 77, 1234])
def test_bfs_loop_through_phase_symbols(gz, orc, device, root):
    lib = _bind(gz)
    g = gz.rmat_graph(13, 16, seed=8)
    gz.load(g)
    lib.gz_execution_prepare(C.byref(gz.options("bfs", root=root)))
    log = _hybrid_loop(gz, lib, True, g.num_vertices, g.num_edges, root)
    par, dep, want_log, _ = orc.bfs(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(root=root))
    p, d = gz.bfs_result()
    assert np.array_equal(d, dep)
    assert np.array_equal(p, par)
    assert log == want_log
