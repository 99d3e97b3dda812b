"""The allocation-free PageRank iteration (gz_pagerank_iteration_noalloc)
against the existing entry points, through the C ABI.

It runs the kernels of the scheduler-aware pull, the merge, the sum S and the
vertex phase over caller-owned arrays with its scratch in a fixed device
array, so its ranks must be bit-identical to a run of the ordinary driver
with the same list and settings.  Two calls in a row, the second on a smaller
graph (the shared scratch keeps the first call's contents), and the library
must make no device allocation or free during either.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind(gz):
    f = gz.lib.gz_pagerank_iteration_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
    gz.lib.gz_device_allocation_events.restype = C.c_uint64
    gz.lib.gz_device_malloc.restype = C.c_void_p
    gz.lib.gz_device_malloc.argtypes = [C.c_size_t]
    gz.lib.gz_device_free.argtypes = [C.c_void_p]
    gz.lib.gz_copy_to_device.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    gz.lib.gz_copy_to_host.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    return f


def _reference_props(gz, g, iterations, damping, chunk):
    """props after `iterations` iterations of the ordinary driver loop."""
    gz.load(g)
    gz.set_chunk_vectors(chunk)
    try:
        gz.run(gz.options("pr", iterations=iterations, damping=damping))
    finally:
        gz.set_chunk_vectors(512)
    return gz.pr_props()


def _noalloc_props(gz, f, g, iterations, damping, chunk):
    gz.load(g)
    V = g.num_vertices
    od = gz.outdegrees()
    nv = float(V)
    props0 = (1.0 / nv) / np.where(od == 0.0, nv, od)
    lib = gz.lib
    d_props = lib.gz_device_malloc((V + 8) * 8)
    d_acc = lib.gz_device_malloc((V + 8) * 8)  # zero-filled
    try:
        lib.gz_copy_to_device(d_props, props0.ctypes.data, V * 8)
        pg = gz.phase_graph
        gz.phase_config.damping = damping
        gz.set_chunk_vectors(chunk)
        lib.gz_synchronize()
        before = lib.gz_device_allocation_events()
        for _ in range(iterations):
            assert f(pg.pull_vectors, pg.pull_vector_count, V, pg.vertex_outdegrees, d_props, d_acc) == 0
        lib.gz_synchronize()
        assert lib.gz_device_allocation_events() == before, "the allocation-free iteration allocated"
        out = np.empty(V, np.float64)
        lib.gz_copy_to_host(out.ctypes.data, d_props, V * 8)
        return out
    finally:
        gz.set_chunk_vectors(512)
        gz.phase_config.damping = 0.875
        lib.gz_device_free(d_props)
        lib.gz_device_free(d_acc)


def test_noalloc_iteration_matches_entry_points(gz, orc, device):
    f = _bind(gz)
    big = gz.rmat_graph(16, 16, seed=3)
    small = gz.rmat_graph(11, 8, seed=4)
    for g, chunk in ((big, 512), (small, 256)):
        want = _reference_props(gz, g, 20, 0.85, chunk)
        got = _noalloc_props(gz, f, g, 20, 0.85, chunk)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), g.num_vertices
        _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=20, damping=0.85))
        od = gz.outdegrees()
        ranks = got * np.where(od == 0.0, float(g.num_vertices), od)
        assert float(np.max(np.abs(ranks - ref) / ref)) <= 1e-12


def test_noalloc_rejects_lists_beyond_its_scratch(gz, device):
    f = _bind(gz)
    g = gz.rmat_graph(8, 4, seed=1)
    gz.load(g)
    pg = gz.phase_graph
    gz.set_chunk_vectors(64)
    try:
        # 2^19 units of 64 vectors is the limit; a count one vector past it is
        # rejected before any launch (the pointers are never read)
        assert f(pg.pull_vectors, (1 << 19) * 64 + 1, g.num_vertices, pg.vertex_outdegrees, None, None) == -1
    finally:
        gz.set_chunk_vectors(512)
