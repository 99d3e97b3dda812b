"""PageRank with fewer launches per iteration against the per-phase launches
they replace (execution_pr.c:99-158: pull, merge, sum, vertex phase):
  - the fused tail (gz_pr_vertex_merged): merge, sum of S and vertex phase
    in one launch after each pull, on by default for small pull lists;
  - every iteration in one cooperative launch (gz_pr_run_persistent), off
    by default.

Bars: bit-identical properties to the per-phase launches (the same
arithmetic in the same order); within 1e-12 of the oracle; a grid barrier
that gives up (forced by a spin limit of zero polls) leaves the kernel and
the run restarts with the per-phase launches, with the same results.
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def _per_phase(gz, opts):
    """Four launches per iteration, no replay."""
    gz.pr_fused_tail(False)
    gz.pr_graph_replay(False)
    try:
        gz.run(opts)
        return gz.pr_props().view(np.uint64).copy()
    finally:
        gz.pr_fused_tail(True)
        gz.pr_graph_replay(True)


def _one_launch(gz, opts, mode=1, per_cu=2):
    gz.pr_persistent(mode, per_cu)
    try:
        n0 = gz.pr_persistent_runs()
        r = gz.run(opts)
        assert gz.pr_persistent_runs() == n0 + 1, "the one-launch path did not run"
        assert r.iterations == r.pull_iterations == opts.iterations
        return gz.pr_props().view(np.uint64).copy()
    finally:
        gz.pr_persistent(0)


def _cases(gz):
    small = dict(graphs.small_cases(gz))
    return {
        "rmat13": gz.rmat_graph(13, 16, seed=41),
        "hub": graphs.hub_graph(gz),
        "planted_trees": graphs.planted_trees(gz),
        "star_in": small["star_in"],
        "single_edge": small["single_edge"],
        "isolated_tail": small["isolated_tail"],
    }


@pytest.mark.parametrize("case", ["rmat13", "hub", "planted_trees", "star_in", "single_edge", "isolated_tail"])
def test_one_launch_equals_per_phase(gz, orc, device, case):
    g = _cases(gz)[case]
    gz.load(g)
    _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=20, damping=0.85))
    try:
        for chunk, vpl in ((64, 1), (512, 2), (4096, 4)):
            gz.set_chunk_vectors(chunk)
            gz.set_vectors_per_lane(vpl)
            gz.load(g)
            for iters in (1, 20):
                opts = gz.options("pr", iterations=iters, damping=0.85)
                want = _per_phase(gz, opts)
                got = _one_launch(gz, opts)
                assert np.array_equal(got, want), (case, chunk, vpl, iters)
            ranks = gz.pr_ranks()
            err = float(np.max(np.abs(ranks - ref) / np.maximum(np.abs(ref), 1e-300)))
            assert err <= 1e-12, (case, chunk, vpl, err)
    finally:
        gz.set_chunk_vectors(512)
        gz.set_vectors_per_lane(1)


@pytest.mark.parametrize("case", ["rmat13", "hub", "planted_trees", "star_in", "single_edge", "isolated_tail"])
def test_fused_tail_equals_per_phase(gz, orc, device, case):
    # the default path for these lists: pull, then one launch for the merge,
    # the sum and the vertex phase; eager and replayed from a captured graph
    g = _cases(gz)[case]
    gz.load(g)
    try:
        for chunk, vpl in ((64, 1), (512, 2), (4096, 4)):
            gz.set_chunk_vectors(chunk)
            gz.set_vectors_per_lane(vpl)
            gz.load(g)
            for iters in (1, 20, 20, 20, 7):  # eager, eager, captured, replayed, eager
                opts = gz.options("pr", iterations=iters, damping=0.85)
                want = _per_phase(gz, opts)
                r = gz.run(opts)
                assert r.iterations == r.pull_iterations == iters
                assert np.array_equal(gz.pr_props().view(np.uint64), want), (case, chunk, vpl, iters)
    finally:
        gz.set_chunk_vectors(512)
        gz.set_vectors_per_lane(1)


def test_fused_tail_limit(gz, device):
    # up to 4096 chunks take the fused tail, more the four launches; both
    # give the four launches' bits
    g = gz.rmat_graph(17, 16, seed=5)  # about 8000 chunks of 64 vectors, 4000 of 128
    seen = set()
    try:
        for chunk in (64, 128):
            gz.set_chunk_vectors(chunk)
            gz.load(g)
            opts = gz.options("pr", iterations=5, damping=0.85)
            got = _one_step(gz, opts)  # the run sizes the merge buffer
            nchunks = int(gz.lib.gz_pull_units_for(gz.phase_graph.pull_vector_count))
            applies = bool(gz.lib.gz_pr_fused_tail_applies())
            assert applies == (nchunks <= 4096), (chunk, nchunks)
            seen.add(applies)
            assert np.array_equal(got, _per_phase(gz, opts)), chunk
    finally:
        gz.set_chunk_vectors(512)
    assert seen == {True, False}


def _one_step(gz, opts):
    gz.run(opts)
    return gz.pr_props().view(np.uint64).copy()


@pytest.mark.parametrize("per_cu", [1, 2, 4])
def test_one_launch_grid_sizes(gz, device, per_cu):
    g = gz.rmat_graph(14, 16, seed=11)
    gz.load(g)
    opts = gz.options("pr", iterations=10, damping=0.85)
    want = _per_phase(gz, opts)
    for _ in range(2):  # a second run reuses the barrier words
        assert np.array_equal(_one_launch(gz, opts, per_cu=per_cu), want)


def test_one_launch_forced_on_a_larger_graph(gz, device):
    # R-MAT s20 needs more than the first stage's 4096 partials per workgroup
    # (several slices of S) and is above the default size bound
    scale = 20
    gz.load_rmat_device(scale, (1 << scale) * 16, 7)
    opts = gz.options("pr", iterations=5, damping=0.85)
    n0 = gz.pr_persistent_runs()
    gz.pr_persistent(1)
    try:
        gz.run(opts)
    finally:
        gz.pr_persistent(0)
    assert gz.pr_persistent_runs() == n0, "the size bound should leave this graph to the per-phase path"
    want = _per_phase(gz, opts)
    assert np.array_equal(_one_launch(gz, opts, mode=2), want)


def test_barrier_that_gives_up_restarts_per_phase(gz, orc, device):
    # a spin limit of zero polls: a workgroup that finds the barrier closed
    # gives up at once, every workgroup leaves the kernel, and the run starts
    # again from the initial ranks with the per-phase launches
    g = gz.rmat_graph(12, 16, seed=3)
    gz.load(g)
    opts = gz.options("pr", iterations=20, damping=0.85)
    want = _per_phase(gz, opts)
    a0, n0 = gz.pr_persistent_gave_up(), gz.pr_persistent_runs()
    gz.pr_persistent(1)
    gz.pr_persistent_spin_limit(0)
    try:
        r = gz.run(opts)
    finally:
        gz.pr_persistent_spin_limit(1 << 22)
        gz.pr_persistent(0)
    assert gz.pr_persistent_gave_up() == a0 + 1
    assert gz.pr_persistent_runs() == n0
    assert r.iterations == 20
    assert np.array_equal(gz.pr_props().view(np.uint64), want)
    # and the next run takes the one-launch path again
    assert np.array_equal(_one_launch(gz, opts), want)
