"""CC push iterations over small frontiers run back to back in one workgroup
(gz_cc_push_tail) against one launch per iteration (execution_cc.c:94-268:
the host reads the vote after every iteration).

Bars: labels bit-exact against the oracle and against the per-iteration
launches; the workgroup hands the host loop a state it continues from (a
statistic above the cap, an iteration limit); the iteration counts add up.
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

DEFAULT_TAIL = 4096  # the library's default entry bound


def _orc_graph(orc, g):
    return orc.Graph(g.num_vertices, g.pull, g.push)


def _labels(gz, opts, max_vote):
    gz.cc_push_tail(max_vote)
    try:
        r = gz.run(opts)
        return gz.cc_labels().copy(), r
    finally:
        gz.cc_push_tail(DEFAULT_TAIL)


def _path(gz, n, shuffle=False):
    # ids increase along the path, so the smallest label crosses all of it;
    # shuffled, the smallest id sits somewhere inside and the vertex order of
    # a round's list is unrelated to the path's
    ids = np.arange(n, dtype=np.uint64)
    if shuffle:
        ids = np.random.default_rng(7).permutation(ids)
    pairs = np.stack([ids[:-1], ids[1:]], axis=1)
    return gz.make_graph(pairs, n, symmetrize=True)


def _star_behind_path(gz, P=200, L=20000):
    pairs = [(i, i + 1) for i in range(P)] + [(P, P + 1 + j) for j in range(L)]
    return gz.make_graph(np.asarray(pairs, dtype=np.uint64), P + 1 + L, symmetrize=True)


def _cases(gz):
    small = dict(graphs.small_cases(gz))
    return {
        "planted_trees": lambda: graphs.planted_trees(gz),
        "planted_trees_s14": lambda: graphs.planted_trees(gz, base_scale=14, edge_factor=16, trees=16, tree_size=300),
        "rmat13": lambda: gz.rmat_graph(13, 16, seed=41),
        "hub": lambda: graphs.hub_graph(gz),
        "path3000": lambda: _path(gz, 3000),
        "path3000_shuffled": lambda: _path(gz, 3000, shuffle=True),
        "star_behind_path": lambda: _star_behind_path(gz),
        "path": lambda: small["path"],
        "star_in": lambda: small["star_in"],
        "star_out": lambda: small["star_out"],
        "isolated_tail": lambda: small["isolated_tail"],
    }


@pytest.mark.parametrize("engine", ["hybrid", "push"])
@pytest.mark.parametrize("case", ["planted_trees", "planted_trees_s14", "rmat13", "hub", "path3000",
                                  "path3000_shuffled", "star_behind_path", "path", "star_in", "star_out",
                                  "isolated_tail"])
def test_tail_equals_oracle_and_per_iteration(gz, orc, device, case, engine):
    g = _cases(gz)[case]()
    gz.load(g)
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options(engine=engine))
    opts = gz.options("cc", engine=engine)
    want, r0 = _labels(gz, opts, 0)
    assert np.array_equal(want, lab), case
    for cap in (16384, 64, 3):
        n0 = gz.cc_push_tail_rounds()
        got, r = _labels(gz, opts, cap)
        assert np.array_equal(got, lab), (case, cap)
        assert r.pull_iterations + r.push_iterations == r.iterations > 0, (case, cap)
        if cap == 16384 and case.startswith(("path3000", "planted")):
            assert gz.cc_push_tail_rounds() > n0, (case, "the tail did not run")
    assert r0.pull_iterations + r0.push_iterations == r0.iterations > 0


def test_tail_runs_many_rounds(gz, orc, device):
    # the smallest label walks 2,999 hops, one per round and at most one
    # more per 1,024-vertex piece of a round's list that it crosses (a
    # piece reads its sources' labels after the pieces before it): the
    # whole run is in one workgroup (statistic 8,998 from the start)
    g = _path(gz, 3000)
    gz.load(g)
    n0 = gz.cc_push_tail_rounds()
    got, r = _labels(gz, gz.options("cc", engine="push"), 16384)
    assert np.all(got == 0)
    ran = gz.cc_push_tail_rounds() - n0
    assert ran == r.iterations and ran >= 750, (ran, r.iterations)
    assert r.push_iterations == r.iterations
    # one launch per iteration, the same labels
    want, r0 = _labels(gz, gz.options("cc", engine="push"), 0)
    assert np.array_equal(got, want)
    assert r0.iterations >= 2


def test_tail_hands_back_above_the_cap(gz, orc, device):
    # pushing: the path's label reaches the star's centre, whose 20,001
    # out-edges put the statistic above the cap: the workgroup stops there,
    # the host loop pushes the centre and its leaves, and the tail resumes
    # afterwards.  (The hybrid engine pulls the label along the whole path
    # in its first sweeps; what is left, the centre or its leaves, is above
    # its threshold, so its run has no tail.)
    g = _star_behind_path(gz)
    G = _orc_graph(orc, g)
    gz.load(g)
    for engine in ("push", "hybrid"):
        n0 = gz.cc_push_tail_rounds()
        got, r = _labels(gz, gz.options("cc", engine=engine), 16384)
        lab, _, _ = orc.cc(G, orc.options(engine=engine))
        assert np.array_equal(got, lab), engine
        assert np.all(got == 0)
        if engine == "push":
            assert gz.cc_push_tail_rounds() > n0
            assert r.push_iterations == r.iterations > gz.cc_push_tail_rounds() - n0
        assert r.pull_iterations + r.push_iterations == r.iterations


@pytest.mark.parametrize("limit", [1, 2, 5, 40, 700])
def test_tail_respects_max_iterations(gz, device, limit):
    # an iteration limit reached inside the tail ends the run there
    g = _path(gz, 3000)
    gz.load(g)
    got, r = _labels(gz, gz.options("cc", engine="push", max_iterations=limit), 16384)
    assert r.iterations == limit == r.push_iterations
    ids = np.arange(3000, dtype=np.uint64)
    assert np.all(got <= ids)
    final, _ = _labels(gz, gz.options("cc", engine="push"), 16384)
    assert np.all(got >= final)
    if limit <= 40:
        # the smallest label advances at most four hops a round, so a short
        # run cannot have finished the path
        assert not np.all(got == 0)


@pytest.mark.parametrize("stat", ["both", "outdeg", "count"])
def test_tail_stat_modes(gz, orc, device, stat):
    # with out-degrees in the statistic the tail applies; counts alone do
    # not bound a frontier's edges, so those runs stay on the host loop
    g = graphs.planted_trees(gz, base_scale=12, edge_factor=8, trees=16, tree_size=200)
    gz.load(g)
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options(engine="hybrid", stat=stat))
    n0 = gz.cc_push_tail_rounds()
    got, _ = _labels(gz, gz.options("cc", engine="hybrid", stat=stat), 16384)
    assert np.array_equal(got, lab), stat
    ran = gz.cc_push_tail_rounds() - n0
    if stat == "count":
        assert ran == 0
    else:
        assert ran > 0


def test_tail_with_kernel_timing_and_records(gz, orc, device):
    # kernel timing keeps the tail (its launch is timed as the CC push);
    # per-iteration records and statistics turn it off
    g = graphs.planted_trees(gz)
    gz.load(g)
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options(engine="hybrid"))
    gz.kernel_timing(True)
    try:
        n0 = gz.cc_push_tail_rounds()
        gz.run(gz.options("cc"))
        kt = gz.kernel_times()
    finally:
        gz.kernel_timing(False)
    assert np.array_equal(gz.cc_labels(), lab)
    assert gz.cc_push_tail_rounds() > n0
    assert kt["push_cc"][1] > 0
    n0 = gz.cc_push_tail_rounds()
    gz.run(gz.options("cc", profile=True, collect_stats=True))
    assert gz.cc_push_tail_rounds() == n0
    assert np.array_equal(gz.cc_labels(), lab)
