"""HIP path versus the CPU oracle, through the C ABI.

Bars: PageRank (FP64) within 1e-12 maximum relative error per vertex (the
summation order differs from the oracle's sequential order); CC labels
bit-exact; BFS depths, parents and the per-iteration frontier statistic
bit-exact against the oracle's single-thread order (the GPU resolves every
race deterministically to the parent a single reference thread picks).
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

PR_TOL = 1e-12


def _orc_graph(orc, g):
    return orc.Graph(g.num_vertices, g.pull, g.push)


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.fixture(scope="module")
def rmat14(gz):
    return gz.rmat_graph(14, 16, seed=11)


@pytest.mark.parametrize("vpl", [1, 2, 4])
@pytest.mark.parametrize("chunk", [64, 512, 4096])
def test_pr_pull_sched_aware(gz, orc, device, rmat14, chunk, vpl):
    g = rmat14
    gz.set_chunk_vectors(chunk)
    gz.set_vectors_per_lane(vpl)
    gz.load(g)
    gz.run(gz.options("pr", iterations=20))
    ranks = gz.pr_ranks()
    gz.set_chunk_vectors(512)
    gz.set_vectors_per_lane(1)
    _, ref, _ = orc.pagerank(_orc_graph(orc, g), orc.options(iterations=20))
    assert _rel(ranks, ref) <= PR_TOL
    assert abs(ranks.sum() - 1.0) < 1e-9


def test_pr_deterministic(gz, device, rmat14):
    gz.load(rmat14)
    gz.run(gz.options("pr", iterations=5))
    a = gz.pr_props()
    gz.run(gz.options("pr", iterations=5))
    b = gz.pr_props()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("variant", ["traditional", "push"])
def test_pr_variants(gz, orc, device, rmat14, variant):
    # the reference's atomic comparison variants.  The traditional pull adds
    # chunk-boundary runs with FP64 atomics (scheduling order moves a sum by
    # rounding only): within 1e-12 of the oracle's traditional pull.
    g = rmat14
    G = _orc_graph(orc, g)
    gz.load(g)
    if variant == "traditional":
        gz.run(gz.options("pr", iterations=10, pull_variant="traditional"))
        _, ref, _ = orc.pagerank(G, orc.options(iterations=10, traditional=True))
        err = _rel(gz.pr_ranks(), ref)
        print(f"PageRank traditional pull: max relative error {err:.3e} against the oracle")
        assert err <= PR_TOL
        return
    # The push adds each edge's message as a 128-bit fixed-point integer, so
    # its accumulators are exact and the same in every order, and S is a
    # tree sum.  The reference's push (phase_edge_push_pr.asm:360-425) adds
    # every message to S in one sequential chain per thread, which moves S by
    # rounding over E terms and, through (1 - S) / V, every rank: on this
    # graph the oracle's own single-thread push differs from its pull by
    # about 1.9e-12, worst at vertices without in-edges.  So the HIP push is
    # held to 1e-12 against the oracle's pull (the same iteration with less
    # rounding), and to the oracle's own push-pull gap plus that bar against
    # the oracle's push.
    gz.run(gz.options("pr", iterations=10, engine="push"))
    got = gz.pr_ranks()
    _, pull, _ = orc.pagerank(G, orc.options(iterations=10))
    _, push, _ = orc.pagerank(G, orc.options(iterations=10, engine="push"))
    gap = _rel(push, pull)
    e_pull, e_push = _rel(got, pull), _rel(got, push)
    print(f"PageRank push: max relative error {e_pull:.3e} against the oracle's pull, {e_push:.3e} against its "
          f"push (the oracle's own push-pull gap {gap:.3e})")
    assert e_pull <= PR_TOL
    assert e_push <= gap + PR_TOL
    first = gz.pr_props().view(np.uint64).copy()
    gz.run(gz.options("pr", iterations=10, engine="push"))
    assert np.array_equal(gz.pr_props().view(np.uint64), first)


def test_pr_damping_085(gz, orc, device):
    g = gz.rmat_graph(12, 16, seed=4)
    gz.load(g)
    gz.run(gz.options("pr", iterations=20, damping=0.85))
    _, ref, _ = orc.pagerank(_orc_graph(orc, g), orc.options(iterations=20, damping=0.85))
    assert _rel(gz.pr_ranks(), ref) <= PR_TOL


@pytest.mark.parametrize("vpl", [1, 2, 4])
def test_pr_hub_spans_units(gz, orc, device, vpl):
    g = graphs.hub_graph(gz)
    gz.set_vectors_per_lane(vpl)
    _, ref, _ = orc.pagerank(_orc_graph(orc, g), orc.options(iterations=20))
    for chunk in (64, 128, 256):
        gz.set_chunk_vectors(chunk)
        gz.load(g)
        gz.run(gz.options("pr", iterations=20))
        got = gz.pr_ranks()
        gz.set_chunk_vectors(512)
        assert _rel(got, ref) <= PR_TOL
    gz.set_vectors_per_lane(1)


@pytest.mark.parametrize("vpl", [1, 2, 4])
def test_cc_pull_vectors_per_lane(gz, orc, device, rmat14, vpl):
    gz.set_vectors_per_lane(vpl)
    for chunk in (128, 512):
        gz.set_chunk_vectors(chunk)
        gz.load(rmat14)
        gz.run(gz.options("cc", engine="pull"))
        got = gz.cc_labels()
        gz.set_chunk_vectors(512)
        lab, _, _ = orc.cc(_orc_graph(orc, rmat14), orc.options(engine="pull"))
        assert np.array_equal(got, lab)
    gz.set_vectors_per_lane(1)


@pytest.mark.parametrize("case", ["rmat14", "planted_trees", "hub", "star_in", "path"])
@pytest.mark.parametrize("engine", ["hybrid", "pull"])
def test_cc_merge_buffer_pull(gz, orc, device, rmat14, case, engine):
    # EDGE_PULL_FORCE_MERGE: plain stores inside units, the unit-boundary
    # runs through the merge buffer with min.  Small units put the hub's and
    # the star centre's runs across many unit boundaries, so the merge sees
    # runs spanning several records followed by a directly stored head.
    if case == "rmat14":
        g = rmat14
    elif case == "planted_trees":
        g = graphs.planted_trees(gz)
    elif case == "hub":
        g = graphs.hub_graph(gz)
    else:
        g = dict(graphs.small_cases(gz))[case]
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options(engine=engine))
    try:
        for chunk, vpl in ((64, 1), (512, 2), (4096, 4)):
            gz.set_chunk_vectors(chunk)
            gz.set_vectors_per_lane(vpl)
            gz.load(g)
            gz.run(gz.options("cc", engine=engine, pull_variant="merge_buffer"))
            assert np.array_equal(gz.cc_labels(), lab), (case, chunk, vpl)
    finally:
        gz.set_chunk_vectors(512)
        gz.set_vectors_per_lane(1)


@pytest.mark.parametrize("name_idx", range(6))
def test_small_cases_all_algorithms(gz, orc, device, name_idx):
    name, g = graphs.small_cases(gz)[name_idx]
    G = _orc_graph(orc, g)
    gz.load(g)
    gz.run(gz.options("pr", iterations=20))
    _, ref, _ = orc.pagerank(G, orc.options(iterations=20))
    assert _rel(gz.pr_ranks(), ref) <= PR_TOL, name
    for engine in ("hybrid", "pull", "push"):
        gz.run(gz.options("cc", engine=engine))
        lab, _, _ = orc.cc(G, orc.options(engine=engine))
        assert np.array_equal(gz.cc_labels(), lab), (name, engine)
        gz.run(gz.options("bfs", engine=engine, root=0))
        par, dep, log, _ = orc.bfs(G, orc.options(engine=engine, root=0))
        p, d = gz.bfs_result()
        assert np.array_equal(d, dep), (name, engine)
        assert np.array_equal(p, par), (name, engine)


@pytest.mark.parametrize("engine", ["hybrid", "pull", "push"])
@pytest.mark.parametrize("skip", [True, False])
def test_cc_labels(gz, orc, device, engine, skip):
    g = graphs.planted_trees(gz)
    gz.set_frontier_skip(skip)
    gz.load(g)
    gz.run(gz.options("cc", engine=engine))
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options(engine=engine))
    gz.set_frontier_skip(True)
    assert np.array_equal(gz.cc_labels(), lab)
    # planted trees: every tree is its own component labelled by its smallest id
    base = 1 << 10
    labs = gz.cc_labels()
    for t in range(8):
        ids = np.arange(base + t * 100, base + (t + 1) * 100)
        assert np.all(labs[ids] == base + t * 100)


@pytest.mark.parametrize("shape", [(10, 8, 100), (14, 16, 300)])
@pytest.mark.parametrize("engine", ["hybrid", "push"])
@pytest.mark.parametrize("stat", ["both", "outdeg", "count"])
def test_cc_small_frontier_push(gz, orc, device, shape, engine, stat):
    # tree frontiers are small after the first sweeps, so with out-degrees in
    # the vote the push reads the bitmap directly; with counts only it
    # compacts.  V = 1,824 ends in a partial bitmap word and fits one
    # wavefront's words; V = 21,184 spans several wavefronts.
    scale, trees, size = shape
    g = graphs.planted_trees(gz, base_scale=scale, trees=trees, tree_size=size)
    gz.load(g)
    gz.run(gz.options("cc", engine=engine, stat=stat, collect_stats=True))
    recs = gz.records()
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options(engine=engine, stat=stat))
    assert np.array_equal(gz.cc_labels(), lab)
    small = [r for r in recs[1:] if r.engine == 2 and r.frontier_stat <= 65536]
    assert small, "no push iteration over a small frontier"
    if engine == "push" and scale == 14 and stat != "count":
        # both widths of the bitmap scan: 64 words per wavefront up to a
        # statistic of 4096, 16 words above it
        assert any(r.frontier_stat <= 4096 for r in small)
        assert any(r.frontier_stat > 4096 for r in small)
    for r in small:
        # on a symmetric graph every activated vertex has an out-edge; a push
        # examines each out-edge of its frontier once, and the vote that built
        # the frontier counted every frontier vertex's out-degree
        assert 0 < r.vectors <= r.edges <= 4 * r.vectors
        if stat != "count":
            assert r.edges <= r.frontier_stat


@pytest.mark.parametrize("L", [1200, 20000])
def test_cc_push_high_degree_source(gz, orc, device, L):
    # a 201-vertex path whose far end holds the smallest id, into the centre
    # of an L-leaf star: the smallest label walks the path one hop per push,
    # and each time it reaches the centre, the centre is a small frontier
    # with many vectors.  With 1,200 leaves one wavefront walks its 301
    # vectors in two batches; with 20,000 its 5,001 vectors are cut into
    # pieces for the overflow queue.
    P = 200
    pairs = [(i, i + 1) for i in range(P)] + [(P, P + 1 + j) for j in range(L)]
    g = gz.make_graph(np.asarray(pairs, dtype=np.uint64), P + 1 + L, symmetrize=True)
    gz.load(g)
    gz.run(gz.options("cc", engine="push", collect_stats=True))
    assert np.all(gz.cc_labels() == 0)
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options(engine="push"))
    assert np.array_equal(gz.cc_labels(), lab)
    recs = gz.records()
    lo, hi = (256, 4096) if L < 4096 else (4096, 65536)
    hub = [r for r in recs[1:] if lo < r.frontier_stat <= hi]
    assert hub, "no push over a small frontier holding the centre"
    for r in recs[1:]:
        assert r.vectors <= r.edges <= 4 * r.vectors
        assert r.edges <= r.frontier_stat


@pytest.mark.parametrize("engine", ["hybrid", "push"])
@pytest.mark.parametrize("L", [1200, 20000])
def test_bfs_push_high_degree_root(gz, orc, device, engine, L):
    # BFS from the centre of an L-leaf star with a 200-vertex path attached:
    # the root's level is one source with L + 1 out-edges, walked by one
    # wavefront (1,200 leaves) or cut into pieces for the overflow queue
    # (20,000 leaves); the path's levels are one small source each
    P = 200
    centre = P
    pairs = [(i, i + 1) for i in range(P)] + [(centre, P + 1 + j) for j in range(L)]
    g = gz.make_graph(np.asarray(pairs, dtype=np.uint64), P + 1 + L, symmetrize=True)
    G = _orc_graph(orc, g)
    gz.load(g)
    gz.run(gz.options("bfs", engine=engine, root=centre, profile=True))
    par, dep, log, _ = orc.bfs(G, orc.options(engine=engine, root=centre))
    p, d = gz.bfs_result()
    assert np.array_equal(d, dep)
    assert np.array_equal(p, par)
    assert [(x.engine, x.frontier_stat) for x in gz.records()] == log
    if engine == "push":
        assert gz.records()[0].engine == 2


@pytest.mark.parametrize("cap", [1, 3, 64])
def test_push_queue_full_fallback(gz, orc, device, cap):
    # a 20,000-leaf star behind a 200-vertex path: the centre's 5,001
    # vectors are cut into 20 pieces for the overflow queue; with room for
    # fewer pieces the lanes that cut the rest push them themselves, and
    # labels, depths and parents stay exact
    P, L = 200, 20000
    pairs = [(i, i + 1) for i in range(P)] + [(P, P + 1 + j) for j in range(L)]
    g = gz.make_graph(np.asarray(pairs, dtype=np.uint64), P + 1 + L, symmetrize=True)
    G = _orc_graph(orc, g)
    gz.load(g)
    gz.push_queue_capacity(cap)
    try:
        gz.run(gz.options("cc", engine="push", collect_stats=True))
        lab, _, _ = orc.cc(G, orc.options(engine="push"))
        assert np.array_equal(gz.cc_labels(), lab)
        assert any(4096 < r.frontier_stat <= 65536 for r in gz.records()[1:])
        for engine in ("push", "hybrid"):
            gz.run(gz.options("bfs", engine=engine, root=P, profile=True))
            par, dep, log, _ = orc.bfs(G, orc.options(engine=engine, root=P))
            p, d = gz.bfs_result()
            assert np.array_equal(d, dep) and np.array_equal(p, par), engine
            assert [(x.engine, x.frontier_stat) for x in gz.records()] == log
    finally:
        gz.push_queue_capacity(131072)
    with pytest.raises(ValueError):
        gz.push_queue_capacity(0)
    with pytest.raises(ValueError):
        gz.push_queue_capacity(131073)


@pytest.mark.parametrize("narrow", [True, False])
def test_compaction_counts_and_offsets(gz, orc, device, rmat14, narrow):
    # the frontier compaction over 32-bit vector counts and over the 64-bit
    # offsets (the path of lists of 2^32 vectors or more): the same results,
    # with the push and the bottom-up levels compacting (count-only votes
    # keep small CC pushes on the compaction too)
    g = rmat14
    G = _orc_graph(orc, g)
    gz.load(g)
    gz.compact_narrow(narrow)
    try:
        engines = set()
        for root in (0, 777):
            gz.run(gz.options("bfs", root=root, profile=True))
            par, dep, log, _ = orc.bfs(G, orc.options(root=root))
            p, d = gz.bfs_result()
            assert np.array_equal(d, dep) and np.array_equal(p, par)
            assert [(x.engine, x.frontier_stat) for x in gz.records()] == log
            engines |= {x.engine for x in gz.records()}
        assert engines == {1, 2}
        gz.run(gz.options("cc", stat="count"))
        lab, _, _ = orc.cc(G, orc.options(stat="count"))
        assert np.array_equal(gz.cc_labels(), lab)
    finally:
        gz.compact_narrow(True)


def test_compaction_many_segments(gz, orc, device):
    # 2^25 + 12,345 vertices (8,195 blocks of 4,096, so 129 blocks per
    # segment of the scan-free count and write passes) holding a star of
    # 150,000 leaves spread over every segment, and a 10-vertex path behind
    # one leaf: the root's level and the leaves' level compact (statistics
    # above the small-frontier bound), the hybrid pulls over the unvisited
    # list, and every depth, parent and statistic must equal the oracle's
    V = (1 << 25) + 12345
    rng = np.random.default_rng(17)
    ids = rng.choice(V, size=150_000 + 11, replace=False).astype(np.uint64)
    root, leaves, path = ids[0], ids[1:150_001], ids[150_001:]
    pairs = [np.stack([np.full(leaves.size, root, dtype=np.uint64), leaves], axis=1),
             np.asarray([(leaves[7], path[0])] + [(path[i], path[i + 1]) for i in range(path.size - 1)],
                        dtype=np.uint64)]
    g = gz.make_graph(np.concatenate(pairs), V, symmetrize=True)
    G = _orc_graph(orc, g)
    gz.load(g)
    for engine in ("push", "hybrid"):
        gz.run(gz.options("bfs", engine=engine, root=int(root), profile=True))
        par, dep, log, _ = orc.bfs(G, orc.options(engine=engine, root=int(root)))
        p, d = gz.bfs_result()
        assert np.array_equal(d, dep), engine
        assert np.array_equal(p, par), engine
        assert [(x.engine, x.frontier_stat) for x in gz.records()] == log, engine
    gz.run(gz.options("cc", engine="push"))
    lab, _, _ = orc.cc(G, orc.options(engine="push"))
    assert np.array_equal(gz.cc_labels(), lab)


def test_cc_directed_rmat(gz, orc, device, rmat14):
    gz.load(rmat14)
    gz.run(gz.options("cc"))
    lab, _, _ = orc.cc(_orc_graph(orc, rmat14), orc.options())
    assert np.array_equal(gz.cc_labels(), lab)


@pytest.mark.parametrize("engine", ["hybrid", "pull", "push"])
@pytest.mark.parametrize("stat", ["both", "outdeg", "count"])
def test_bfs_bit_exact(gz, orc, device, rmat14, engine, stat):
    g = rmat14
    G = _orc_graph(orc, g)
    gz.load(g)
    for root in (0, 5, 777):
        gz.run(gz.options("bfs", engine=engine, stat=stat, root=root, profile=True))
        par, dep, log, r = orc.bfs(G, orc.options(engine=engine, stat=stat, root=root))
        p, d = gz.bfs_result()
        assert np.array_equal(d, dep)
        assert np.array_equal(p, par)
        recs = gz.records()
        assert [(x.engine, x.frontier_stat) for x in recs] == log


@pytest.mark.parametrize("skip", [True, False])
def test_bfs_frontier_skip_modes(gz, orc, device, skip):
    g = graphs.hub_graph(gz)
    gz.set_frontier_skip(skip)
    gz.load(g)
    gz.run(gz.options("bfs", root=2))
    par, dep, _, _ = orc.bfs(_orc_graph(orc, g), orc.options(root=2))
    gz.set_frontier_skip(True)
    p, d = gz.bfs_result()
    assert np.array_equal(d, dep) and np.array_equal(p, par)


def test_device_layout_matches_host_packer(gz, orc, device, rmat14):
    g = rmat14
    gz.load(g)
    for pull, pairs in ((True, g.pull), (False, g.push)):
        dv = gz.device_vectors(pull)
        assert np.array_equal(dv, orc.pack(pairs, by_dst=pull))
        off = gz.device_offsets(pull)
        assert off[-1] == dv.shape[0]
        assert np.all(np.diff(off.astype(np.int64)) >= 0)


def test_cli_output_matches_oracle(gz, orc, device, tmp_path, rmat14):
    import subprocess, os
    g = gz.rmat_graph(10, 8, seed=9)
    base = str(tmp_path / "g")
    gz.write_graph_files(base, g)
    exe = os.path.join(os.path.dirname(gz.LIB_PATH), "grazelle")
    out = str(tmp_path / "o.txt")
    res = subprocess.run([exe, "-i", base, "-a", "cc", "-o", out], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options())
    want = "".join(f"{i} {float(x):.0f}\n" for i, x in enumerate(lab))
    assert open(out).read() == want
    res = subprocess.run([exe, "-i", base, "-N", "20", "-o", out], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "PageRank Sum" in res.stdout
    _, ranks, _ = orc.pagerank(_orc_graph(orc, g), orc.options(iterations=20))
    got = [float(l.split()[1]) for l in open(out)]
    assert np.allclose(got, ranks, rtol=1e-5, atol=0)


@pytest.mark.parametrize("spec", ["5", "4,2", "3,3"])
def test_cli_source_sliced_matches_oracle(gz, orc, device, tmp_path, spec):
    # -L: device ingress with renumbering and the source-sliced pull; output
    # in original ids (labels exact, ranks to the output's 6 digits)
    import subprocess, os
    g = gz.rmat_graph(11, 8, seed=19)
    base = str(tmp_path / "g")
    gz.write_graph_files(base, g)
    exe = os.path.join(os.path.dirname(gz.LIB_PATH), "grazelle")
    out = str(tmp_path / "o.txt")
    res = subprocess.run([exe, "-i", base, "-L", spec, "-a", "cc", "-o", out], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "Source-sliced pull" in res.stdout
    lab, _, _ = orc.cc(_orc_graph(orc, g), orc.options())
    assert open(out).read() == "".join(f"{i} {float(x):.0f}\n" for i, x in enumerate(lab))
    res = subprocess.run([exe, "-i", base, "-L", spec, "-N", "20", "-o", out], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    _, ranks, _ = orc.pagerank(_orc_graph(orc, g), orc.options(iterations=20))
    got = [float(l.split()[1]) for l in open(out)]
    assert np.allclose(got, ranks, rtol=1e-5, atol=0)


@pytest.mark.parametrize("variant", ["sched_aware", "traditional", "source_sliced", "push"])
def test_pr_graph_replay_matches_eager(gz, device, variant):
    # iterations 2.. replayed from a captured graph: bit-identical to eager
    # launches (within 1e-12 for the atomic variants, whose own eager runs
    # differ by rounding), over several replays (cache hits, another
    # iteration count) with eager work and synchronisation in between
    g = gz.rmat_graph(13, 16, seed=41)
    gz.load(g)
    kw = {"engine": "push"} if variant == "push" else {"pull_variant": variant}
    if variant == "source_sliced":
        gz.build_source_slices(128, 2, 3, 32)
        gz.slices_set_hot_cache(4096)
    try:
        eager = {}
        gz.pr_graph_replay(False)
        for iters in (20, 7):
            gz.run(gz.options("pr", iterations=iters, **kw))
            eager[iters] = gz.pr_props().view(np.uint64).copy()
        gz.pr_graph_replay(True)
        gz.run(gz.options("cc"))  # the allocations the CC runs below make
        c0, l0 = gz.graph_replay_counts()
        # a configuration runs eagerly, then is captured when it repeats:
        # 20 eager, 20 captured, 7 eager, 7 captured, 20 eager, 7 replays,
        # 20 captured, 20 replays
        for iters in (20, 20, 7, 7, 20, 7, 20, 20):
            gz.run(gz.options("pr", iterations=iters, **kw))
            got = gz.pr_props()
            if variant in ("push", "traditional"):
                want = eager[iters].view(np.float64)
                assert np.max(np.abs(got - want) / np.abs(want)) <= 1e-12, iters
            else:
                assert np.array_equal(got.view(np.uint64), eager[iters]), iters
            gz.run(gz.options("cc"))
            gz.lib.gz_synchronize()
        c1, l1 = gz.graph_replay_counts()
        assert (c1 - c0, l1 - l0) == (3, 5)
    finally:
        gz.pr_graph_replay(True)
        gz.slices_set_hot_cache(0)
