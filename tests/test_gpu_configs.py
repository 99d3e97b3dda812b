"""The HIP path on each benchmark configuration's own workload, against the
oracle (reference anchor: the drivers of execution_pr.c:92-168,
execution_cc.c:94-268 and execution_bfs.c:120-276).

Each large graph is built exactly as bench.py builds it (device ingress, with
the same renumbering and source-sliced layout), its Vector-Sparse lists are
copied back, and the oracle runs over those lists (device ingress is pinned
bit-identical to the host packer in test_gpu_ingress.py).  Renumbered graphs
are compared through the device-to-original id map: PageRank ranks are
permutation-equivariant, BFS depths are numbering-free, and CC on the oracle
starts from the original ids as the product does.

  C1  R-MAT s16 ef16, PageRank plain pull, 20 iterations, d = 0.85
  C2  symmetrised R-MAT s22 + 64 planted 1000-vertex trees, hybrid CC, plain
      pull and the source-sliced pull (2^19 windows) on the renumbered graph
  C3  R-MAT s24, BFS from the 64 seeded roots, fixed threshold and learned
  C4  the 1024-graph batch: PageRank (10 iterations) and CC on the union
  C5  the full-scale graph (41.7M vertices, 1.47B edges) in the headline
      layout: PageRank, CC and BFS
and the headline kernel instance, the sliced pull with a 20480-source LDS
cache and five LDS blocks, against the oracle on R-MAT s21.

Bars: PageRank within 1e-12 maximum relative error per vertex, and ranks
summing to 1 within SUM_TOL (rounding of the sum itself over up to 41.7M
terms); CC labels bit-exact; BFS depths bit-exact, parents bit-exact where the
numbering is the oracle's (fixed threshold, not renumbered), a valid BFS tree
otherwise.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PR_TOL = 1e-12
SUM_TOL = 1e-10
UNREACHED = np.iinfo(np.uint64).max
ORACLE_THREADS = max(1, min(16, os.cpu_count() or 1))

# the headline layout (bench.py): two tiers of 2^20-vertex windows, five LDS
# blocks of 20480 sources, a 20480-source LDS cache, chunks of 2048 vectors
HEAD_SLICES, HEAD_TIERS, HEAD_BLOCKS, HOT, SLICED_CHUNK = 20, 2, 5, 20480, 2048


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _download_oracle_graph(gz, orc):
    """Oracle graph over the device lists (device numbering)."""
    return orc.Graph.from_vectors(gz.num_vertices(), gz.device_vectors(True), gz.device_vectors(False))


def _sliced(gz, run):
    gz.set_chunk_vectors(SLICED_CHUNK)
    gz.slices_set_hot_cache(HOT)
    try:
        return run()
    finally:
        gz.set_chunk_vectors(512)
        gz.slices_set_hot_cache(0)


def _to_device_order(values_by_original, orig):
    return values_by_original[orig.astype(np.int64)]


def _parents_to_device(parents_by_original, orig):
    dev_of = np.empty(orig.shape[0], np.int64)
    dev_of[orig.astype(np.int64)] = np.arange(orig.shape[0])
    p = _to_device_order(parents_by_original, orig)
    return np.where(p >= 0, dev_of[np.maximum(p, 0)], -1)


# ---- C1 ---------------------------------------------------------------------------

@pytest.mark.parametrize("ingress", ["host", "device"])
def test_c1_pr_rmat16(gz, orc, device, ingress):
    g = gz.workload_graph("pr-rmat16")
    assert g.num_vertices == 65536 and 900_000 < g.num_edges < 1_000_000
    if ingress == "host":
        gz.load(g)
    else:
        gz.load_workload_device("pr-rmat16")
        assert np.array_equal(gz.device_vectors(True), orc.pack(g.pull, True))
    _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=20, damping=0.85))
    # the default chunks and the benchmark's (256 vectors); both take the
    # fused merge-sum-vertex launch
    try:
        for chunk in (512, 256):
            gz.set_chunk_vectors(chunk)
            if ingress == "host":
                gz.load(g)
            gz.run(gz.options("pr", iterations=20, damping=0.85))
            assert gz.lib.gz_pr_fused_tail_applies()
            ranks = gz.pr_ranks()
            assert _rel(ranks, ref) <= PR_TOL, chunk
            assert abs(ranks.sum() - 1.0) < SUM_TOL
    finally:
        gz.set_chunk_vectors(512)


# ---- C2 ---------------------------------------------------------------------------

C2_BASE, C2_TREES, C2_TREE = 1 << 22, 64, 1000


@pytest.fixture(scope="module")
def c2(gz, orc, device):
    gz.load_workload_device("cc-rmat22-sym-trees")
    V = gz.num_vertices()
    assert V == C2_BASE + C2_TREES * C2_TREE
    G = _download_oracle_graph(gz, orc)
    labels, _, _ = orc.cc(G, orc.options(threads=ORACLE_THREADS))
    return V, labels


def _check_trees(labels):
    for t in range(C2_TREES):
        lo = C2_BASE + t * C2_TREE
        assert np.all(labels[lo:lo + C2_TREE] == lo), t


def test_c2_cc_plain_pull(gz, orc, c2):
    V, want = c2
    _check_trees(want)
    gz.load_workload_device("cc-rmat22-sym-trees")
    for engine in ("hybrid", "pull", "push"):
        r = gz.run(gz.options("cc", engine=engine))
        assert np.array_equal(gz.cc_labels(), want), engine
        assert r.iterations > 0


def test_c2_cc_source_sliced_renumbered(gz, orc, c2):
    V, want = c2
    gz.load_workload_device("cc-rmat22-sym-trees", relabel=True, slices_log2=19, slice_tiers=1)
    assert gz.source_slices_layout()[0] == 1 << 19
    r = _sliced(gz, lambda: gz.run(gz.options("cc", pull_variant="source_sliced")))
    got = gz.cc_labels()
    assert r.pull_iterations >= 1
    assert np.array_equal(got, want)
    _check_trees(got)


# ---- C3 ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def c3(gz, orc, device):
    import selector_train as st

    gz.load_workload_device("bfs-rmat24")
    assert gz.num_vertices() == 1 << 24
    G = _download_oracle_graph(gz, orc)
    roots = st.pick_roots(64, 24)
    return G, roots


def test_c3_bfs_fixed_threshold_against_oracle(gz, orc, c3):
    G, roots = c3
    gz.load_workload_device("bfs-rmat24")
    for root in roots[:4]:
        gz.run(gz.options("bfs", root=root, profile=True))
        p, d = gz.bfs_result()
        par, dep, log, _ = orc.bfs(G, orc.options(root=root))
        assert np.array_equal(d, dep), root
        assert np.array_equal(p, par), root
        assert [(x.engine, x.frontier_stat) for x in gz.records()] == log, root
        assert {e for e, _ in log} == {1, 2}, "the hybrid used one engine only"


def test_c3_bfs_learned_equals_fixed_all_roots(gz, orc, c3):
    import selector_train as st

    G, roots = c3
    gz.load_workload_device("bfs-rmat24")
    fixed = {}
    for root in roots:
        gz.run(gz.options("bfs", root=root))
        fixed[root] = gz.bfs_result()[1]
    gz.selector_load(st.DEFAULT_WEIGHTS)
    try:
        engines = set()
        for i, root in enumerate(roots):
            gz.run(gz.options("bfs", engine="learned", root=root, profile=True))
            p, d = gz.bfs_result()
            engines |= {x.engine for x in gz.records()}
            assert np.array_equal(d, fixed[root]), root
            if i < 16:
                assert orc.bfs_certificate_violations(G, root, p, d, ORACLE_THREADS) == 0, root
        assert engines == {1, 2}
    finally:
        gz.selector_clear()


# ---- C4 ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def c4(gz, orc, device):
    specs = gz.batch_specs()
    assert len(specs) == 1024
    checked = list(range(0, 1024, 16))
    singles = {i: gz.rmat_graph(*specs[i]) for i in checked}
    refs = {}
    for i, g in singles.items():
        G = orc.Graph(g.num_vertices, g.pull, g.push)
        refs[i] = (orc.pagerank(G, orc.options(iterations=10, damping=0.85))[1], orc.cc(G, orc.options())[0])
    return specs, singles, refs


@pytest.mark.parametrize("cfg", [
    dict(vectors_per_lane=1, workgroup_size=256, chunk=(128, 256, 512, 1024)),
    dict(vectors_per_lane=2, workgroup_size=512, chunk=(128, 512, 1024, 2048)),
])
def test_c4_batch_1024(gz, orc, c4, cfg):
    specs, singles, refs = c4
    bases = gz.load_batch_device(specs)
    E = gz.num_edges()
    assert E > 400_000_000
    for i, g in singles.items():
        lo, hi = int(bases[i]), int(bases[i + 1])
        assert hi - lo == g.num_vertices
    gz.batch_set_config(**cfg)
    b = bases.astype(np.int64)
    for use_graph in (False, True, True):  # eager, captured, replayed
        gz.batch_pr(10, damping=0.85, use_graph=use_graph)
        ranks = gz.batch_pr_ranks()
        sums = np.add.reduceat(ranks, b[:-1])
        assert np.max(np.abs(sums - 1.0)) < SUM_TOL
        for i in singles:
            assert _rel(ranks[b[i]:b[i + 1]], refs[i][0]) <= PR_TOL, (i, use_graph)
    gz.run(gz.options("cc"))
    labels = gz.cc_labels().astype(np.int64)
    owner = np.repeat(np.arange(len(specs)), np.diff(b))
    assert np.all(labels >= b[owner]) and np.all(labels < b[owner + 1])
    for i in singles:
        assert np.array_equal(labels[b[i]:b[i + 1]] - b[i], refs[i][1].astype(np.int64)), i


# ---- the headline kernel instance -----------------------------------------------

def test_headline_kernel_instance_against_oracle(gz, orc, device):
    # gz_k_edge_pull_pr_sliced_hot<20480> with its five LDS blocks: windows
    # of 2^17 hold the blocks ((5 + 1) * 20480 <= 2^17), two tiers fill the
    # 2^21 vertices of R-MAT s21
    scale, seed = 21, 5
    g = gz.rmat_graph(scale, 16, seed)
    gz.load_rmat_device(scale, (1 << scale) * 16, seed, relabel=True, slices_log2=17, slice_tiers=2, slice_blocks=5)
    assert gz.source_slices_layout() == (1 << 17, 2, 5, HOT)
    vecs, _ = gz.source_slices_info()
    assert vecs.shape[0] == 8 * (2 + 5) + 1 and np.all(vecs[8:48] > 0)  # every block slice holds vectors
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    _, ref, _ = orc.pagerank(G, orc.options(iterations=20, damping=0.85))
    for vpl in (1, 2):
        gz.set_vectors_per_lane(vpl)
        try:
            _sliced(gz, lambda: gz.run(gz.options("pr", iterations=20, damping=0.85, pull_variant="source_sliced")))
        finally:
            gz.set_vectors_per_lane(1)
        ranks = gz.pr_ranks()
        assert _rel(ranks, ref) <= PR_TOL, vpl
        assert abs(ranks.sum() - 1.0) < SUM_TOL
    # CC over the same layout
    sym = gz.make_graph(gz.rmat_edges(scale, (1 << scale) * 8, seed), 1 << scale, symmetrize=True)
    gz.load_pairs_device(sym.num_vertices, sym.pull, relabel=True, slices_log2=17, slice_tiers=2, slice_blocks=5)
    _sliced(gz, lambda: gz.run(gz.options("cc", pull_variant="source_sliced")))
    want, _, _ = orc.cc(orc.Graph(sym.num_vertices, sym.pull, sym.push), orc.options(threads=ORACLE_THREADS))
    assert np.array_equal(gz.cc_labels(), want)


# ---- C5 ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def c5(gz, orc, device):
    gz.load_workload_device("full-rmat26-fold41.7m", relabel=True, slices_log2=HEAD_SLICES, slice_tiers=HEAD_TIERS,
                            slice_blocks=HEAD_BLOCKS)
    V, E = gz.num_vertices(), gz.num_edges()
    assert V == gz.FULL_VERTICES and 1_450_000_000 < E < 1_490_000_000
    assert gz.source_slices_layout() == (1 << HEAD_SLICES, HEAD_TIERS, HEAD_BLOCKS, HOT)
    orig = gz.original_ids()
    G = _download_oracle_graph(gz, orc)
    return G, orig


def test_c5_pagerank_headline_layout(gz, orc, c5):
    G, orig = c5
    opts = dict(iterations=20, damping=0.85)
    _sliced(gz, lambda: gz.run(gz.options("pr", pull_variant="source_sliced", **opts)))
    sliced = gz.pr_ranks()
    gz.run(gz.options("pr", **opts))
    plain = gz.pr_ranks()
    _, ref, _ = orc.pagerank(G, orc.options(threads=ORACLE_THREADS, **opts))
    assert _rel(_to_device_order(sliced, orig), ref) <= PR_TOL
    assert _rel(_to_device_order(plain, orig), ref) <= PR_TOL
    assert abs(sliced.sum() - 1.0) < SUM_TOL
    # the in-degree tail of the workload (about 770k at the top)
    indeg = gz.indegrees()
    assert int(indeg.sum()) == gz.num_edges()
    print(f"C5 max in-degree {int(indeg.max())}")
    assert 400_000 < int(indeg.max()) < 1_600_000


def test_c5_cc_headline_layout(gz, orc, c5):
    G, orig = c5
    want, _, _ = orc.cc(G, orc.options(threads=ORACLE_THREADS), init=orig)
    gz.run(gz.options("cc"))
    assert np.array_equal(_to_device_order(gz.cc_labels(), orig), want)
    _sliced(gz, lambda: gz.run(gz.options("cc", pull_variant="source_sliced")))
    assert np.array_equal(_to_device_order(gz.cc_labels(), orig), want)


def test_c5_bfs_16_roots(gz, orc, c5):
    import selector_train as st

    G, orig = c5
    roots = st.pick_roots(16, 26)
    fixed = {}
    for root in roots:
        gz.run(gz.options("bfs", root=root))
        fixed[root] = gz.bfs_result()
    gz.selector_load(st.DEFAULT_WEIGHTS)
    try:
        for root in roots:
            gz.run(gz.options("bfs", engine="learned", root=root))
            assert np.array_equal(gz.bfs_result()[1], fixed[root][1]), root
    finally:
        gz.selector_clear()
    for root in roots[:2]:
        droot = gz.device_vertex(root)
        _, dep, _, _ = orc.bfs(G, orc.options(threads=ORACLE_THREADS, root=droot))
        p, d = fixed[root]
        d_dev = _to_device_order(d, orig)
        assert np.array_equal(d_dev, dep), root
        assert orc.bfs_certificate_violations(G, droot, _parents_to_device(p, orig), d_dev, ORACLE_THREADS) == 0, root
