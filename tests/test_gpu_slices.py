"""Source-sliced PageRank pull versus the host packer and the CPU oracle.

Layout (T tiers, G LDS blocks of H sources): slice 8*level + x must be the
host packer's pull list of the edges whose source falls in that slice (level
0: tier-0 window x without its blocks, levels 1..G: block b of window x,
[x*S + b*H, x*S + (b+1)*H), level G + t: tier-t window x), the last slice of
all other edges, with each destination replaced by its rank among the slice's
destinations, concatenated in slice order; the mask must name the slices
holding each destination.
PageRank over it must match the oracle within 1e-12 maximum relative error
per vertex (each destination's sum is regrouped by slice), be bit-identical
run to run, and the interleaved renumbering must deal each tier's 8*S ranks
of the degree order round robin into the tier's slices.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PR_TOL = 1e-12


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.fixture(scope="module")
def rmat14(gz):
    return gz.rmat_graph(14, 16, seed=11)


def _layout_params(V, S, T, G, H):
    S = min(S, V // (8 * T), 1 << 23)
    if G and (H == 0 or S // H < 1):
        G = 0
    G = min(G, S // H - 1) if G else 0
    return S, G


def _expected_layout(gz, pull, V, S, T=1, G=0, H=0):
    S, G = _layout_params(V, S, T, G, H)
    src = pull[:, 0].astype(np.int64)
    cold = 8 * (G + T)
    if S:
        w = src // S
        local = src - w * S
        tier, x = w // 8, w % 8
        level = np.where(tier > 0, G + tier, 0)
        if G:
            level = np.where((tier == 0) & (local >= H) & (local < (G + 1) * H), local // H, level)
        sl = np.where(src < 8 * T * S, 8 * level + x, cold)
    else:
        sl = np.full(len(src), cold)
    vecs, mask = [], np.zeros(V, np.uint64)
    for x in range(cold + 1):
        part = pull[sl == x]
        if len(part):
            dsts, local = np.unique(part[:, 1], return_inverse=True)
            ranked = np.stack([part[:, 0], local.astype(np.uint64)], axis=1)
            vecs.append(gz.pack_vectors(ranked, True))
            mask[dsts.astype(np.int64)] |= np.uint64(1 << x)
        else:
            vecs.append(np.zeros((0, 4), np.uint64))
    return vecs, mask


@pytest.mark.parametrize("S,T,G,H", [(1, 1, 0, 0), (64, 1, 0, 0), (700, 1, 0, 0), (1 << 20, 1, 0, 0),
                                     (64, 2, 0, 0), (300, 3, 0, 0), (1 << 20, 3, 0, 0),
                                     (1024, 1, 3, 64), (2048, 1, 6, 100), (512, 2, 2, 100), (300, 3, 4, 10),
                                     (128, 1, 5, 64), (100, 1, 2, 200)])
def test_slices_layout_matches_host_packer(gz, device, rmat14, S, T, G, H):
    g = rmat14
    gz.load(g)
    gz.build_source_slices(S, T, G, H)
    s_eff, g_eff = _layout_params(g.num_vertices, S, T, G, H)
    assert gz.source_slices_layout() == (s_eff, T, g_eff, H if G else 0)
    counts, got_s = gz.source_slices_info()
    assert len(counts) == 8 * (T + g_eff) + 1
    assert got_s == s_eff
    vecs, mask = gz.source_slices_download()
    want, want_mask = _expected_layout(gz, g.pull, g.num_vertices, S, T, G, H)
    assert [int(c) for c in counts] == [len(w) for w in want]
    assert np.array_equal(vecs, np.concatenate(want).reshape(-1, 4))
    assert np.array_equal(mask, want_mask)


@pytest.mark.parametrize("spread", [True, False])
@pytest.mark.parametrize("vpl", [1, 2, 4])
@pytest.mark.parametrize("chunk", [64, 512])
@pytest.mark.parametrize("S,T", [(64, 1), (1024, 1), (128, 2), (256, 3)])
def test_sliced_pr_parity(gz, orc, device, rmat14, S, T, chunk, vpl, spread):
    g = rmat14
    gz.set_chunk_vectors(chunk)
    gz.set_vectors_per_lane(vpl)
    gz.slices_set_cold_spread(spread)
    try:
        gz.load(g)
        gz.build_source_slices(S, T)
        gz.run(gz.options("pr", iterations=20, pull_variant="source_sliced"))
        ranks = gz.pr_ranks()
    finally:
        gz.set_chunk_vectors(512)
        gz.set_vectors_per_lane(1)
        gz.slices_set_cold_spread(True)
    _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=20))
    assert _rel(ranks, ref) <= PR_TOL
    assert abs(ranks.sum() - 1.0) < 1e-9


def test_sliced_pr_deterministic_and_close_to_plain(gz, device, rmat14):
    gz.load(rmat14)
    gz.build_source_slices(256)
    gz.run(gz.options("pr", iterations=5, pull_variant="source_sliced"))
    a = gz.pr_props()
    gz.run(gz.options("pr", iterations=5, pull_variant="source_sliced"))
    b = gz.pr_props()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    gz.run(gz.options("pr", iterations=5))
    c = gz.pr_props()
    assert _rel(a, c) <= PR_TOL


def test_sliced_pr_sym_trees_and_empty_slices(gz, orc, device):
    # a graph whose low-id slices hold no sources at all
    V = 5000
    rng = np.random.default_rng(3)
    src = rng.integers(2000, V, 40000)
    dst = rng.integers(0, V, 40000)
    g = gz.make_graph(np.stack([src, dst], 1).astype(np.uint64), V)
    gz.load(g)
    gz.build_source_slices(100)
    counts, _ = gz.source_slices_info()
    assert int(counts[0]) == 0 and int(counts[8]) > 0
    gz.run(gz.options("pr", iterations=10, pull_variant="source_sliced"))
    _, ref, _ = orc.pagerank(orc.Graph(V, g.pull, g.push), orc.options(iterations=10))
    assert _rel(gz.pr_ranks(), ref) <= PR_TOL


@pytest.mark.parametrize("k,T", [(4, 1), (8, 1), (5, 2), (6, 3)])
def test_interleaved_renumbering_and_pr(gz, orc, device, k, T):
    scale, ef, seed = 13, 16, 23
    nv = 1 << scale
    raw = gz.rmat_edges(scale, nv * ef, seed)
    g = gz.make_graph(raw, nv)
    gz.load_rmat_device(scale, nv * ef, seed, relabel=True, slices_log2=k, slice_tiers=T)
    S = min(1 << k, nv // (8 * T))
    assert gz.source_slices_info()[1] == S
    assert len(gz.source_slices_info()[0]) == 8 * T + 1
    # rank t*8S + q of the raw out-degree order (descending, ties by id) gets
    # id t*8S + (q % 8) * S + q // 8 for t < T, else keeps its rank
    cnt = np.bincount(raw[:, 0].astype(np.int64), minlength=nv)
    order = np.lexsort((np.arange(nv), -cnt))
    r = np.arange(nv)
    t, q = r // (8 * S), r % (8 * S)
    want = np.where(r < 8 * S * T, t * 8 * S + (q % 8) * S + q // 8, r)
    got = np.array([gz.lib.gz_graph_device_vertex(int(v)) for v in order])
    assert np.array_equal(got, want)
    gz.run(gz.options("pr", iterations=15, pull_variant="source_sliced"))
    _, ref, _ = orc.pagerank(orc.Graph(nv, g.pull, g.push), orc.options(iterations=15))
    assert _rel(gz.pr_ranks(), ref) <= PR_TOL
    # the plain pull on the same renumbered graph
    gz.run(gz.options("pr", iterations=15))
    assert _rel(gz.pr_ranks(), ref) <= PR_TOL


@pytest.fixture(scope="module")
def rmat17(gz):
    return gz.rmat_graph(17, 16, seed=5)


@pytest.mark.parametrize("hot", [4096, 8192, 16384, 20480])
@pytest.mark.parametrize("S,T", [(4096, 1), (16384, 1), (4096, 3)])
def test_sliced_pr_hot_cache(gz, orc, device, rmat17, S, T, hot):
    g = rmat17
    gz.load(g)
    gz.build_source_slices(S, T)
    gz.slices_set_hot_cache(hot)
    try:
        gz.run(gz.options("pr", iterations=10, pull_variant="source_sliced"))
        a = gz.pr_props()
        ranks = gz.pr_ranks()
        gz.run(gz.options("pr", iterations=10, pull_variant="source_sliced"))
        b = gz.pr_props()
    finally:
        gz.slices_set_hot_cache(0)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=10))
    assert _rel(ranks, ref) <= PR_TOL


@pytest.mark.parametrize("chunk", [64, 1024])
def test_sliced_pr_hot_cache_small_slices(gz, orc, device, rmat14, chunk):
    # windows smaller than the cache, many chunks per XCD
    g = rmat14
    gz.set_chunk_vectors(chunk)
    gz.load(g)
    gz.build_source_slices(64)
    gz.slices_set_hot_cache(4096)
    try:
        gz.run(gz.options("pr", iterations=20, pull_variant="source_sliced"))
        ranks = gz.pr_ranks()
    finally:
        gz.slices_set_hot_cache(0)
        gz.set_chunk_vectors(512)
    _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=20))
    assert _rel(ranks, ref) <= PR_TOL


# ---- Connected Components over the sliced layout --------------------------

import graphs  # noqa: E402


@pytest.fixture(scope="module")
def cc_trees(gz):
    return graphs.planted_trees(gz, base_scale=12, edge_factor=8, trees=16, tree_size=200)


@pytest.mark.parametrize("hot", [0, 4096, 20480])
@pytest.mark.parametrize("S,T", [(64, 1), (512, 1), (128, 2)])
@pytest.mark.parametrize("engine", ["pull", "hybrid"])
def test_sliced_cc_labels_bit_exact(gz, orc, device, cc_trees, S, T, hot, engine):
    g = cc_trees
    gz.load(g)
    gz.build_source_slices(S, T)
    gz.slices_set_hot_cache(hot)
    try:
        gz.run(gz.options("cc", engine=engine, pull_variant="source_sliced"))
        got = gz.cc_labels()
    finally:
        gz.slices_set_hot_cache(0)
    labels, _, _ = orc.cc(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(engine=engine))
    assert np.array_equal(got, labels)


@pytest.mark.parametrize("hot", [0, 20480])
def test_sliced_cc_directed_and_renumbered(gz, orc, device, hot):
    # directed R-MAT: labels are min{u : u reaches v}
    scale, ef, seed = 12, 8, 31
    nv = 1 << scale
    raw = gz.rmat_edges(scale, nv * ef, seed)
    g = gz.make_graph(raw, nv)
    want, _, _ = orc.cc(orc.Graph(nv, g.pull, g.push), orc.options(engine="pull"))
    gz.load(g)
    gz.build_source_slices(32, 1)
    gz.slices_set_hot_cache(hot)
    try:
        gz.run(gz.options("cc", engine="pull", pull_variant="source_sliced"))
        assert np.array_equal(gz.cc_labels(), want)
        gz.load_rmat_device(scale, nv * ef, seed, relabel=True, slices_log2=5)
        gz.run(gz.options("cc", engine="hybrid", pull_variant="source_sliced"))
        assert np.array_equal(gz.cc_labels(), want)
    finally:
        gz.slices_set_hot_cache(0)


def test_sliced_matches_plain_at_scale_20(gz, device):
    # size-independent properties at a size the oracle would be slow on:
    # sliced PageRank equals the plain pull to rounding, ranks sum to 1, and
    # sliced CC labels equal the plain CC labels
    scale = 20
    gz.load_rmat_device(scale, 16 << scale, 77, relabel=True, slices_log2=15)
    gz.run(gz.options("pr", iterations=10))
    plain = gz.pr_ranks()
    gz.set_chunk_vectors(2048)
    gz.slices_set_hot_cache(20480)
    try:
        gz.run(gz.options("pr", iterations=10, pull_variant="source_sliced"))
        sliced = gz.pr_ranks()
        gz.run(gz.options("cc"))
        cc_plain = gz.cc_labels()
        gz.run(gz.options("cc", pull_variant="source_sliced"))
        cc_sliced = gz.cc_labels()
    finally:
        gz.slices_set_hot_cache(0)
        gz.set_chunk_vectors(512)
    assert _rel(sliced, plain) <= PR_TOL
    assert abs(sliced.sum() - 1.0) < 1e-9
    assert np.array_equal(cc_sliced, cc_plain)


# ---- LDS blocks ----------------------------------------------------------------


@pytest.mark.parametrize("hot", [0, 4096])
@pytest.mark.parametrize("vpl", [1, 2, 4])
@pytest.mark.parametrize("chunk", [64, 512])
@pytest.mark.parametrize("S,T,G,H", [(2048, 1, 3, 64), (1024, 2, 4, 100), (512, 3, 4, 64), (256, 1, 6, 32)])
def test_sliced_pr_blocks_parity(gz, orc, device, rmat14, S, T, G, H, chunk, vpl, hot):
    # blocks of H <= the cache are held in LDS one after another; without a
    # cache (hot 0) their gathers read memory
    g = rmat14
    gz.set_chunk_vectors(chunk)
    gz.set_vectors_per_lane(vpl)
    gz.load(g)
    gz.build_source_slices(S, T, G, H)
    gz.slices_set_hot_cache(hot)
    try:
        gz.run(gz.options("pr", iterations=20, pull_variant="source_sliced"))
        a = gz.pr_props()
        ranks = gz.pr_ranks()
        gz.run(gz.options("pr", iterations=20, pull_variant="source_sliced"))
        b = gz.pr_props()
    finally:
        gz.slices_set_hot_cache(0)
        gz.set_chunk_vectors(512)
        gz.set_vectors_per_lane(1)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=20))
    assert _rel(ranks, ref) <= PR_TOL
    assert abs(ranks.sum() - 1.0) < 1e-9


@pytest.mark.parametrize("hot", [0, 4096])
@pytest.mark.parametrize("S,T,G,H", [(256, 1, 3, 32), (128, 2, 5, 16)])
@pytest.mark.parametrize("engine", ["pull", "hybrid"])
def test_sliced_cc_blocks_bit_exact(gz, orc, device, cc_trees, S, T, G, H, hot, engine):
    g = cc_trees
    gz.load(g)
    gz.build_source_slices(S, T, G, H)
    assert gz.source_slices_layout()[2] == G
    gz.slices_set_hot_cache(hot)
    try:
        gz.run(gz.options("cc", engine=engine, pull_variant="source_sliced"))
        got = gz.cc_labels()
    finally:
        gz.slices_set_hot_cache(0)
    labels, _, _ = orc.cc(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(engine=engine))
    assert np.array_equal(got, labels)


@pytest.mark.parametrize("hot", [16384, 20480])
def test_sliced_blocks_at_scale_20(gz, device, hot):
    # blocks of 20480 sources in two 2^16-vertex tiers: held in LDS with the
    # 20480 cache, read from memory with the smaller one; PageRank equals the
    # plain pull to rounding and CC labels equal the plain labels
    scale = 20
    gz.load_rmat_device(scale, 16 << scale, 78, relabel=True, slices_log2=16, slice_tiers=2, slice_blocks=2)
    assert gz.source_slices_layout() == (1 << 16, 2, 2, gz.SLICE_BLOCK_VERTICES)
    gz.run(gz.options("pr", iterations=10))
    plain = gz.pr_ranks()
    gz.run(gz.options("cc"))
    cc_plain = gz.cc_labels()
    gz.set_chunk_vectors(2048)
    gz.slices_set_hot_cache(hot)
    try:
        gz.run(gz.options("pr", iterations=10, pull_variant="source_sliced"))
        sliced = gz.pr_ranks()
        gz.run(gz.options("cc", pull_variant="source_sliced"))
        cc_sliced = gz.cc_labels()
    finally:
        gz.slices_set_hot_cache(0)
        gz.set_chunk_vectors(512)
    assert _rel(sliced, plain) <= PR_TOL
    assert abs(sliced.sum() - 1.0) < 1e-9
    assert np.array_equal(cc_sliced, cc_plain)
