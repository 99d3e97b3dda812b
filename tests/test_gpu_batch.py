"""Batches of independent graphs (disjoint union, per-graph PageRank state).

Every graph of a batch must get exactly what it gets alone: PageRank within
1e-12 of the oracle run on that graph by itself, CC labels and BFS depths
bit-exact, for every batch configuration (vectors per lane, workgroup size,
chunk length per size bucket), eagerly and replayed from a captured graph.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PR_TOL = 1e-12

SPECS = [(6, 4, 11), (9, 16, 12), (7, 30, 13), (12, 8, 14), (10, 20, 15), (8, 5, 16), (13, 6, 17), (11, 32, 18),
         (5, 3, 19), (12, 17, 20)]


@pytest.fixture(scope="module")
def singles(gz):
    return [gz.rmat_graph(s, ef, seed) for s, ef, seed in SPECS]


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _pr_refs(orc, singles, iterations, damping=0.875):
    out = []
    for g in singles:
        _, ranks, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push),
                                   orc.options(iterations=iterations, damping=damping))
        out.append(ranks)
    return out


def test_union_layout_is_the_graphs_back_to_back(gz, device, singles):
    bases = gz.load_batch_device(SPECS)
    assert gz.num_edges() == sum(g.num_edges for g in singles)
    vecs = gz.device_vectors(True)
    offs = gz.device_offsets(True)
    for (g, b0, b1) in zip(singles, bases[:-1], bases[1:]):
        lo, hi = int(offs[b0]), int(offs[b1])
        mine = gz.pack_vectors(g.pull + np.uint64(b0), True)
        assert np.array_equal(vecs[lo:hi], mine)


@pytest.mark.parametrize("cfg", [
    dict(vectors_per_lane=1, workgroup_size=256, chunk=(512, 512, 512, 512)),
    dict(vectors_per_lane=1, workgroup_size=64, chunk=(64, 128, 192, 4096)),
    dict(vectors_per_lane=2, workgroup_size=512, chunk=(128, 256, 1024, 2048)),
    dict(vectors_per_lane=4, workgroup_size=1024, chunk=(256, 256, 512, 1024)),
])
def test_batched_pagerank_matches_single_graph_oracle(gz, orc, device, singles, cfg):
    bases = gz.load_batch_device(SPECS)
    gz.batch_set_config(**cfg)
    refs = _pr_refs(orc, singles, 10)
    for use_graph in (False, True):
        gz.batch_pr(10, use_graph=use_graph)
        ranks = gz.batch_pr_ranks()
        for g, ref in enumerate(refs):
            got = ranks[int(bases[g]):int(bases[g + 1])]
            assert _rel(got, ref) <= PR_TOL, (cfg, use_graph, g)


def test_graph_replays_are_identical(gz, device):
    gz.load_batch_device(SPECS)
    gz.batch_set_config()
    gz.batch_pr(7, use_graph=False)
    eager = gz.batch_pr_ranks()
    for _ in range(3):
        gz.batch_pr(7, use_graph=True)  # the second and third calls replay the captured graph
        assert np.array_equal(gz.batch_pr_ranks().view(np.uint64), eager.view(np.uint64))


def test_batched_cc_and_bfs(gz, orc, device):
    specs = [(s, ef, seed) for s, ef, seed in SPECS]
    bases = gz.load_batch_device(specs, symmetrize=True)
    syms = [gz.rmat_graph(s, ef, seed, symmetrize=True) for s, ef, seed in specs]
    gz.run(gz.options("cc"))
    labels = gz.cc_labels()
    for g, sg in enumerate(syms):
        ref, _, _ = orc.cc(orc.Graph(sg.num_vertices, sg.pull, sg.push), orc.options())
        got = labels[int(bases[g]):int(bases[g + 1])] - bases[g]
        assert np.array_equal(got, ref), g
    # BFS from a root in graph 3 reaches only graph 3
    g = 3
    sg = syms[g]
    root = int(np.nonzero(np.bincount(sg.pull[:, 0].astype(np.int64), minlength=sg.num_vertices))[0][0])
    gz.run(gz.options("bfs", root=int(bases[g]) + root))
    _, d = gz.bfs_result()
    _, dep, _, _ = orc.bfs(orc.Graph(sg.num_vertices, sg.pull, sg.push), orc.options(root=root))
    assert np.array_equal(d[int(bases[g]):int(bases[g + 1])], dep)
    outside = np.ones(d.shape[0], bool)
    outside[int(bases[g]):int(bases[g + 1])] = False
    assert np.all(d[outside] == np.iinfo(np.uint64).max)


def test_autotuner_probe(gz, device):
    gz.load_batch_device(SPECS)
    for b in range(gz.BATCH_BUCKETS):
        t = gz.batch_time_pull(1, 256, (256, 256, 256, 256), b, reps=2)
        assert t >= 0.0
    assert gz.batch_time_pull(3, 256, (256, 256, 256, 256), 0) == -1.0  # invalid vectors per lane
    # the probe leaves the batch configured as before
    gz.batch_set_config(vectors_per_lane=2, workgroup_size=128, chunk=(256, 512, 512, 1024))
    gz.batch_time_pull(1, 64, (64, 64, 64, 64), 1)
    assert gz.batch_get_config() == {"vectors_per_lane": 2, "workgroup_size": 128, "chunk": [256, 512, 512, 1024]}
