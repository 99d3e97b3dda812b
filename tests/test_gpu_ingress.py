"""Device ingress versus the host packer.

The device pipeline (radix sort of (group, neighbour) keys, dedupe, binary-
searched group offsets, scatter pack) must give the Vector-Sparse lists, CSR
offsets and FP64 outdegrees of gz_edges_sort + gz_pack_vectors bit for bit,
invalid lanes included.
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def _host_state(gz, g):
    gz.load(g)
    return _state(gz)


def _state(gz):
    return {
        "V": gz.num_vertices(),
        "E": gz.num_edges(),
        "pull": gz.device_vectors(True),
        "push": gz.device_vectors(False),
        "pull_off": gz.device_offsets(True),
        "push_off": gz.device_offsets(False),
        "outdeg": gz.outdegrees(),
    }


def _assert_same(a, b):
    assert a["V"] == b["V"] and a["E"] == b["E"]
    for k in ("pull", "push", "pull_off", "push_off"):
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["outdeg"].view(np.uint64), b["outdeg"].view(np.uint64))


@pytest.mark.parametrize("symmetrize", [False, True])
@pytest.mark.parametrize("scale,ef,seed", [(10, 16, 3), (14, 8, 7)])
def test_rmat_device_equals_host(gz, device, scale, ef, seed, symmetrize):
    nv = 1 << scale
    raw = gz.rmat_edges(scale, nv * ef, seed)
    host = _host_state(gz, gz.make_graph(raw, nv, symmetrize=symmetrize))
    gz.load_rmat_device(scale, nv * ef, seed, symmetrize=symmetrize)
    dev = _state(gz)
    _assert_same(dev, host)
    gz.load_pairs_device(nv, raw, symmetrize=symmetrize)
    _assert_same(_state(gz), host)


def test_folded_rmat_with_extra_pairs(gz, device):
    # ids folded into a non-power-of-two range, plus planted trees after it
    scale, nv_fold = 12, 3001
    raw = gz.rmat_edges(scale, 40000, 9, num_vertices=nv_fold)
    trees = gz.planted_trees(nv_fold, 5, 50, 4)
    V = nv_fold + 5 * 50 + 7  # a few isolated vertices at the top
    host = _host_state(gz, gz.make_graph(np.concatenate([raw, trees]), V, symmetrize=True))
    gz.load_rmat_device(scale, 40000, 9, symmetrize=True, num_vertices=V, extra=trees, rmat_vertices=nv_fold)
    _assert_same(_state(gz), host)


def test_cc_workload_device_equals_host(gz, device):
    host = _host_state(gz, gz.workload_graph("cc-rmat16-sym-trees"))
    gz.load_workload_device("cc-rmat16-sym-trees")
    _assert_same(_state(gz), host)


@pytest.mark.parametrize("pairs,V", [
    ([], 5),
    ([(3, 3), (4, 4)], 6),                       # self-loops only
    ([(0, 1)] * 9 + [(1, 0)], 2),                # duplicates
    ([(0, 1)], 1 << 20),                         # nearly all vertices isolated
    ([(i, 7) for i in range(100) if i != 7], 100),  # one destination, 25 vectors
])
def test_edge_cases(gz, device, pairs, V):
    p = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    host = _host_state(gz, gz.make_graph(p, V))
    gz.load_pairs_device(V, p)
    _assert_same(_state(gz), host)


def test_pagerank_on_device_built_graph(gz, orc, device):
    g = graphs.planted_trees(gz)
    gz.load_pairs_device(g.num_vertices, g.pull)
    gz.run(gz.options("pr", iterations=10))
    ranks = gz.pr_ranks()
    _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=10))
    assert float(np.max(np.abs(ranks - ref) / ref)) <= 1e-12


# ---- renumbering by out-degree ------------------------------------------------

def _bfs_tree_ok(g, root, parents, depths):
    """parents[v] is an in-neighbour of v one level up; root and unreached -1."""
    src, dst = g.pull[:, 0].astype(np.int64), g.pull[:, 1].astype(np.int64)
    reached = depths != np.iinfo(np.uint64).max
    assert parents[root] == -1 and depths[root] == 0
    v = np.nonzero(reached & (np.arange(g.num_vertices) != root))[0]
    p = parents[v]
    assert np.all(p >= 0)
    assert np.array_equal(depths[p], depths[v] - 1)
    edges = set(zip(src.tolist(), dst.tolist()))
    assert all((int(a), int(b)) in edges for a, b in zip(p, v))
    assert np.all(parents[~reached] == -1)


@pytest.fixture(scope="module")
def sym_trees(gz):
    return graphs.planted_trees(gz)


def test_relabel_edges_map_back(gz, device):
    scale, ef, seed = 12, 8, 21
    nv = 1 << scale
    raw = gz.rmat_edges(scale, nv * ef, seed)
    g = gz.make_graph(raw, nv)
    gz.load_rmat_device(scale, nv * ef, seed, relabel=True)
    vecs = gz.device_vectors(True)
    valid = (vecs >> np.uint64(63)).astype(bool)
    nbr = vecs & np.uint64((1 << 48) - 1)
    piece = (vecs >> np.uint64(48)) & np.uint64(0x7FFF)
    shared = piece[:, 0] | (piece[:, 1] << np.uint64(15)) | (piece[:, 2] << np.uint64(30)) | \
        ((piece[:, 3] & np.uint64(7)) << np.uint64(45))
    dst_new = np.repeat(shared[:, None], 4, axis=1)[valid]
    src_new = nbr[valid]
    orig = np.empty(nv, dtype=np.uint64)
    for v in range(nv):
        orig[gz.lib.gz_graph_device_vertex(v)] = v
    got = np.stack([orig[src_new.astype(np.int64)], orig[dst_new.astype(np.int64)]], axis=1)
    got = got[np.lexsort((got[:, 0], got[:, 1]))]
    assert np.array_equal(got, g.pull)
    # new ids are ordered by raw out-degree, highest first, ties by old id
    raw_cnt = np.bincount(raw[:, 0].astype(np.int64), minlength=nv)[orig.astype(np.int64)]
    assert np.all(np.diff(raw_cnt) <= 0)
    ties = np.diff(raw_cnt) == 0
    assert np.all(np.diff(orig.astype(np.int64))[ties] > 0)


def test_relabel_pagerank(gz, orc, device, sym_trees):
    g = sym_trees
    gz.load_pairs_device(g.num_vertices, g.pull, relabel=True)
    gz.run(gz.options("pr", iterations=10))
    ranks = gz.pr_ranks()
    _, ref, _ = orc.pagerank(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(iterations=10))
    assert float(np.max(np.abs(ranks - ref) / ref)) <= 1e-12


@pytest.mark.parametrize("engine", ["hybrid", "pull", "push"])
def test_relabel_cc_labels_bit_exact(gz, orc, device, sym_trees, engine):
    g = sym_trees
    gz.load_pairs_device(g.num_vertices, g.pull, relabel=True)
    gz.run(gz.options("cc", engine=engine))
    labels, _, _ = orc.cc(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(engine=engine))
    assert np.array_equal(gz.cc_labels(), labels)


@pytest.mark.parametrize("engine", ["hybrid", "pull", "push"])
def test_relabel_bfs_depths_and_tree(gz, orc, device, engine):
    g = gz.rmat_graph(13, 16, seed=4)
    gz.load_pairs_device(g.num_vertices, g.pull, relabel=True)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    for root in (0, 5, 77):
        gz.run(gz.options("bfs", engine=engine, root=root))
        par, dep, _, _ = orc.bfs(G, orc.options(engine=engine, root=root))
        p, d = gz.bfs_result()
        assert np.array_equal(d, dep)
        _bfs_tree_ok(g, root, p, d)
