"""Breadth-first searches from many roots in one pass per level
(gz_bfs_batch_run) against the oracle and against single-root runs.

Bars: every search's depths bit-exact against the oracle's BFS from that
root and against the single-root HIP path; parents a valid BFS tree (the
oracle's BFS certificate: each parent an in-neighbour one level up, no edge
that would shorten a depth).  Roots may repeat and may have no out-edges.
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def _check(gz, orc, G, roots, exact_single=True):
    res = gz.bfs_batch(roots)
    depths_max = 0
    for i, root in enumerate(roots):
        par, dep = gz.bfs_batch_result(i)
        _, want, _, _ = orc.bfs(G, orc.options(root=int(root)))
        assert np.array_equal(dep, want), (i, root)
        assert orc.bfs_certificate_violations(G, int(root), par, dep) == 0, (i, root)
        reached = dep[dep != np.iinfo(np.uint64).max]
        depths_max = max(depths_max, int(reached.max()))
        if exact_single and i < 8:
            gz.run(gz.options("bfs", root=int(root)))
            assert np.array_equal(gz.bfs_result()[1], dep), (i, root)
    # one level per depth, and the last level finds nothing
    assert res.iterations == depths_max + 1
    assert res.pull_iterations + res.push_iterations == res.iterations
    return res


@pytest.mark.parametrize("count", [1, 7, 64])
def test_batch_depths_rmat14(gz, orc, device, count):
    g = gz.rmat_graph(14, 16, seed=11)
    gz.load(g)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    rng = np.random.default_rng(count)
    roots = rng.integers(0, g.num_vertices, size=count)
    if count >= 7:
        roots[3] = roots[1]                       # a repeated root
        od = np.bincount(g.pull[:, 0].astype(np.int64), minlength=g.num_vertices)
        roots[5] = int(np.nonzero(od == 0)[0][0])  # a root without out-edges
    res = _check(gz, orc, G, roots)
    if count == 64:
        assert res.pull_iterations > 0 and res.push_iterations > 0


def test_batch_deep_star_path(gz, orc, device):
    # 200 levels along the path, then a 20,000-leaf star: depths well past
    # any R-MAT graph's, one root at each end and one at the centre
    P, L = 200, 20000
    pairs = [(i, i + 1) for i in range(P)] + [(P, P + 1 + j) for j in range(L)]
    g = gz.make_graph(np.asarray(pairs, dtype=np.uint64), P + 1 + L, symmetrize=True)
    gz.load(g)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    _check(gz, orc, G, [0, P, P + 7, 0])


def test_batch_directed_planted_trees(gz, orc, device):
    g = graphs.planted_trees(gz, base_scale=12)
    gz.load(g)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    _check(gz, orc, G, [0, 1, 4096, 4096 + 150, 4096 + 799, 17])


def test_batch_renumbered_graph(gz, orc, device):
    # device ingress with renumbering: roots and results in original ids
    scale, seed = 13, 21
    g = gz.rmat_graph(scale, 16, seed)
    gz.load_rmat_device(scale, (1 << scale) * 16, seed, relabel=True)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    roots = np.random.default_rng(5).integers(0, g.num_vertices, size=32)
    res = gz.bfs_batch(roots)
    for i, root in enumerate(roots):
        par, dep = gz.bfs_batch_result(i)
        _, want, _, _ = orc.bfs(G, orc.options(root=int(root)))
        assert np.array_equal(dep, want), (i, root)
        assert orc.bfs_certificate_violations(G, int(root), par, dep) == 0, (i, root)
    assert res.iterations > 0


def test_batch_rejects_bad_requests(gz, device):
    P = 300
    g = gz.make_graph(np.asarray([(i, i + 1) for i in range(P)], dtype=np.uint64), P + 1)
    gz.load(g)
    with pytest.raises(ValueError):
        gz.bfs_batch([])
    with pytest.raises(ValueError):
        gz.bfs_batch(list(range(65)))
    with pytest.raises(ValueError):
        gz.bfs_batch([0, P + 1])
    with pytest.raises(ValueError):
        gz.bfs_batch([0])          # 300 levels: past the 254 a byte records
    res = gz.bfs_batch([100])      # 200 levels fit
    assert res.iterations == 201
