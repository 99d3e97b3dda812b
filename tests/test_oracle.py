"""Pins of the CPU oracle (the checker of the HIP path).

The reference repository has no tests, golden files or fixtures, and its
phase kernels cannot be built here, so each oracle function is pinned by
what the mathematics fixes, with checks chosen so that a plausible slip
fails one of them:
  - scheduler split: exhaustive partition check over small (vectors, units);
  - PageRank: closed form of a 2-vertex graph, the stationary vector of the
    Google matrix (numpy eigensolver) after convergence, one-step agreement
    with a scipy sparse mat-vec for many unit splittings (pins pull + merge +
    vertex phase), Sigma rank = 1, push == pull;
  - CC: scipy connected components (symmetric) and brute-force reachability
    (directed), planted trees;
  - BFS: scipy unweighted shortest paths for depths, tree validity, the
    single-thread parent rules by brute force, and the per-iteration
    frontier statistic / engine sequence recomputed from the depths;
  - selector MLP: gradient against central finite differences.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import graphs


def adjacency(g):
    V = g.num_vertices
    return sp.csr_matrix((np.ones(g.num_edges), (g.push[:, 0].astype(np.int64), g.push[:, 1].astype(np.int64))),
                         shape=(V, V))


# ---- scheduler ----------------------------------------------------------------

def test_unit_range_partition(orc):
    for n in range(0, 120):
        for u in range(1, 40):
            ranges = [orc.unit_range(n, u, k) for k in range(u)]
            assert ranges[0][0] == 0 and ranges[-1][1] == n
            for (a0, a1), (b0, b1) in zip(ranges, ranges[1:]):
                assert a1 == b0
            sizes = [b - a for a, b in ranges]
            assert max(sizes) - min(sizes) <= 1
            assert sizes == sorted(sizes, reverse=True)


# ---- PageRank -------------------------------------------------------------------

def google_stationary(g, d):
    """Stationary vector of the damped walk with sinks jumping uniformly."""
    V = g.num_vertices
    A = adjacency(g).toarray()
    out = A.sum(axis=1)
    P = np.zeros((V, V))
    for u in range(V):
        P[u] = A[u] / out[u] if out[u] > 0 else 1.0 / V
    G = d * P + (1 - d) / V
    w, vecs = np.linalg.eig(G.T)
    x = np.real(vecs[:, np.argmin(np.abs(w - 1))])
    return x / x.sum()


def test_pr_two_vertex_closed_form(gz, orc):
    g = graphs.edge_lists(gz, [(0, 1)], 2)
    G = orc.Graph(2, g.pull, g.push)
    for d in (0.875, 0.85):
        _, ranks, _ = orc.pagerank(G, orc.options(iterations=200, damping=d))
        assert abs(ranks[0] - 1.0 / (2.0 + d)) < 1e-14
        assert abs(ranks[1] - (1.0 + d) / (2.0 + d)) < 1e-14


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pr_stationary_vector(gz, orc, seed):
    rng = np.random.default_rng(seed)
    V = 40
    pairs = rng.integers(0, V, size=(120, 2), dtype=np.uint64)
    g = gz.make_graph(pairs, V)
    G = orc.Graph(V, g.pull, g.push)
    for d in (0.875, 0.85):
        _, ranks, _ = orc.pagerank(G, orc.options(iterations=400, damping=d))
        assert np.max(np.abs(ranks - google_stationary(g, d))) < 1e-12
        assert abs(ranks.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("threads,units", [(1, 1), (1, 32), (3, 7), (8, 32), (5, 101)])
def test_pr_one_step_matches_sparse_matvec(gz, orc, threads, units):
    g = graphs.hub_graph(gz, num_vertices=800, hub_in=3000, seed=threads)
    V = g.num_vertices
    G = orc.Graph(V, g.pull, g.push)
    od = np.bincount(g.pull[:, 0].astype(np.int64), minlength=V).astype(np.float64)
    odp = np.where(od == 0, V, od)
    p0 = (1.0 / V) / odp
    acc = adjacency(g).T @ p0
    S = acc.sum()
    d = 0.875
    want = ((acc + (1 - S) / V) * d + (1 - d) / V) / odp
    props, _, _ = orc.pagerank(G, orc.options(iterations=1, threads=threads, units_per_thread=units))
    assert np.max(np.abs(props - want) / want) < 1e-13


def test_pr_variants_agree(gz, orc):
    g = gz.rmat_graph(11, 16, seed=8)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    _, base, _ = orc.pagerank(G, orc.options(iterations=20))
    for o in (orc.options(iterations=20, engine="push"), orc.options(iterations=20, traditional=True),
              orc.options(iterations=20, threads=8)):
        _, r, _ = orc.pagerank(G, o)
        assert np.max(np.abs(r - base) / base) < 1e-12
    # vertices without in-edges all end with the same rank
    no_in = np.setdiff1d(np.arange(g.num_vertices), g.pull[:, 1].astype(np.int64))
    assert np.ptp(base[no_in]) < 1e-18


# ---- CC ---------------------------------------------------------------------------

def min_reach_labels(g):
    """label(v) = min{u : u reaches v} by brute force."""
    V = g.num_vertices
    D = csg.shortest_path(adjacency(g), unweighted=True)
    lab = np.empty(V, dtype=np.uint64)
    for v in range(V):
        lab[v] = np.min(np.nonzero(~np.isinf(D[:, v]))[0])
    return lab


@pytest.mark.parametrize("engine", ["hybrid", "pull", "push"])
def test_cc_symmetric_components(gz, orc, engine):
    g = graphs.planted_trees(gz)
    lab, log, r = orc.cc(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(engine=engine))
    n, comp = csg.connected_components(adjacency(g), directed=False)
    mins = np.full(n, np.iinfo(np.int64).max)
    np.minimum.at(mins, comp, np.arange(g.num_vertices))
    assert np.array_equal(lab, mins[comp].astype(np.uint64))
    base = 1 << 10
    for t in range(8):
        assert np.all(lab[base + t * 100: base + (t + 1) * 100] == base + t * 100)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cc_directed_reachability(gz, orc, seed):
    rng = np.random.default_rng(seed)
    V = 60
    g = gz.make_graph(rng.integers(0, V, size=(90, 2), dtype=np.uint64), V)
    for engine in ("hybrid", "pull", "push"):
        for threads in (1, 4):
            lab, _, _ = orc.cc(orc.Graph(V, g.pull, g.push), orc.options(engine=engine, threads=threads))
            assert np.array_equal(lab, min_reach_labels(g))


# ---- BFS --------------------------------------------------------------------------

def bfs_expectations(g, root):
    V = g.num_vertices
    dist = csg.shortest_path(adjacency(g), unweighted=True, indices=root)
    depth = np.full(V, np.iinfo(np.uint64).max, dtype=np.uint64)
    fin = ~np.isinf(dist)
    depth[fin] = dist[fin].astype(np.uint64)
    return depth


@pytest.mark.parametrize("engine", ["hybrid", "pull", "push"])
@pytest.mark.parametrize("stat", ["both", "outdeg", "count"])
def test_bfs_depths_parents_and_votes(gz, orc, engine, stat):
    g = gz.rmat_graph(11, 8, seed=13)
    V, E = g.num_vertices, g.num_edges
    G = orc.Graph(V, g.pull, g.push)
    od = np.bincount(g.pull[:, 0].astype(np.int64), minlength=V)
    vecs = orc.pack(g.pull, True)
    for root in (0, 3, 100):
        par, dep, log, r = orc.bfs(G, orc.options(engine=engine, stat=stat, root=root))
        want = bfs_expectations(g, root)
        assert np.array_equal(dep, want)
        # tree validity
        src, dst = g.pull[:, 0].astype(np.int64), g.pull[:, 1].astype(np.int64)
        edges = set(zip(src.tolist(), dst.tolist()))
        for v in np.nonzero((want != np.iinfo(np.uint64).max))[0]:
            if v == root:
                assert par[v] == -1
            else:
                assert (int(par[v]), int(v)) in edges and dep[par[v]] == dep[v] - 1
        assert np.all(par[want == np.iinfo(np.uint64).max] == -1)
        # frontier statistic per iteration, recomputed from the depths
        stat_of = lambda vs: (0 if stat == "outdeg" else len(vs)) + (0 if stat == "count" else int(od[vs].sum()))
        votes = [stat_of(np.array([root]))]
        k = 1
        while True:
            vs = np.nonzero(dep == k)[0]
            votes.append(stat_of(vs))
            if len(vs) == 0:
                break
            k += 1
        assert [v for _, v in log] == votes[:len(log)] and len(log) == len(votes) - 1
        thr = V // 2 if stat == "count" else E // 5
        for (eng, vote) in log:
            if engine == "hybrid":
                assert eng == (1 if vote > thr else 2)
            else:
                assert eng == (1 if engine == "pull" else 2)
        # single-thread parent rules
        for v in np.nonzero((dep != np.iinfo(np.uint64).max) & (np.arange(V) != root))[0][:300]:
            lvl = int(dep[v])
            eng = log[lvl - 1][0]
            ins = src[dst == v]
            frontier = ins[dep[ins] == lvl - 1]
            if eng == 2:
                assert par[v] == frontier.min()
            else:
                rows = vecs[np.nonzero(((vecs[:, 0] >> np.uint64(48)) & np.uint64(0x7FFF)).astype(np.int64)
                                       | (((vecs[:, 1] >> np.uint64(48)) & np.uint64(0x7FFF)).astype(np.int64) << 15)
                                       == v)[0]]
                for row in rows:
                    ids = [int(x) & ((1 << 48) - 1) for x in row if int(x) >> 63]
                    q = [s for s in ids if dep[s] == lvl - 1]
                    if q:
                        assert par[v] == max(q)
                        break


def test_bfs_threads_same_depths(gz, orc):
    g = gz.rmat_graph(12, 16, seed=2)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    _, d1, _, _ = orc.bfs(G, orc.options(root=1))
    _, d8, _, _ = orc.bfs(G, orc.options(root=1, threads=8))
    assert np.array_equal(d1, d8)


def test_bfs_bad_root(gz, orc):
    g = graphs.edge_lists(gz, [(0, 1)], 2)
    with pytest.raises(ValueError):
        orc.bfs(orc.Graph(2, g.pull, g.push), orc.options(root=5))


# ---- selector MLP -----------------------------------------------------------------

def test_mlp_gradient_finite_differences(orc):
    rng = np.random.default_rng(0)
    n_params = 32 * 8 + 32 + 32 * 32 + 32 + 2 * 32 + 2
    p = (rng.standard_normal(n_params) * 0.3).astype(np.float32)
    X = rng.standard_normal((16, 8)).astype(np.float32)
    T = rng.standard_normal((16, 2)).astype(np.float32)
    loss, g = orc.mlp_gradient(p, X, T)
    for i in rng.choice(n_params, 40, replace=False):
        h = 1e-2
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (orc.mlp_gradient(pp, X, T)[0] - orc.mlp_gradient(pm, X, T)[0]) / (2 * h)
        assert abs(num - g[i]) <= 2e-3 + 2e-2 * abs(g[i])
    y = orc.mlp_forward(p, X[0])
    assert y.shape == (2,)


def _np_mlp(p, X):
    """The selector's forward pass restated with numpy matmuls over the
    parameter layout the selector header documents: W1[32][8] b1[32]
    W2[32][32] b2[32] W3[2][32] b3[2], row-major, ReLU after both hidden
    layers, a linear output."""
    p = p.astype(np.float64)
    o = 0
    W1 = p[o:o + 256].reshape(32, 8); o += 256
    b1 = p[o:o + 32]; o += 32
    W2 = p[o:o + 1024].reshape(32, 32); o += 1024
    b2 = p[o:o + 32]; o += 32
    W3 = p[o:o + 64].reshape(2, 32); o += 64
    b3 = p[o:o + 2]; o += 2
    assert o == p.shape[0]
    h1 = np.maximum(X.astype(np.float64) @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3


def test_mlp_forward_matches_numpy(orc):
    # random weights with non-zero biases; the hidden layers are non-square
    # (32x8) or asymmetric (32x32), so a transposed matrix, a swapped bias or
    # a missing ReLU changes the outputs
    rng = np.random.default_rng(11)
    p = (rng.standard_normal(1410) * 0.4).astype(np.float32)
    X = rng.standard_normal((200, 8)).astype(np.float32)
    want = _np_mlp(p, X)
    got = np.stack([orc.mlp_forward(p, x) for x in X])
    assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, float(np.max(np.abs(want))))
    # the loss the gradient routine reports is the mean squared error of
    # the same forward pass over both outputs
    T = rng.standard_normal((200, 2)).astype(np.float32)
    loss, _ = orc.mlp_gradient(p, X, T)
    assert loss == pytest.approx(float(np.mean((want - T) ** 2)), rel=1e-5)


def test_mlp_forward_hand_case(orc):
    # y0 = relu(relu(x0 - x1) * 2 + 1) * 3 - 4, y1 = relu(x2) via the second
    # unit of each layer; all other weights zero
    p = np.zeros(1410, np.float32)
    W1, b1, W2, b2, W3, b3 = 0, 256, 288, 1312, 1344, 1408
    p[W1 + 0 * 8 + 0], p[W1 + 0 * 8 + 1] = 1.0, -1.0
    p[W1 + 1 * 8 + 2] = 1.0
    p[W2 + 0 * 32 + 0], p[b2 + 0] = 2.0, 1.0
    p[W2 + 1 * 32 + 1] = 1.0
    p[W3 + 0 * 32 + 0], p[b3 + 0] = 3.0, -4.0
    p[W3 + 1 * 32 + 1] = 1.0
    for x0, x1, x2 in ((3.0, 1.0, 2.0), (1.0, 3.0, -2.0), (0.5, 0.25, 7.0)):
        x = np.array([x0, x1, x2, 9, 9, 9, 9, 9], np.float32)
        y = orc.mlp_forward(p, x)
        assert y[0] == pytest.approx(max(max(x0 - x1, 0) * 2 + 1, 0) * 3 - 4, abs=1e-6)
        assert y[1] == pytest.approx(max(x2, 0), abs=1e-6)


# ---- graphs over packed lists, CC from given labels ------------------------------

def test_graph_from_vectors_matches_pairs(gz, orc):
    g = gz.rmat_graph(11, 12, seed=23)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    H = orc.Graph.from_vectors(g.num_vertices, orc.pack(g.pull, True), orc.pack(g.push, False))
    _, r1, _ = orc.pagerank(G, orc.options(iterations=10))
    _, r2, _ = orc.pagerank(H, orc.options(iterations=10))
    assert np.array_equal(r1.view(np.uint64), r2.view(np.uint64))
    assert np.array_equal(orc.cc(G, orc.options())[0], orc.cc(H, orc.options())[0])
    for root in (0, 9):
        a = orc.bfs(G, orc.options(root=root))
        b = orc.bfs(H, orc.options(root=root))
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    # without a pull list: E and outdegrees from the push list
    P = orc.Graph.from_vectors(g.num_vertices, None, orc.pack(g.push, False))
    assert np.array_equal(orc.cc(P, orc.options(engine="push"))[0], orc.cc(G, orc.options(engine="push"))[0])


def test_cc_from_original_ids_on_renumbered_graph(gz, orc):
    # renumber a directed graph by a random permutation (new = perm[old]);
    # CC on the renumbered copy, started from the original ids, gives each
    # vertex the label the original graph gives it: the smallest original id
    # that reaches it
    g = gz.rmat_graph(10, 6, seed=31)
    V = g.num_vertices
    perm = np.random.default_rng(2).permutation(V).astype(np.uint64)
    orig = np.empty(V, np.uint64)
    orig[perm] = np.arange(V, dtype=np.uint64)
    h = gz.make_graph(perm[g.pull.astype(np.int64)], V)
    G = orc.Graph(V, g.pull, g.push)
    H = orc.Graph(V, h.pull, h.push)
    for engine in ("hybrid", "pull", "push"):
        want = orc.cc(G, orc.options(engine=engine))[0]
        got = orc.cc(H, orc.options(engine=engine), init=orig)[0]
        assert np.array_equal(got[perm.astype(np.int64)], want)
    # started from the ids, the renumbered copy's labels differ
    assert not np.array_equal(orc.cc(H, orc.options())[0][perm.astype(np.int64)], orc.cc(G, orc.options())[0])


def test_bfs_certificate(gz, orc):
    # accepts the BFS the oracle computes and each planted mistake fails it:
    # a depth one too large, one too small, a parent that is no in-neighbour,
    # a parent one level too high, an unreached vertex marked reached
    g = gz.rmat_graph(11, 8, seed=37)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    root = 5
    par, dep, _, _ = orc.bfs(G, orc.options(root=root))
    assert orc.bfs_certificate_violations(G, root, par, dep) == 0
    _, d8, _, _ = orc.bfs(G, orc.options(root=root, threads=8, engine="push"))
    p8 = orc.bfs(G, orc.options(root=root, threads=8, engine="push"))[0]
    assert orc.bfs_certificate_violations(G, root, p8, d8, threads=4) == 0
    reached = np.nonzero((dep != np.iinfo(np.uint64).max) & (np.arange(g.num_vertices) != root))[0]
    deep = reached[np.argmax(dep[reached])]
    src, dst = g.pull[:, 0].astype(np.int64), g.pull[:, 1].astype(np.int64)
    for mutate in ("deeper", "shallower", "not_edge", "wrong_level", "unreached"):
        p, d = par.copy(), dep.copy()
        if mutate == "deeper":
            d[deep] += 1
        elif mutate == "shallower":
            d[deep] -= 1
        elif mutate == "not_edge":
            ins = set(src[dst == deep].tolist())
            p[deep] = next(u for u in reached if u not in ins and u != deep)
        elif mutate == "wrong_level":
            ins = src[dst == deep]
            cand = [u for u in ins if dep[u] != np.iinfo(np.uint64).max and dep[u] + 1 != dep[deep]]
            if not cand:
                continue
            p[deep] = cand[0]
        else:
            un = np.nonzero(dep == np.iinfo(np.uint64).max)[0]
            if un.size == 0:
                continue
            d[un[0]], p[un[0]] = 3, root
        assert orc.bfs_certificate_violations(G, root, p, d) > 0, mutate
