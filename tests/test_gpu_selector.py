"""Learned engine selector on the device.

The MLP kernels (forward, full-batch gradient, Adam) are checked against the
oracle's double-precision MLP (itself pinned by finite differences in
test_oracle.py).  The device feature pass is checked against what BFS depths
fix exactly (frontier sizes per level, unvisited counts).  The probe engine
must leave results and the engine sequence of the fixed-threshold hybrid
unchanged, and the learned engine must keep depths and labels exact.  The
last test measures what the committed network gains over the fixed
threshold on held-out graphs.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "grazelle-ppopp18_amd",
                       "selector_bfs.bin")


def _params(seed):
    import selector_train as st
    p = st.init_params(seed)
    rng = np.random.default_rng(seed)
    p += rng.normal(0, 0.05, size=p.shape).astype(np.float32)  # non-zero biases
    return p


def test_forward_matches_oracle(gz, orc, device):
    p = _params(3)
    X = np.random.default_rng(4).normal(size=(300, 8)).astype(np.float32)
    Y = gz.selector_forward(p, X)
    ref = np.stack([orc.mlp_forward(p, x) for x in X])
    assert np.max(np.abs(Y - ref)) <= 1e-4 * max(1.0, float(np.max(np.abs(ref))))


def test_gradient_matches_oracle(gz, orc, device):
    p = _params(5)
    rng = np.random.default_rng(6)
    X = rng.normal(size=(1000, 8)).astype(np.float32)
    T = rng.normal(size=(1000, 2)).astype(np.float32)
    loss, g = gz.selector_gradient(p, X, T)
    rloss, rg = orc.mlp_gradient(p, X, T)
    assert abs(loss - rloss) <= 1e-4 * max(1.0, abs(rloss))
    # float atomics in a different order: compare against the gradient scale
    assert np.max(np.abs(g - rg)) <= 1e-4 * max(1e-3, float(np.max(np.abs(rg))))
    assert np.dot(g, rg) / (np.linalg.norm(g) * np.linalg.norm(rg)) > 0.99999


def test_training_fits_a_known_function(gz, device):
    rng = np.random.default_rng(8)
    X = rng.normal(size=(2000, 8)).astype(np.float32)
    T = np.stack([X[:, 0] - 0.5 * X[:, 1], np.maximum(X[:, 2], 0) + 0.1 * X[:, 3]], axis=1).astype(np.float32)
    import selector_train as st
    p, hist = gz.selector_train(st.init_params(1), X, T, 1500, 3e-3)
    assert hist[-1] < 0.05 * hist[0]
    Y = gz.selector_forward(p, X)
    assert float(np.mean((Y - T) ** 2)) == pytest.approx(float(hist[-1]), rel=0.2, abs=1e-3)


@pytest.fixture(scope="module")
def rmat13(gz):
    return gz.rmat_graph(13, 16, seed=17)


def _star_path(gz, leaves=20000, path=200, other=40000):
    """A path 0..path into the centre of a star with `leaves` leaves, plus a
    disconnected star with `other` leaves whose edges raise E/5 above the
    centre's frontier statistic (1 + leaves + 1), so that the hybrid pushes
    the centre's level."""
    pairs = [(i, i + 1) for i in range(path)] + [(path, path + 1 + j) for j in range(leaves)]
    c2 = path + 1 + leaves
    pairs += [(c2, c2 + 1 + j) for j in range(other)]
    return gz.make_graph(np.asarray(pairs, dtype=np.uint64), c2 + 1 + other, symmetrize=True)


@pytest.mark.parametrize("case", ["rmat13", "star_path", "rmat18"])
def test_probe_features_and_sequence(gz, orc, device, rmat13, case):
    # the star-and-path graph from the far end of the path, and R-MAT s18,
    # reach push levels whose frontier statistic lies between 4,096 and
    # 65,536: the small-frontier push with the overflow queue, whose queue
    # kernel writes its features after the bitmap kernel's wavefronts
    if case == "rmat13":
        g, root = rmat13, 3
    elif case == "star_path":
        g, root = _star_path(gz), 0
    else:
        # root 0 has 15,919 out-edges: its level and the fourth are pushes
        # with statistics of 15,920 and 30,459
        g, root = gz.rmat_graph(18, 16, seed=29), 0
    gz.load(g)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    gz.run(gz.options("bfs", engine="hybrid", root=root, profile=True))
    hybrid = [r.engine for r in gz.records()]
    gz.run(gz.options("bfs", engine="probe", root=root))
    recs = gz.records()
    p, d = gz.bfs_result()
    _, dep, _, _ = orc.bfs(G, orc.options(root=root))
    assert np.array_equal(d, dep)
    assert [r.engine for r in recs] == hybrid
    assert all(r.pull_ms > 0 and r.push_ms > 0 for r in recs)
    # features pinned by the depths: level i's frontier is depth i-1
    V, E = g.num_vertices, g.num_edges
    od = np.bincount(g.pull[:, 0].astype(np.int64), minlength=V)
    ind = np.bincount(g.pull[:, 1].astype(np.int64), minlength=V)
    reached = dep != np.iinfo(np.uint64).max
    for i, r in enumerate(recs):
        f = np.array(r.features, dtype=np.float64)
        level = reached & (dep == i)
        seen = reached & (dep <= i)
        exp = [np.log1p(level.sum()), np.log1p(od[level].sum()), np.log1p(V - seen.sum()),
               np.log1p(E - ind[seen].sum()), i + 1, 0 if i == 0 else hybrid[i - 1]]
        assert np.allclose(f[:6], exp, rtol=1e-5, atol=1e-5), (i, f, exp)
        assert f[7] == pytest.approx(E / V, rel=1e-6)
    if case != "rmat13":
        # at least one level went through the overflow queue's statistic range
        assert any(4096 < r.frontier_stat <= 65536 and r.engine == 2 for r in recs), case


@pytest.mark.parametrize("case", ["rmat13", "star_path"])
def test_wide_probe_times_every_variant(gz, orc, device, rmat13, case):
    # the wide probe times the uncompacted pull, the compacted push and the
    # uncompacted push from the same state as the two engines; the search
    # still follows the fixed threshold and its depths stay exact
    g, root = (rmat13, 3) if case == "rmat13" else (_star_path(gz), 0)
    gz.load(g)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    gz.run(gz.options("bfs", engine="hybrid", root=root, profile=True))
    hybrid = [r.engine for r in gz.records()]
    gz.probe_wide(True)
    try:
        gz.run(gz.options("bfs", engine="probe", root=root))
    finally:
        gz.probe_wide(False)
    recs = gz.records()
    _, dep, _, _ = orc.bfs(G, orc.options(root=root))
    assert np.array_equal(gz.bfs_result()[1], dep)
    assert [r.engine for r in recs] == hybrid
    assert all(r.pull_ms > 0 and r.push_ms > 0 and min(r.variant_ms) > 0 for r in recs)
    # without the switch no variant is timed
    gz.run(gz.options("bfs", engine="probe", root=root))
    assert all(max(r.variant_ms) == 0 for r in gz.records())


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="selector weights not trained yet")
def test_learned_engine_keeps_results_exact(gz, orc, device, rmat13):
    g = rmat13
    gz.selector_load(WEIGHTS)
    try:
        gz.load(g)
        G = orc.Graph(g.num_vertices, g.pull, g.push)
        for root in (0, 3, 99):
            gz.run(gz.options("bfs", engine="learned", root=root, profile=True))
            p, d = gz.bfs_result()
            _, dep, _, _ = orc.bfs(G, orc.options(root=root))
            assert np.array_equal(d, dep)
            assert all(r.engine in (1, 2) for r in gz.records())
        sym = gz.make_graph(gz.rmat_edges(12, 1 << 15, 5), 1 << 12, symmetrize=True)
        gz.load(sym)
        gz.run(gz.options("cc", engine="learned"))
        labels, _, _ = orc.cc(orc.Graph(sym.num_vertices, sym.pull, sym.push), orc.options())
        assert np.array_equal(gz.cc_labels(), labels)
    finally:
        gz.selector_clear()


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="selector weights not trained yet")
def test_unused_feature_skips_its_inputs(gz, orc, device, rmat13, tmp_path):
    # the committed model does not use feature 3 (std 0): the engine feeds it
    # 0 and skips the in-degree reads; a copy that marks it used (std 1)
    # reads them.  Depths stay exact either way, and with std 0 the model's
    # output cannot depend on the feature's value
    import selector_train as st
    params, mean, std = st.load_weights(WEIGHTS)
    assert std[3] == 0 and np.all(np.delete(std, 3) > 0)
    used = tmp_path / "used.bin"
    std_used = std.copy()
    std_used[3] = 1.0
    st.save_weights(str(used), params, mean, std_used)
    x = np.random.default_rng(4).normal(size=(16, 8)).astype(np.float32)
    x2 = x.copy()
    x2[:, 3] += 5.0
    assert np.array_equal(st.normalize(x, mean, std), st.normalize(x2, mean, std))
    g = rmat13
    gz.load(g)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    try:
        for path, flag in ((WEIGHTS, False), (str(used), True)):
            gz.selector_load(path)
            assert gz.selector_feature_used(3) == flag
            assert all(gz.selector_feature_used(i) for i in (0, 1, 2, 4, 5, 6, 7))
            for root in (0, 3):
                gz.run(gz.options("bfs", engine="learned", root=root))
                _, dep, _, _ = orc.bfs(G, orc.options(root=root))
                assert np.array_equal(gz.bfs_result()[1], dep)
    finally:
        gz.selector_clear()
    assert not gz.selector_feature_used(3)


@pytest.mark.parametrize("prefer", ["push", "pull"])
def test_trust_band_bounds_the_model(gz, orc, device, rmat13, tmp_path, prefer):
    # a model that always prefers one engine (zero weights, output biases
    # +-10): inside the band [E/5 / below, above * E/5] the engine follows
    # it, outside the fixed rule's choice stands (pull above, push below);
    # depths stay exact
    import selector_train as st
    params = np.zeros(gz.SEL_PARAMS, dtype=np.float32)
    params[-2:] = [10.0, -10.0] if prefer == "push" else [-10.0, 10.0]
    path = tmp_path / f"{prefer}.bin"
    st.save_weights(str(path), params, np.zeros(8, np.float32), np.ones(8, np.float32))
    g = rmat13
    gz.load(g)
    G = orc.Graph(g.num_vertices, g.pull, g.push)
    thr = g.num_edges // 5
    model = 2 if prefer == "push" else 1
    gz.selector_load(str(path))
    try:
        for above, below in ((gz.TRUST_ABOVE, gz.TRUST_BELOW), (4.0, 4.0), (1.5, 1.5), (1.0, 0.0), (0.0, 8.0),
                             (0.0, 0.0)):
            gz.selector_set_trust(above, below)
            for root in (0, 3):
                gz.run(gz.options("bfs", engine="learned", root=root, profile=True))
                _, dep, _, _ = orc.bfs(G, orc.options(root=root))
                assert np.array_equal(gz.bfs_result()[1], dep)
                for r in gz.records():
                    want = model
                    if above >= 1.0 and r.frontier_stat > above * thr:
                        want = 1
                    elif below >= 1.0 and r.frontier_stat * below < thr:
                        want = 2
                    assert r.engine == want, (above, below, root, r.iteration, r.frontier_stat, thr)
    finally:
        gz.selector_set_trust(gz.TRUST_ABOVE, gz.TRUST_BELOW)
        gz.selector_clear()


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="selector weights not trained yet")
def test_learned_gain_on_held_out_graphs(gz, device):
    """Probe both engines on graphs the network never saw and compare the
    edge-phase total of the fixed threshold with the network's choices (both
    read off the same measured per-level times)."""
    import selector_train as st
    params, mean, std = st.load_weights(WEIGHTS)
    data = st.collect(st.training_graphs(10, 90000), 2, log=lambda s: None)
    pt = st.policy_times(data, params, mean, std)
    print(f"held-out: fixed {pt['fixed']:.3f} ms, learned {pt['learned']:.3f} ms, per-level best {pt['best']:.3f} ms,"
          f" gain {pt['fixed'] / pt['learned']:.3f}x of at most {pt['fixed'] / pt['best']:.3f}x")
    assert pt["learned"] <= pt["fixed"]
    assert pt["learned"] <= 1.03 * pt["best"]
