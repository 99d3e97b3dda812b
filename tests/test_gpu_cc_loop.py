"""CC with the engine chosen on the device and its iterations replayed from
a captured HIP graph (execution_cc.c:124-142 decides on the host).

A configuration runs eagerly the first time, is captured the second time in
a row and replayed after that; every run's labels must equal the oracle's
(bit-exact), and the captured runs must report consistent iteration counts
under the same engine rule.
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def device_loop(gz, device):
    gz.cc_graph_replay(True)
    yield
    gz.cc_graph_replay(False)


def _cases(gz):
    P, L = 200, 20000
    star = [(i, i + 1) for i in range(P)] + [(P, P + 1 + j) for j in range(L)]
    return {
        "rmat14_directed": gz.rmat_graph(14, 16, seed=11),
        "planted_trees": graphs.planted_trees(gz, base_scale=12),
        "star_path": gz.make_graph(np.asarray(star, dtype=np.uint64), P + 1 + L, symmetrize=True),
    }


@pytest.mark.parametrize("case", ["rmat14_directed", "planted_trees", "star_path"])
@pytest.mark.parametrize("engine", ["hybrid", "pull", "push"])
def test_device_loop_labels(gz, orc, device, case, engine):
    g = _cases(gz)[case]
    gz.load(g)
    want, _, _ = orc.cc(orc.Graph(g.num_vertices, g.pull, g.push), orc.options(engine=engine))
    c0, l0 = gz.graph_replay_counts()
    runs = []
    for _ in range(4):  # eager, eager (allocations settle), captured, replayed
        r = gz.run(gz.options("cc", engine=engine))
        assert np.array_equal(gz.cc_labels(), want), case
        assert r.pull_iterations + r.push_iterations == r.iterations > 0
        if engine == "pull":
            assert r.push_iterations == 0
        if engine == "push":
            assert r.pull_iterations == 0
        runs.append(r)
    c1, l1 = gz.graph_replay_counts()
    assert c1 - c0 >= 1 and l1 - l0 >= 2, "the CC loop was never captured"


def test_device_loop_iteration_cap(gz, orc, device):
    g = graphs.planted_trees(gz, base_scale=12)
    gz.load(g)
    full = None
    for _ in range(3):
        full = gz.run(gz.options("cc"))
    assert full.iterations > 3
    for _ in range(3):
        r = gz.run(gz.options("cc", max_iterations=3))
        assert r.iterations == 3


def test_device_loop_matches_host_loop(gz, orc, device):
    # the same graph with the device loop switched off: the same labels
    g = gz.rmat_graph(13, 16, seed=8)
    gz.load(g)
    gz.cc_graph_replay(False)
    host = gz.run(gz.options("cc"))
    want = gz.cc_labels()
    gz.cc_graph_replay(True)
    for _ in range(3):
        r = gz.run(gz.options("cc"))
        assert np.array_equal(gz.cc_labels(), want)
    assert r.iterations > 0 and host.iterations > 0
