"""Seeded test graphs shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np


def edge_lists(gz, pairs, num_vertices, symmetrize=False):
    pairs = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    return gz.make_graph(pairs, num_vertices, symmetrize=symmetrize)


def hub_graph(gz, num_vertices=3000, hub_in=9000, seed=3):
    """Vertex 1 receives edges from almost every vertex (its vectors span many
    scheduling units); the rest is a sparse random graph with sinks and
    sources."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, num_vertices, size=hub_in, dtype=np.uint64)
    hub = np.stack([src, np.full(hub_in, 1, dtype=np.uint64)], axis=1)
    rest = rng.integers(0, num_vertices, size=(4 * num_vertices, 2), dtype=np.uint64)
    return gz.make_graph(np.concatenate([hub, rest]), num_vertices)


def planted_trees(gz, base_scale=10, edge_factor=8, trees=8, tree_size=100, seed=5):
    """Symmetrised R-MAT plus disjoint random trees with ids after 2^base_scale
    (the shape of the connected-components configuration, scaled down)."""
    nv0 = 1 << base_scale
    rm = gz.rmat_edges(base_scale, nv0 * edge_factor, seed)
    rng = np.random.default_rng(seed)
    extra = []
    for t in range(trees):
        base = nv0 + t * tree_size
        for k in range(1, tree_size):
            p = int(rng.integers(0, k))
            extra.append((base + k, base + p))
    pairs = np.concatenate([rm, np.asarray(extra, dtype=np.uint64)])
    return gz.make_graph(pairs, nv0 + trees * tree_size, symmetrize=True)


def small_cases(gz):
    """Hand-made edge cases: (name, EdgeLists)."""
    cases = []
    cases.append(("single_edge", edge_lists(gz, [(0, 1)], 2)))
    cases.append(("two_cycle", edge_lists(gz, [(0, 1), (1, 0)], 2)))
    cases.append(("path", edge_lists(gz, [(i, i + 1) for i in range(99)], 100)))
    cases.append(("star_in", edge_lists(gz, [(i, 0) for i in range(1, 300)], 300)))
    cases.append(("star_out", edge_lists(gz, [(0, i) for i in range(1, 300)], 300)))
    cases.append(("isolated_tail", edge_lists(gz, [(0, 1), (1, 2), (2, 0), (5, 6)], 200)))
    return cases
