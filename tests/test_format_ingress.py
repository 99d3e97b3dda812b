"""Vector-Sparse format, packer, vertex index and edge-list files (CPU).

The packer of the product (libgrazelle.so) and of the oracle are checked
against each other bit for bit, and both against what the format fixes
independently: decoding recovers the grouped edge list, the vector count is
sum over groups of ceil(degree/4), and the index/offsets agree with a brute
force over the decoded vectors.
"""
import os

import numpy as np
import pytest

import graphs

M48 = np.uint64((1 << 48) - 1)


def decode(vectors):
    """Decode vectors with the bit layout written out in the format spec."""
    v = vectors.astype(np.uint64)
    lanes = [v[:, k] for k in range(4)]
    shared = ((lanes[0] >> np.uint64(48)) & np.uint64(0x7FFF)) \
        | (((lanes[1] >> np.uint64(48)) & np.uint64(0x7FFF)) << np.uint64(15)) \
        | (((lanes[2] >> np.uint64(48)) & np.uint64(0x7FFF)) << np.uint64(30)) \
        | (((lanes[3] >> np.uint64(48)) & np.uint64(0x7)) << np.uint64(45))
    out = []
    for k in range(4):
        valid = (lanes[k] >> np.uint64(63)) == 1
        out.append((np.nonzero(valid)[0], shared[valid], lanes[k][valid] & M48))
    return shared, out


def edges_from_vectors(vectors, by_dst):
    _, lanes = decode(vectors)
    rows = []
    for k, (idx, sh, nb) in enumerate(lanes):
        rows.append(np.stack([idx, np.full(idx.shape, k), sh, nb], axis=1))
    r = np.concatenate(rows)
    r = r[np.lexsort((r[:, 1], r[:, 0]))]
    sh, nb = r[:, 2], r[:, 3]
    return np.stack([nb, sh], axis=1) if by_dst else np.stack([sh, nb], axis=1)


@pytest.fixture(scope="module")
def g(gz):
    return gz.rmat_graph(12, 16, seed=21)


@pytest.mark.parametrize("by_dst", [True, False])
def test_pack_roundtrip_and_parity(gz, orc, g, by_dst):
    pairs = g.pull if by_dst else g.push
    v = gz.pack_vectors(pairs, by_dst)
    assert np.array_equal(v, orc.pack(pairs, by_dst))
    assert np.array_equal(edges_from_vectors(v, by_dst).astype(np.uint64), pairs)
    key = pairs[:, 1] if by_dst else pairs[:, 0]
    _, counts = np.unique(key, return_counts=True)
    assert v.shape[0] == int(np.sum((counts + 3) // 4))
    assert gz.lib.gz_count_vectors(gz._ptr(np.ascontiguousarray(pairs)), pairs.shape[0], int(by_dst)) == v.shape[0]
    # valid lanes form a prefix in every vector
    valid = (v >> np.uint64(63)).astype(np.int64)
    assert np.all(np.diff(valid, axis=1) <= 0)


def test_packing_efficiency_rule(gz):
    # degrees 1..9 for consecutive destinations: vectors per group ceil(d/4)
    pairs = []
    for d in range(1, 10):
        pairs += [(s, d) for s in range(100, 100 + d)]
    el = graphs.edge_lists(gz, pairs, 200)
    v = gz.pack_vectors(el.pull, True)
    assert v.shape[0] == sum((d + 3) // 4 for d in range(1, 10))


@pytest.mark.parametrize("vid", [0, 1, 0x7FFF, 0x8000, (1 << 30) - 1, 1 << 30, (1 << 45) - 1, 1 << 45,
                                 (1 << 48) - 1, 0xA5A5A5A5A5A5])
def test_shared_vertex_pieces(gz, orc, vid):
    pairs = np.array([[5, vid], [7, vid], [(1 << 48) - 2, vid]], dtype=np.uint64)
    v = gz.pack_vectors(pairs, True)
    shared, lanes = decode(v)
    assert int(shared[0]) == vid
    assert int(orc.lib.orc_shared_vertex(v[0:1].ctypes.data_as(orc.vp))) == vid
    assert [int(x) for x in lanes[0][2]] == [5] and [int(x) for x in lanes[2][2]] == [(1 << 48) - 2]
    assert int(v[0, 3]) >> 63 == 0  # lane 3 invalid


def test_vertex_index_and_offsets(gz, orc, g):
    V = g.num_vertices
    for by_dst, pairs in ((True, g.pull), (False, g.push)):
        v = gz.pack_vectors(pairs, by_dst)
        shared, _ = decode(v)
        idx, s, e = gz.vertex_index(v, V)
        oidx, os_, oe = orc.vertex_index(v, V)
        assert np.array_equal(idx, oidx) and (s, e) == (os_, oe)
        first = {}
        for i, x in enumerate(shared.tolist()):
            first.setdefault(x, i)
        lo, hi = int(shared[0]), int(shared[-1])
        assert (s, e) == (lo, hi)
        for x in range(V + 64):
            if x in first:
                want = first[x]
            elif x <= hi:
                want = 0x7FFFFFFFFFFFFFFF
            else:
                want = 0xFFFFFFFFFFFFFFFF
            assert int(idx[x]) == want
        off = gz.vertex_offsets(v, V)
        for x in range(V):
            seg = shared[int(off[x]):int(off[x + 1])]
            assert np.all(seg == x)
            assert seg.shape[0] == (int(np.sum(shared == x)))


def test_outdegrees(gz, orc, g):
    od = orc.outdegrees(g.pull, g.num_vertices)
    want = np.bincount(g.pull[:, 0].astype(np.int64), minlength=g.num_vertices + 8).astype(np.float64)
    assert np.array_equal(od, want)


def test_sort_and_dedupe(gz):
    rng = np.random.default_rng(0)
    pairs = rng.integers(0, 50, size=(3000, 2), dtype=np.uint64)
    for by_dst in (True, False):
        got = gz.sort_edges(pairs, by_dst=by_dst)
        keep = pairs[pairs[:, 0] != pairs[:, 1]]
        u = np.unique(keep, axis=0)
        order = np.lexsort((u[:, 0], u[:, 1])) if by_dst else np.lexsort((u[:, 1], u[:, 0]))
        assert np.array_equal(got, u[order])
    big = np.array([[1 << 40, 3], [2, 1 << 40], [2, 1 << 40]], dtype=np.uint64)
    got = gz.sort_edges(big, by_dst=False)
    assert got.tolist() == [[2, 1 << 40], [1 << 40, 3]]


def test_edge_list_files(gz, tmp_path, g):
    import ctypes as C
    path = str(tmp_path / "x-pull")
    gz.write_edge_list(path, g.num_vertices, g.pull)
    raw = np.fromfile(path, dtype="<u8")
    assert raw[0] == g.num_vertices and raw[1] == g.num_edges
    assert np.array_equal(raw[2:].reshape(-1, 2), g.pull)
    V, E, P = C.c_uint64(), C.c_uint64(), C.POINTER(C.c_uint64)()
    assert gz.lib.gz_edge_list_read(path.encode(), C.byref(V), C.byref(E), C.byref(P)) == 0
    assert (V.value, E.value) == (g.num_vertices, g.num_edges)
    assert gz.lib.gz_edge_list_read(str(tmp_path / "missing").encode(), C.byref(V), C.byref(E), C.byref(P)) == -1
    raw[:1].tofile(str(tmp_path / "short"))
    assert gz.lib.gz_edge_list_read(str(tmp_path / "short").encode(), C.byref(V), C.byref(E), C.byref(P)) == -2
    raw[:10].tofile(str(tmp_path / "trunc"))
    assert gz.lib.gz_edge_list_read(str(tmp_path / "trunc").encode(), C.byref(V), C.byref(E), C.byref(P)) == -3


def test_rmat_generator(gz):
    a = gz.rmat_edges(16, 200000, seed=5)
    b = gz.rmat_edges(16, 200000, seed=5)
    c = gz.rmat_edges(16, 200000, seed=6)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # drawing a later slice of the same stream reproduces the tail
    t = gz.rmat_edges(16, 1000, seed=5, edge_offset=199000)
    assert np.array_equal(t, a[199000:])
    # top-level quadrant frequencies follow a/b/c/d = 0.57/0.19/0.19/0.05
    top_s = (a[:, 0] >> np.uint64(15)).astype(np.int64)
    top_d = (a[:, 1] >> np.uint64(15)).astype(np.int64)
    freq = [np.mean((top_s == i) & (top_d == j)) for i, j in ((0, 0), (0, 1), (1, 0), (1, 1))]
    assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.01)
    assert a.max() < (1 << 16)
    # modulo fold into a non power-of-two vertex count
    f = gz.rmat_edges(16, 10000, seed=5, num_vertices=50000)
    assert f.max() < 50000
