"""CPU emulation of the wavefront index arithmetic used by the HIP kernels.

The pull kernels split the vector list into chunks of `chunk` vectors, walk a
chunk 64 vectors per step, combine runs of equal destinations with a
segmented scan, carry the open run across steps, store completed runs and
leave the chunk's last run in a merge-buffer record; the frontier kernels map
"slots" of a compacted active-vertex list back to vectors.  These tests replay
that arithmetic step by step in numpy, assert every index stays inside its
buffer, and compare the outcome with the oracle and with brute force.  They
exercise the runs that end exactly on a step boundary and runs that span
several chunks.
"""
import numpy as np
import pytest

import graphs

WAVE = 64


def shared_of(vectors):
    v = vectors.astype(np.uint64)
    return (((v[:, 0] >> np.uint64(48)) & np.uint64(0x7FFF))
            | (((v[:, 1] >> np.uint64(48)) & np.uint64(0x7FFF)) << np.uint64(15))
            | (((v[:, 2] >> np.uint64(48)) & np.uint64(0x7FFF)) << np.uint64(30))
            | (((v[:, 3] >> np.uint64(48)) & np.uint64(0x7)) << np.uint64(45))).astype(np.int64)


def emulate_pull_sum(vectors, props, chunk, V):
    """Replay of gz_k_edge_pull_pr<true>: returns acc (direct stores only),
    merge records and the S partials."""
    n = vectors.shape[0]
    d_all = shared_of(vectors)
    valid = (vectors >> np.uint64(63)).astype(bool)
    ids = (vectors & np.uint64((1 << 48) - 1)).astype(np.int64)
    gv = np.where(valid, props[np.where(valid, ids, 0)], 0.0)
    g_all = (gv[:, 2] + gv[:, 3]) + (gv[:, 0] + gv[:, 1])
    acc = np.zeros(V)
    written = np.zeros(V, dtype=np.int64)
    nchunks = (n + chunk - 1) // chunk
    merge = []
    for c in range(nchunks):
        b, e = c * chunk, min(c * chunk + chunk, n)
        assert 0 <= b < n
        first_d = d_all[b]
        carry_d, carry_v = first_d, 0.0
        for base in range(b, e, WAVE):
            lanes = np.arange(WAVE)
            i = base + lanes
            ok = i < e
            d = np.where(ok, d_all[np.minimum(i, n - 1)], -1)
            g = np.where(ok, g_all[np.minimum(i, n - 1)], 0.0)
            if base != b and d[0] != carry_d:
                acc[carry_d] = carry_v
                written[carry_d] += 1
            heads = np.ones(WAVE, dtype=bool)
            heads[1:] = d[1:] != d[:-1]
            s = np.zeros(WAVE)
            run = 0.0
            for L in range(WAVE):
                run = g[L] if heads[L] else run + g[L]
                s[L] = run
            s = np.where(d == carry_d, s + carry_v, s)
            last = min(WAVE - 1, e - 1 - base)
            for L in range(WAVE):
                seg_end = L == WAVE - 1 or heads[L + 1]
                if ok[L] and seg_end and L != last:
                    assert 0 <= d[L] < V
                    acc[d[L]] = s[L]
                    written[d[L]] += 1
            carry_d, carry_v = d[last], s[last]
        merge.append((first_d, carry_d, carry_v, 0))
    return acc, written, np.array(merge, dtype=[("initial", "<u8"), ("final", "<u8"), ("partial", "<f8"),
                                                ("pad", "<u8")])


@pytest.mark.parametrize("chunk", [64, 128, 512])
def test_pull_chunk_emulation_matches_dense(gz, orc, chunk):
    g = graphs.hub_graph(gz, num_vertices=700, hub_in=2500, seed=chunk)
    V = g.num_vertices
    vecs = orc.pack(g.pull, True)
    rng = np.random.default_rng(1)
    props = rng.random(V + 8)
    acc, written, merge = emulate_pull_sum(vecs, props, chunk, V)
    # every destination with in-edges is either stored directly or finalised
    # by exactly one merge run head
    orc.merge(merge, acc)
    dense = np.zeros(V)
    np.add.at(dense, g.pull[:, 1].astype(np.int64), props[g.pull[:, 0].astype(np.int64)])
    has_in = np.bincount(g.pull[:, 1].astype(np.int64), minlength=V) > 0
    assert np.allclose(acc[has_in], dense[has_in], rtol=1e-13, atol=0)
    assert np.all(written[~has_in] == 0)


def test_pull_emulation_boundary_runs(gz, orc):
    """Runs ending exactly at lane 63 of a step (the case a lost carry breaks)."""
    V = 300
    pairs = []
    for dst in range(1, 200):
        pairs += [(s, dst) for s in range(200, 204)]  # one full vector per destination
    el = graphs.edge_lists(gz, pairs, V)
    vecs = orc.pack(el.pull, True)
    props = np.arange(V + 8, dtype=np.float64)
    acc, _, merge = emulate_pull_sum(vecs, props, 512, V)
    orc.merge(merge, acc)
    for dst in range(1, 200):
        assert acc[dst] == 200 + 201 + 202 + 203


def emulate_walk(vectors, props, chunk, V, vpl):
    """Replay of gz_walk_chunk (gz_walk.h) with the PageRank sum op: lanes hold
    `vpl` consecutive vectors; lane-local runs, a segmented scan of lane tails,
    head fix-ups, and the carry between steps.  Returns acc after the direct
    stores plus the merge records."""
    n = vectors.shape[0]
    d_all = shared_of(vectors)
    valid = (vectors >> np.uint64(63)).astype(bool)
    ids = (vectors & np.uint64((1 << 48) - 1)).astype(np.int64)
    gv = np.where(valid, props[np.where(valid, ids, 0)], 0.0)
    g_all = (gv[:, 2] + gv[:, 3]) + (gv[:, 0] + gv[:, 1])
    acc = np.zeros(V)
    stores = np.zeros(V, dtype=np.int64)

    def store(d, s):
        assert 0 <= d < V
        acc[d] = s
        stores[d] += 1

    step = WAVE * vpl
    merge = []
    for c in range((n + chunk - 1) // chunk):
        b, e = c * chunk, min(c * chunk + chunk, n)
        first_d = d_all[b]
        carry_d, carry_v = first_d, 0.0
        for base in range(b, e, step):
            h_d = [-1] * WAVE; t_d = [-1] * WAVE; h_s = [0.0] * WAVE; t_s = [0.0] * WAVE
            has = [False] * WAVE
            for L in range(WAVE):
                nvalid = 0
                for k in range(vpl):
                    i = base + L * vpl + k
                    if i >= e:
                        break
                    d, g = d_all[i], g_all[i]
                    if nvalid == 0:
                        h_d[L] = t_d[L] = d; h_s[L] = t_s[L] = g
                    elif d == t_d[L]:
                        t_s[L] += g
                        if t_d[L] == h_d[L]:
                            h_s[L] = t_s[L]
                    else:
                        if t_d[L] != h_d[L]:
                            store(t_d[L], t_s[L])
                        t_d[L], t_s[L] = d, g
                    nvalid += 1
                has[L] = nvalid > 0
            if base != b and h_d[0] != carry_d:
                store(carry_d, carry_v)
            S = [0.0] * WAVE
            seg_start = 0
            for L in range(WAVE):
                if L == 0 or t_d[L] != t_d[L - 1]:
                    seg_start = L
                    S[L] = t_s[L]
                else:
                    S[L] = S[L - 1] + t_s[L]
                if seg_start == 0 and t_d[L] == carry_d and has[L]:
                    pass
            # carry into the first segment
            for L in range(WAVE):
                first_seg = all(t_d[k] == t_d[0] for k in range(L + 1))
                if first_seg and t_d[L] == carry_d and has[L]:
                    S[L] += carry_v
            last = sum(has) - 1
            for L in range(WAVE):
                if not has[L]:
                    continue
                single = h_d[L] == t_d[L]
                if not single:
                    head = h_s[L]
                    if L == 0:
                        if h_d[0] == carry_d:
                            head += carry_v
                    elif t_d[L - 1] == h_d[L]:
                        head += S[L - 1]
                    store(h_d[L], head)
                next_h = h_d[(L + 1) % WAVE]
                if L != last and next_h != t_d[L]:
                    store(t_d[L], S[L])
            carry_d, carry_v = t_d[last], S[last]
        merge.append((first_d, carry_d, carry_v, 0))
    return acc, stores, np.array(merge, dtype=[("initial", "<u8"), ("final", "<u8"), ("partial", "<f8"),
                                                ("pad", "<u8")])


@pytest.mark.parametrize("vpl", [1, 2, 4])
@pytest.mark.parametrize("chunk", [128, 512, 1024])
def test_walker_emulation_matches_dense(gz, orc, vpl, chunk):
    g = graphs.hub_graph(gz, num_vertices=600, hub_in=2000, seed=vpl * 7 + chunk)
    V = g.num_vertices
    vecs = orc.pack(g.pull, True)
    props = np.random.default_rng(3).random(V + 8)
    acc, stores, merge = emulate_walk(vecs, props, chunk, V, vpl)
    orc.merge(merge, acc)
    dense = np.zeros(V)
    np.add.at(dense, g.pull[:, 1].astype(np.int64), props[g.pull[:, 0].astype(np.int64)])
    has_in = np.bincount(g.pull[:, 1].astype(np.int64), minlength=V) > 0
    assert np.allclose(acc[has_in], dense[has_in], rtol=1e-13, atol=0)
    assert np.all(stores[~has_in] == 0)
    # every destination is stored at most once per chunk it touches
    assert stores.max() <= 1 + (vecs.shape[0] + chunk - 1) // chunk


@pytest.mark.parametrize("vpl", [2, 4])
def test_walker_emulation_boundary_shapes(gz, orc, vpl):
    """Runs of 1..9 vectors so that heads, tails, single-run lanes and step
    boundaries all occur."""
    V = 2000
    rng = np.random.default_rng(vpl)
    pairs, dst = [], 1
    while dst < 1500:
        deg = int(rng.integers(1, 37))
        pairs += [(1500 + s, dst) for s in range(deg)]
        dst += 1
    el = graphs.edge_lists(gz, pairs, V)
    vecs = orc.pack(el.pull, True)
    props = np.arange(V + 8, dtype=np.float64)
    for chunk in (64 * vpl, 256, 768):
        acc, _, merge = emulate_walk(vecs, props, chunk, V, vpl)
        orc.merge(merge, acc)
        dense = np.zeros(V)
        np.add.at(dense, el.pull[:, 1].astype(np.int64), props[el.pull[:, 0].astype(np.int64)])
        assert np.array_equal(acc[1:1500], dense[1:1500])


def compact(bitmap_bits, offsets):
    """Replay of gz_frontier_compact: active vertices owning vectors and the
    exclusive prefix of their vector counts."""
    V = offsets.shape[0] - 1
    act = [v for v in range(V) if bitmap_bits[v] and offsets[v + 1] > offsets[v]]
    counts = np.array([offsets[v + 1] - offsets[v] for v in act], dtype=np.int64)
    prefix = np.zeros(len(act) + 1, dtype=np.int64)
    prefix[1:] = np.cumsum(counts)
    return np.array(act, dtype=np.int64), prefix


def owner_search(prefix, n, j):
    lo, hi = 0, n
    while hi - lo > 1:
        mid = lo + (hi - lo) // 2
        if prefix[mid] <= j:
            lo = mid
        else:
            hi = mid
    return lo


def active_owner(prefix, n, a, base):
    starts = 0
    for lane in range(WAVE):
        k = a + 1 + lane
        if k <= n:
            r = int(prefix[k]) - base
            assert r >= 0
            if r < WAVE:
                starts |= 1 << r
    return [a + bin(starts & ((2 << lane) - 1)).count("1") for lane in range(WAVE)]


@pytest.mark.parametrize("chunk", [64, 256])
@pytest.mark.parametrize("density", [0.02, 0.5, 1.0])
def test_active_slot_mapping(gz, orc, chunk, density):
    g = graphs.hub_graph(gz, num_vertices=900, hub_in=4000, seed=7)
    V = g.num_vertices
    vecs = orc.pack(g.push, False)
    off = gz.vertex_offsets(vecs, V).astype(np.int64)
    rng = np.random.default_rng(int(density * 100))
    bits = rng.random(V) < density
    act, prefix = compact(bits, off)
    n, T = len(act), int(prefix[-1])
    want = np.concatenate([np.arange(off[v], off[v + 1]) for v in act]) if n else np.zeros(0, np.int64)
    seen = np.full(T, -1, dtype=np.int64)
    for c in range((T + chunk - 1) // chunk):
        j0, jend = c * chunk, min(c * chunk + chunk, T)
        a = owner_search(prefix, n, j0)
        for base in range(j0, jend, WAVE):
            owners = active_owner(prefix, n, a, base)
            for lane in range(WAVE):
                j = base + lane
                if j < jend:
                    o = owners[lane]
                    assert 0 <= o < n and prefix[o] <= j < prefix[o + 1]
                    vi = off[act[o]] + (j - prefix[o])
                    assert 0 <= vi < vecs.shape[0]
                    seen[j] = vi
            a = owners[WAVE - 1]
    assert np.array_equal(seen, want)


def emulate_compact_blocks(bitmap_bits, offsets, vpb=4096, block=256):
    """Replay of the three compaction kernels' index arithmetic: a block
    covers vpb consecutive vertices, vpb / block per thread; per-block counts, an
    exclusive scan of the block counts, then each thread's exclusive offset
    from a block scan of the per-thread counts, written in vertex order."""
    V = offsets.shape[0] - 1
    blocks = (V + vpb - 1) // vpb
    per = vpb // block

    def quad(v0):
        c = [0] * per
        if v0 >= V:
            return c
        for k in range(per):
            v = v0 + k
            if v < V and bitmap_bits[v]:
                c[k] = int(offsets[min(v + 1, V)] - offsets[min(v, V)])
        return c

    quads = [[quad(b * vpb + per * t) for t in range(block)] for b in range(blocks)]
    sums = [(sum(sum(1 for x in q if x) for q in qs), sum(sum(q) for q in qs)) for qs in quads]
    base_v = np.concatenate([[0], np.cumsum([s[0] for s in sums])]).astype(np.int64)
    base_c = np.concatenate([[0], np.cumsum([s[1] for s in sums])]).astype(np.int64)
    n = int(base_v[-1])
    lst = np.full(n, -1, dtype=np.int64)
    prefix = np.full(n + 1, -1, dtype=np.int64)
    prefix[n] = base_c[-1]  # the scan kernel's sentinel
    for b, qs in enumerate(quads):
        na = [sum(1 for x in q if x) for q in qs]
        nc = [sum(q) for q in qs]
        ea = np.concatenate([[0], np.cumsum(na)])
        ec = np.concatenate([[0], np.cumsum(nc)])
        for t, q in enumerate(qs):
            pos, vpos = int(base_v[b] + ea[t]), int(base_c[b] + ec[t])
            for k, c in enumerate(q):
                if c:
                    assert lst[pos] == -1
                    lst[pos] = b * vpb + per * t + k
                    prefix[pos] = vpos
                    pos += 1
                    vpos += c
    return lst, prefix


@pytest.mark.parametrize("V,density", [(1, 1.0), (700, 0.3), (1024, 1.0), (2500, 0.05), (3333, 0.9)])
def test_compaction_blocks_match_definition(gz, orc, V, density):
    rng = np.random.default_rng(V)
    pairs = rng.integers(0, V, size=(3 * V, 2), dtype=np.uint64)
    el = graphs.edge_lists(gz, pairs, V)
    off = gz.vertex_offsets(orc.pack(el.push, False), V).astype(np.int64)
    bits = rng.random(V) < density
    act, prefix = compact(bits, off)
    lst, pre = emulate_compact_blocks(bits, off)
    assert np.array_equal(lst, act)
    assert np.array_equal(pre, prefix)
