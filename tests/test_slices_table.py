"""Chunk table of the source-sliced pull (host computation, no GPU).

Properties the kernel relies on: every chunk of every slice is in exactly one
slot, with its merge record numbered in slice order; a workgroup w only holds
chunks of slices x = w % 8, 8 + x, ... (level by level, in order) or of the
cold slice 8L; the XCDs get an even number of chunks whenever the cold slice can
fill them; four slots per workgroup and empty slots marked by count 0.
"""
import numpy as np
import pytest

import grazelle as gz

XCDS = 8


def _xcd_sequences(t):
    out = []
    for x in range(XCDS):
        seq = []
        for wg in range(x, len(t) // 4, XCDS):
            seq.extend(t[4 * wg:4 * wg + 4])
        out.append(seq)
    return out


def _check(vectors, L, spread):
    t = gz.slices_chunk_table(vectors, L, spread)
    T = (len(vectors) - 1) // 8
    cold_slice = 8 * T
    vectors = np.asarray(vectors, dtype=np.int64)
    chunks = (vectors + L - 1) // L
    base = np.concatenate([[0], np.cumsum(vectors)])
    rec = np.concatenate([[0], np.cumsum(chunks)])
    assert len(t) % (4 * XCDS) == 0
    used = t[t["count"] > 0]
    assert len(used) == chunks.sum()
    sl = (used["record"] >> gz.SLICE_RECORD_BITS).astype(np.int64)
    r = (used["record"] & ((1 << gz.SLICE_RECORD_BITS) - 1)).astype(np.int64)
    assert np.array_equal(np.sort(r), np.arange(chunks.sum()))
    for x in range(len(vectors)):
        m = sl == x
        k = r[m] - rec[x]
        assert np.all((k >= 0) & (k < chunks[x]))
        assert np.array_equal(used["begin"][m].astype(np.int64), base[x] + k * L)
        want = np.minimum(L, vectors[x] - k * L)
        assert np.array_equal(used["count"][m].astype(np.int64), want)
    # workgroup -> XCD round robin; an XCD holds its own slices in every tier
    slot = np.flatnonzero(t["count"] > 0)
    xcd = (slot // 4) % XCDS
    assert np.all(((sl % XCDS == xcd) & (sl < cold_slice)) | (sl == cold_slice))
    # per-XCD chunk counts
    per = np.bincount(xcd, minlength=XCDS)
    own = np.array([sum(int(chunks[8 * k + x]) for k in range(T)) for x in range(XCDS)])
    target = max(int(-(-chunks.sum() // XCDS)), int(own.max(initial=0)))
    assert per.max() <= target
    if own.max(initial=0) <= -(-chunks.sum() // XCDS):
        assert per.max() - per.min() <= 1
    for x, seq in enumerate(_xcd_sequences(t)):
        used = np.array([c["count"] > 0 for c in seq], bool)
        n = int(used.sum())
        # a XCD's chunks occupy a prefix of its own slot sequence
        assert used[:n].all() and not used[n:].any()
        slices = np.array([int(c["record"] >> gz.SLICE_RECORD_BITS) for c in seq[:n]], np.int64)
        hot = slices[slices != cold_slice]
        # tier by tier, each slice's chunks in order
        assert np.all(np.diff(hot) >= 0)
        if not spread:
            cold = slices == cold_slice
            if cold.any():
                assert cold[int(np.argmax(cold)):].all()
    return t


@pytest.mark.parametrize("tiers", [1, 2, 3, 5, 7])
@pytest.mark.parametrize("spread", [True, False])
@pytest.mark.parametrize("L", [64, 512])
def test_table_random(spread, L, tiers):
    rng = np.random.default_rng(5 + tiers)
    n = 8 * tiers + 1
    for _ in range(40):
        v = rng.integers(0, 40 * L, size=n)
        v[rng.random(n) < 0.2] = 0
        _check(v, L, spread)


def test_table_rejects_bad_sizes():
    with pytest.raises(ValueError):
        gz.slices_chunk_table([1] * 10, 64)
    assert len(gz.slices_chunk_table([1] * 65, 64)) == 0  # eight levels are not supported
    assert len(gz.slices_chunk_table([1] * 57, 64)) > 0


@pytest.mark.parametrize("spread", [True, False])
def test_table_edge_cases_tiers(spread):
    _check([0] * 17, 64, spread)
    _check([64 * 3] * 8 + [0] * 8 + [100], 64, spread)
    _check([0] * 8 + [64 * 3] * 8 + [100], 64, spread)
    _check([64] * 24 + [64 * 40], 64, spread)


@pytest.mark.parametrize("spread", [True, False])
def test_table_edge_cases(spread):
    _check([0] * 9, 64, spread)
    _check([0] * 8 + [1000], 64, spread)
    _check([1000] + [0] * 8, 64, spread)
    _check([1] * 9, 64, spread)
    _check([64 * 50] * 8 + [0], 64, spread)
    t = _check([64 * 4] * 8 + [0], 64, spread)
    # eight full workgroups, one per XCD
    assert len(t) == 32 and np.all(t["count"] == 64)


def test_table_spread_interleaves():
    t = gz.slices_chunk_table([64 * 30] * 8 + [64 * 80], 64, True)
    for x in range(XCDS):
        own = []
        for wg in range(x, len(t) // 4, XCDS):
            own.extend(int(c["record"] >> gz.SLICE_RECORD_BITS) for c in t[4 * wg:4 * wg + 4] if c["count"])
        cold = np.flatnonzero(np.array(own) == XCDS)
        assert len(cold) == 10
        # spread: the cold chunks are not bunched at either end
        assert cold[0] < 5 and cold[-1] > len(own) - 5


def test_hot_cache_sizes():
    with pytest.raises(ValueError):
        gz.slices_set_hot_cache(1000)
    for v in (4096, 8192, 16384, 20480, 0):
        gz.slices_set_hot_cache(v)
