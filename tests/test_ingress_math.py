"""CPU emulation of the device ingress index arithmetic (gz_ingress.hip).

The device pipeline never runs here; this restates its per-thread formulas
in numpy on the same keys and checks them against the host packer, so an
index mistake (wrong lane, wrong vector, missed padding lane) shows on the
CPU before a kernel runs:
  keys        = dst << b | src (pull), sorted, duplicates and self-loops dropped
  eoff[v]     = lower_bound(keys, v << b), eoff[V] = m
  voff        = exclusive scan of ceil((eoff[v+1] - eoff[v]) / 4)
  edge i      -> vector voff[g] + r // 4, lane r % 4, r = i - eoff[g]
  padding     -> lanes [min(4, deg - r), 4) of the vector whose lane 0 is i
"""
import numpy as np
import pytest

VALID = np.uint64(1 << 63)


def _bits(nv):
    b = 1
    while b < 64 and (nv - 1) >> b:
        b += 1
    return b


def _piece(g, l):
    mask = 0x7 if l == 3 else 0x7FFF
    return ((g >> np.uint64(15 * l)) & np.uint64(mask)) << np.uint64(48)


def emulate(pairs, nv, by_dst=True):
    b = np.uint64(_bits(nv))
    s, d = pairs[:, 0].astype(np.uint64), pairs[:, 1].astype(np.uint64)
    g, o = (d, s) if by_dst else (s, d)
    keys = np.sort((g << b) | o)
    keep = np.ones(keys.shape[0], bool)
    keep[1:] = keys[1:] != keys[:-1]
    keep &= (keys >> b) != (keys & ((np.uint64(1) << b) - np.uint64(1)))
    keys = keys[keep]
    m = keys.shape[0]
    gg = keys >> b

    # eoff and voff looked up for the groups that occur (the kernels index
    # the full [V+1] arrays; only these entries are read by the pack)
    def eoff(v):
        return np.searchsorted(keys, v << b, side="left").astype(np.uint64)

    groups = np.unique(gg)
    deg = eoff(groups + np.uint64(1)) - eoff(groups)
    cnt = (deg + np.uint64(3)) // np.uint64(4)
    first_vec = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.uint64)
    total = int(cnt.sum())
    lanes = np.full(total * 4, 0xDEAD, dtype=np.uint64)
    gi = np.searchsorted(groups, gg)
    i = np.arange(m, dtype=np.uint64)
    r = i - eoff(gg)
    j = first_vec[gi] + r // np.uint64(4)
    l = (r & np.uint64(3)).astype(np.int64)
    low = (np.uint64(1) << b) - np.uint64(1)
    for lane in range(4):
        sel = l == lane
        lanes[(4 * j[sel]).astype(np.int64) + lane] = VALID | _piece(gg[sel], lane) | (keys[sel] & low)
    lead = l == 0
    left = (deg[gi[lead]] - r[lead]).astype(np.int64)
    for q in range(1, 4):
        sel = left <= q
        lanes[(4 * j[lead][sel]).astype(np.int64) + q] = _piece(gg[lead][sel], q)
    voff = None
    if nv <= 1 << 20:
        all_v = np.arange(nv + 1, dtype=np.uint64)
        e_all = eoff(all_v[:-1])
        e_all = np.append(e_all, np.uint64(m))
        voff = np.concatenate([[0], np.cumsum((e_all[1:] - e_all[:-1] + np.uint64(3)) // np.uint64(4))])
        voff = voff.astype(np.uint64)
    return lanes.reshape(-1, 4), voff


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pack_math_matches_host_packer(gz, seed):
    rng = np.random.default_rng(seed)
    nv = int(rng.integers(2, 3000))
    pairs = rng.integers(0, nv, size=(int(rng.integers(0, 20000)), 2), dtype=np.uint64)
    pairs[: pairs.shape[0] // 4, 1] = 7 % nv  # a hub destination spanning many vectors
    for by_dst in (True, False):
        lanes, voff = emulate(pairs, nv, by_dst)
        grouped = gz.sort_edges(pairs, by_dst=by_dst)
        ref = gz.pack_vectors(grouped, by_dst)
        assert lanes.shape == ref.shape
        assert np.array_equal(lanes, ref)
        assert np.array_equal(voff, gz.vertex_offsets(ref, nv))


def test_pack_math_high_ids(gz):
    # ids above 2^15 and 2^30 exercise every shared-vertex piece
    nv = (1 << 31) + 5
    ids = np.array([0, 1, (1 << 15) + 3, (1 << 30) + 9, nv - 1], dtype=np.uint64)
    pairs = np.array([(a, c) for a in ids for c in ids], dtype=np.uint64)
    lanes, _ = emulate(pairs, nv)
    ref = gz.pack_vectors(gz.sort_edges(pairs, by_dst=True), True)
    assert np.array_equal(lanes, ref)
