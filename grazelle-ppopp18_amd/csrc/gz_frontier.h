/*
 * gz_frontier.h -- device side of the frontier skip.
 *
 * The reference skips inactive top-level vertices inside its vector loop: a
 * strong frontier check (phasehelpers.inc:137-213) TZCNT-scans the bitmap for
 * the next active vertex and jumps to its first vector through the vertex
 * index.  On the GPU the same skip is done up front: gz_frontier_compact()
 * turns a bitmap into a list of active vertices that own vectors plus an
 * exclusive prefix of their vector counts, and the edge kernels then walk the
 * concatenated vector ranges ("slots") with every lane busy, however skewed
 * the degrees are.
 *
 * Slot j belongs to active entry a with prefix[a] <= j < prefix[a+1]; its
 * vector is offsets[vertex[a]] + (j - prefix[a]).  Within a slot range the
 * shared vertex is nondecreasing, so the wavefront segmented logic of the
 * full-sweep kernels applies unchanged.
 */
#ifndef GZ_FRONTIER_H
#define GZ_FRONTIER_H

#include "gz_internal.h"

/* Vertices per thread and per block of the compaction passes. */
#define GZ_COMPACT_VPT 16
#define GZ_COMPACT_VPB (GZ_COMPACT_VPT * GZ_BLOCK)

static inline uint64_t gz_compact_blocks(uint64_t num_vertices)
{
    return (num_vertices + GZ_COMPACT_VPB - 1) / GZ_COMPACT_VPB;
}

__device__ __forceinline__ uint64_t gz_wave_or_u64(uint64_t x)
{
#pragma unroll
    for (int o = GZ_WAVE / 2; o > 0; o >>= 1) x |= __shfl_xor(x, o, GZ_WAVE);
    return x;
}

/* Largest a in [0, n) with prefix[a] <= j (prefix[n] > j is required). */
__device__ __forceinline__ uint64_t gz_active_search(const uint64_t* __restrict__ prefix, uint64_t n, uint64_t j)
{
    uint64_t lo = 0, hi = n;  /* invariant: prefix[lo] <= j < prefix[hi] */
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (prefix[mid] <= j)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

/* Owner entry of slot base+lane, given a with prefix[a] <= base and at most
 * 64 entries starting inside [base, base+64).  Wavefront-collective. */
__device__ __forceinline__ uint64_t gz_active_owner(const uint64_t* __restrict__ prefix, uint64_t n, uint64_t a,
                                                    uint64_t base, int lane)
{
    const uint64_t k = a + 1 + (uint64_t)lane;
    uint64_t bit = 0;
    if (k <= n) {
        const uint64_t r = prefix[k] - base;
        if (r < (uint64_t)GZ_WAVE) bit = 1ull << r;
    }
    const uint64_t starts = gz_wave_or_u64(bit);
    return a + (uint64_t)__popcll(starts & ((2ull << lane) - 1ull));
}

#endif /* GZ_FRONTIER_H */
