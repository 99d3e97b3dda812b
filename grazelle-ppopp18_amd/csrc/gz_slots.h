/*
 * gz_slots.h -- the vector walk shared by the frontier-driven kernels.
 *
 * A kernel walks "slots": either every vector of a list (sweep mode, slot j is
 * vector j) or the concatenated vector ranges of the compacted active
 * vertices (active mode, see gz_frontier.h).  Work is split into chunks of
 * `chunk_len` slots handed out grid-stride to wavefronts; inside a chunk a
 * wavefront advances 64 slots per step.  gz_slot_step() gives each lane the
 * vector index of its slot and keeps the owner entry for the next step.
 */
#ifndef GZ_SLOTS_H
#define GZ_SLOTS_H

#include "gz_frontier.h"

struct gz_slot_source {
    const uint64_t* offsets;  /* [V+1] vector offsets of the list */
    const uint64_t* list;     /* active vertices (active mode) */
    const uint64_t* prefix;   /* [n+1] exclusive prefix of their vector counts */
    const uint64_t* totals;   /* [0] active vertices, [1] active vectors */
    uint64_t sweep_count;     /* vectors in the list (sweep mode) */
};

template <bool kActive>
__device__ __forceinline__ uint64_t gz_slot_total(const gz_slot_source& s)
{
    return kActive ? s.totals[1] : s.sweep_count;
}

template <bool kActive>
__device__ __forceinline__ uint64_t gz_slot_begin(const gz_slot_source& s, uint64_t j0)
{
    return kActive ? gz_active_search(s.prefix, s.totals[0], j0) : 0;
}

/* Wavefront-collective: all 64 lanes must call it.  Returns the vector index
 * of slot base+lane (meaningful only when that slot is below the chunk end)
 * and, in active mode, writes the owner entry of that slot to *entry. */
template <bool kActive>
__device__ __forceinline__ uint64_t gz_slot_step(const gz_slot_source& s, uint64_t& a, uint64_t base, int lane,
                                                 bool valid, uint64_t* entry)
{
    if (!kActive) {
        *entry = base + (uint64_t)lane;
        return base + (uint64_t)lane;
    }
    const uint64_t n = s.totals[0];
    const uint64_t my_a = gz_active_owner(s.prefix, n, a, base, lane);
    uint64_t vec = 0;
    if (valid) vec = s.offsets[s.list[my_a]] + (base + (uint64_t)lane - s.prefix[my_a]);
    a = __shfl((unsigned long long)my_a, GZ_WAVE - 1, GZ_WAVE);
    *entry = my_a;
    return vec;
}

#endif /* GZ_SLOTS_H */
