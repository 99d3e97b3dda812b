/*
 * gz_bfs.hip -- Breadth-First Search phases on gfx950 (level-synchronous,
 * hybrid top-down / bottom-up).
 *
 *   perform_edge_pull_phase_bfs  phase_edge_pull_bfs.asm:217-311 (step :53-211)
 *   perform_edge_push_phase_bfs  phase_edge_push_bfs.asm:208-302 (step :53-202)
 *   perform_vertex_phase_bfs     phase_vertex_bfs.asm:64-170
 *
 * State per vertex is one uint64 word, (depth << 48) | parent, with
 * 0xffffffffffffffff for "not visited" and parent 0xffffffffffff for "none"
 * (the root).  Frontiers are the reference's bitmaps: HasInfo (visited in the
 * previous level), WantsInfo (not yet visited), HasInfo* (visited in this
 * level).
 *
 * Push (top-down): for each active source and each valid lane whose
 * destination is in WantsInfo, atomic-min the destination word with
 * (level << 48 | source).  Exactly one lane sees the "not visited" old value;
 * it sets HasInfo* and counts the frontier statistic once (the reference's
 * weak check and clear are not atomic, phase_edge_push_bfs.asm:108-156).
 * When the driver reports a small frontier statistic, the push reads the
 * HasInfo words directly instead of compacting them (gz_sparse_push.h), with
 * the same per-edge work.
 * WantsInfo stays fixed during the kernel and loses the visited vertices in a
 * second pass, so every visiting source competes and the surviving parent is
 * the smallest one: the parent the reference picks when one thread walks the
 * sources in order.
 *
 * Pull (bottom-up): the unvisited destinations are compacted, their vectors
 * walked; a vector qualifies when its destination wants information and some
 * valid source is in HasInfo.  The parent is the largest qualifying source of
 * the destination's first qualifying vector (phase_edge_pull_bfs.asm:175-196
 * takes the max within the vector; a single thread stops at the first such
 * vector because it clears WantsInfo).  Pass A records each destination's
 * first qualifying vector with an atomic min over vector indices, pass B
 * resolves the parent and updates the frontiers.  Both engines therefore give
 * the same words on every run.
 */
#include <string.h>

#include "gz_frontier.h"
#include "gz_host_internal.h"
#include "gz_slots.h"
#include "gz_sparse_push.h"

/* The largest valid source of v in HasInfo (the asm's vmaxpd over the
 * qualifying lanes, phase_edge_pull_bfs.asm:175-196), or ~0 when none is;
 * counts v's valid lanes into `edges`. */
__device__ __forceinline__ uint64_t gz_vec_frontier_parent(const gz_lanes& v, const uint64_t* __restrict__ has_info,
                                                           uint32_t& edges)
{
    uint64_t parent = ~0ull;
    const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (gz_lane_valid(l[k])) {
            edges += 1;
            const uint64_t s = gz_lane_id(l[k]);
            if (gz_bit_test(has_info, s) && (parent == ~0ull || s > parent)) parent = s;
        }
    }
    return parent;
}

__device__ __forceinline__ bool gz_vec_has_frontier_source(const gz_lanes& v, const uint64_t* __restrict__ has_info,
                                                           uint32_t& edges)
{
    bool q = false;
    const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (gz_lane_valid(l[k])) {
            edges += 1;
            q = q || gz_bit_test(has_info, gz_lane_id(l[k]));
        }
    }
    return q;
}

template <bool kActive>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_push_bfs(const gz_vec_t* __restrict__ vecs, gz_slot_source src,
                                                               uint64_t chunk_len,
                                                               const uint64_t* __restrict__ has_info,
                                                               uint64_t* __restrict__ wants_info,
                                                               uint64_t* __restrict__ words,
                                                               uint64_t* __restrict__ next_frontier,
                                                               const double* __restrict__ outdeg, uint32_t stat_mode,
                                                               uint64_t level, uint64_t* __restrict__ partials,
                                                               uint64_t* __restrict__ stats, uint32_t collect,
                                                               gz_feature_sink sink)
{
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t W = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    const uint64_t T = gz_slot_total<kActive>(src);
    uint64_t vote = 0;
    gz_feature_acc feat{0, 0, 0};
    uint32_t vectors = 0, edges = 0;
    for (uint64_t chunk = gw; chunk * chunk_len < T; chunk += W) {
        const uint64_t j0 = chunk * chunk_len;
        const uint64_t jend = (j0 + chunk_len < T) ? j0 + chunk_len : T;
        uint64_t a = gz_slot_begin<kActive>(src, j0);
        for (uint64_t base = j0; base < jend; base += GZ_WAVE) {
            const bool valid = base + (uint64_t)lane < jend;
            uint64_t entry;
            const uint64_t vi = gz_slot_step<kActive>(src, a, base, lane, valid, &entry);
            if (!valid) continue;
            const gz_lanes v = gz_load_vector(vecs, vi);
            const uint64_t s = gz_shared_vertex(v);
            if (!kActive && !gz_bit_test(has_info, s)) continue;
            vectors += 1;
            const unsigned long long key = ((unsigned long long)level << 48) | (unsigned long long)s;
            const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!gz_lane_valid(l[k])) continue;
                edges += 1;
                const uint64_t d = gz_lane_id(l[k]);
                if (!gz_bit_test(wants_info, d)) continue;
                /* words only decrease: a value already <= key means a source
                 * no larger than s reached d at this level, and the atomic
                 * would change nothing */
                if (__hip_atomic_load(&words[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= key) continue;
                const unsigned long long old = atomicMin((unsigned long long*)&words[d], key);
                if (old == GZ_BFS_UNVISITED) {
                    gz_bit_set(next_frontier, d);
                    vote += gz_activate_stat(feat, sink, outdeg, d, stat_mode);
                }
            }
        }
    }
    const uint64_t total = gz_wave_sum_u64(vote);
    if (lane == 0) partials[gw] = total;
    gz_feature_flush(feat, sink, gw);
    if (collect) {
        const uint64_t tv = gz_wave_sum_u64(vectors), te = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * gw] = tv;
            stats[2 * gw + 1] = te;
        }
    }
}

/* The BFS push's per-edge work for the sparse push: the key (level << 48 |
 * source), the destination's WantsInfo word and its state read ahead, and the
 * atomic min, as in gz_k_edge_push_bfs. */
struct gz_bfs_push_op {
    const uint64_t* wants_info;
    uint64_t* words;
    uint64_t* next_frontier;
    const double* outdeg;
    uint32_t stat_mode;
    uint64_t level;
    gz_feature_sink sink;
    __device__ __forceinline__ uint64_t message(uint64_t s) const { return (level << 48) | s; }
    __device__ __forceinline__ void gather(uint64_t d, uint64_t* g) const
    {
        g[0] = wants_info[d >> 6];
        g[1] = __hip_atomic_load(&words[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void apply(uint64_t d, uint64_t key, const uint64_t* g, uint64_t& vote,
                                          gz_feature_acc& feat) const
    {
        if (!((g[0] >> (d & 63)) & 1ull)) return;
        /* words only decrease: a value already <= key means a source no
         * larger than this one reached d at this level */
        if (g[1] <= key) return;
        const unsigned long long old = atomicMin((unsigned long long*)&words[d], (unsigned long long)key);
        if (old == GZ_BFS_UNVISITED) {
            gz_bit_set(next_frontier, d);
            vote += gz_activate_stat(feat, sink, outdeg, d, stat_mode);
        }
    }
    __device__ __forceinline__ void finish(uint64_t wave, const gz_feature_acc& feat) const
    {
        gz_feature_flush(feat, sink, wave);
    }
};

/* WantsInfo &= ~HasInfo*: the push kernel reads WantsInfo as the snapshot of
 * "not visited before this level" and leaves it unchanged, so that every
 * visiting source takes part in the atomic min whatever the order of the
 * lanes; the visited vertices leave WantsInfo here, after the kernel. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_bfs_retire_wants(uint64_t* __restrict__ wants_info,
                                                                 const uint64_t* __restrict__ next_frontier,
                                                                 uint64_t words)
{
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x)
        wants_info[w] &= ~next_frontier[w];
}

/* Pass A0 of the bottom-up step over the compacted destinations: each
 * destination's first vector.  When it qualifies it is the destination's
 * first qualifying vector, so A0 visits the destination at once, as pass B
 * would: the parent is the vector's largest frontier source, the word, the
 * frontier bits, the vote and the selector's features are written here (in a
 * level with a large frontier, most destinations stop here, as the single
 * thread of the reference stops at its first qualifying vector).  A
 * destination whose first vector does not qualify and that has more vectors
 * gets its bit in `pending` (cleared beforehand); pass A walks those
 * destinations' vectors and pass B visits the ones it resolves.  Runs on pass
 * A's grid; the wave's counts go to its stats slot and pass A adds its own;
 * its vote and features go to wave slot gw, pass B's to the slots after the
 * grid's. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_bfs_first(const gz_vec_t* __restrict__ vecs,
                                                                     gz_slot_source src,
                                                                     const uint64_t* __restrict__ has_info,
                                                                     uint64_t* __restrict__ wants_info,
                                                                     uint64_t* __restrict__ pending,
                                                                     uint64_t* __restrict__ words,
                                                                     uint64_t* __restrict__ next_frontier,
                                                                     const double* __restrict__ outdeg,
                                                                     uint32_t stat_mode, uint64_t level,
                                                                     uint64_t* __restrict__ partials,
                                                                     gz_feature_sink sink,
                                                                     uint64_t* __restrict__ stats, uint32_t collect)
{
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t n = src.totals[0];
    uint32_t vectors = 0, edges = 0;
    uint64_t vote = 0;
    gz_feature_acc feat{0, 0, 0};
    /* wave-uniform loop: the list is in vertex order, so the vertices of
     * neighbouring lanes share bitmap words; one lane per run of equal words
     * applies the run's bits with one atomic per bitmap */
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(uint64_t)(GZ_WAVE - 1));
    /* the next slot's destination and vector range are loaded one iteration
     * ahead (the destination while this vector loads, its range while this
     * vector's frontier bits load), so that an iteration waits for its vector
     * and bits only */
    uint64_t dn = 0, vin = 0, ven = 0;
    if (first + (uint64_t)lane < n) {
        dn = src.list[first + (uint64_t)lane];
        vin = src.offsets[dn];
        ven = src.offsets[dn + 1];
    }
    for (uint64_t i0 = first; i0 < n; i0 += stride) {
        const uint64_t i = i0 + (uint64_t)lane;
        bool pend = false, found = false;
        const uint64_t d = dn, vi = vin, vend = ven;
        const uint64_t inext = i + stride;
        if (i < n) {
            const gz_lanes v = gz_load_vector(vecs, vi);
            if (inext < n) dn = src.list[inext];
            vectors += 1;
            const uint64_t parent = gz_vec_frontier_parent(v, has_info, edges);
            if (inext < n) {
                vin = src.offsets[dn];
                ven = src.offsets[dn + 1];
            }
            if (parent != ~0ull) {
                found = true;
                words[d] = ((uint64_t)level << 48) | parent;
                vote += gz_activate_stat(feat, sink, outdeg, d, stat_mode);
            } else {
                pend = vend - vi > 1;
            }
        }
        const uint64_t word = (i < n) ? (d >> 5) : ~0ull;
        uint32_t pbits = pend ? (1u << (d & 31)) : 0u;
        uint32_t fbits = found ? (1u << (d & 31)) : 0u;
#pragma unroll
        for (int o = 1; o < GZ_WAVE; o <<= 1) {
            const uint32_t xp = __shfl_down(pbits, o, GZ_WAVE);
            const uint32_t xf = __shfl_down(fbits, o, GZ_WAVE);
            const uint64_t xw = (uint64_t)__shfl_down((unsigned long long)word, o, GZ_WAVE);
            if (lane + o < GZ_WAVE && xw == word) {
                pbits |= xp;
                fbits |= xf;
            }
        }
        const uint64_t prev_word = (uint64_t)__shfl_up((unsigned long long)word, 1, GZ_WAVE);
        if (i < n && (lane == 0 || prev_word != word)) {
            if (pbits) atomicOr(reinterpret_cast<unsigned int*>(pending) + word, pbits);
            if (fbits) {
                atomicAnd(reinterpret_cast<unsigned int*>(wants_info) + word, ~fbits);
                atomicOr(reinterpret_cast<unsigned int*>(next_frontier) + word, fbits);
            }
        }
    }
    const uint64_t total = gz_wave_sum_u64(vote);
    if (lane == 0) partials[gw] = total;
    gz_feature_flush(feat, sink, gw);
    if (collect) {
        const uint64_t tv = gz_wave_sum_u64(vectors), te = gz_wave_sum_u64(edges);
        if (gz_lane() == 0) {
            stats[2 * gw] = tv;
            stats[2 * gw + 1] = te;
        }
    }
}

/* Pass A of the bottom-up step. */
template <bool kActive>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_bfs_find(const gz_vec_t* __restrict__ vecs,
                                                                    gz_slot_source src, uint64_t chunk_len,
                                                                    const uint64_t* __restrict__ has_info,
                                                                    const uint64_t* __restrict__ wants_info,
                                                                    uint64_t* __restrict__ first_vector,
                                                                    uint64_t* __restrict__ stats, uint32_t collect)
{
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t W = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    const uint64_t T = gz_slot_total<kActive>(src);
    uint32_t vectors = 0, edges = 0;
    for (uint64_t chunk = gw; chunk * chunk_len < T; chunk += W) {
        const uint64_t j0 = chunk * chunk_len;
        const uint64_t jend = (j0 + chunk_len < T) ? j0 + chunk_len : T;
        uint64_t a = gz_slot_begin<kActive>(src, j0);
        uint64_t carry_d = ~0ull;
        bool carry_found = false;
        for (uint64_t base = j0; base < jend; base += GZ_WAVE) {
            const bool valid = base + (uint64_t)lane < jend;
            uint64_t entry;
            const uint64_t vi = gz_slot_step<kActive>(src, a, base, lane, valid, &entry);
            uint64_t d = ~0ull;
            bool cand = false;
            if (valid && kActive) {
                /* the destination comes from the compacted list, so a vector
                 * whose destination already has a smaller qualifying vector
                 * (pass A0's first vector, or another chunk's) is not loaded;
                 * pass A0 has tested every first vector */
                d = src.list[entry];
                const uint64_t known = __hip_atomic_load(&first_vector[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (known > vi && base + (uint64_t)lane != src.prefix[entry]) {
                    const gz_lanes v = gz_load_vector(vecs, vi);
                    vectors += 1;
                    cand = gz_vec_has_frontier_source(v, has_info, edges);
                }
            } else if (valid) {
                const gz_lanes v = gz_load_vector(vecs, vi);
                d = gz_shared_vertex(v);
                if (gz_bit_test(wants_info, d)) {
                    vectors += 1;
                    /* a smaller qualifying vector of d is already known (from
                     * another chunk): this one cannot win, skip its gathers */
                    const uint64_t known = __hip_atomic_load(&first_vector[d], __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT);
                    if (known > vi) cand = gz_vec_has_frontier_source(v, has_info, edges);
                }
            }
            const uint64_t d_prev = __shfl_up((unsigned long long)d, 1, GZ_WAVE);
            const uint64_t heads = __ballot(lane == 0 || d != d_prev);
            const uint64_t cands = __ballot(cand);
            const int seg = gz_segment_start(heads, lane);
            const uint64_t before = ((1ull << lane) - 1ull) & ~((1ull << seg) - 1ull);
            const bool carried = (d == carry_d) && carry_found;
            if (cand && !(cands & before) && !carried)
                atomicMin((unsigned long long*)&first_vector[d], (unsigned long long)vi);
            const int last = (int)((jend - 1 - base) < (uint64_t)(GZ_WAVE - 1) ? (jend - 1 - base) : (uint64_t)(GZ_WAVE - 1));
            const uint64_t last_d = __shfl((unsigned long long)d, last, GZ_WAVE);
            const int last_seg = __shfl(seg, last, GZ_WAVE);
            const uint64_t last_mask = ((last == 63) ? ~0ull : ((2ull << last) - 1ull)) & ~((1ull << last_seg) - 1ull);
            carry_found = (cands & last_mask) != 0 || (last_d == carry_d && carry_found);
            carry_d = last_d;
        }
    }
    if (collect) {
        const uint64_t tv = gz_wave_sum_u64(vectors), te = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * gw] = (kActive ? stats[2 * gw] : 0) + tv;
            stats[2 * gw + 1] = (kActive ? stats[2 * gw + 1] : 0) + te;
        }
    }
}

/* Pass B: visit every destination that found a qualifying vector. */
template <bool kActive>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_bfs_visit(const gz_vec_t* __restrict__ vecs,
                                                                     const uint64_t* __restrict__ list,
                                                                     const uint64_t* __restrict__ totals,
                                                                     uint64_t num_vertices,
                                                                     const uint64_t* __restrict__ has_info,
                                                                     uint64_t* __restrict__ wants_info,
                                                                     uint64_t* __restrict__ first_vector,
                                                                     uint64_t* __restrict__ words,
                                                                     uint64_t* __restrict__ next_frontier,
                                                                     const double* __restrict__ outdeg,
                                                                     uint32_t stat_mode, uint64_t level,
                                                                     uint64_t* __restrict__ partials,
                                                                     gz_feature_sink sink)
{
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t n = kActive ? totals[0] : num_vertices;
    uint64_t vote = 0;
    gz_feature_acc feat{0, 0, 0};
    /* wave-uniform loop: the list is in vertex order, so visited vertices of
     * neighbouring lanes share bitmap words; one lane per run of equal words
     * applies the whole run's bits with a single atomic per bitmap */
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(uint64_t)(GZ_WAVE - 1)); i0 < n;
         i0 += stride) {
        const uint64_t i = i0 + (uint64_t)lane;
        const uint64_t d = (i < n) ? (kActive ? list[i] : i) : 0;
        const uint64_t fv = (i < n) ? first_vector[d] : ~0ull;
        const bool found = fv != ~0ull;
        if (found) {
            first_vector[d] = ~0ull;
            const gz_lanes v = gz_load_vector(vecs, fv);
            uint64_t parent = 0;
            const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (gz_lane_valid(l[k])) {
                    const uint64_t s = gz_lane_id(l[k]);
                    if (gz_bit_test(has_info, s) && s > parent) parent = s;
                }
            }
            words[d] = ((uint64_t)level << 48) | parent;
            vote += gz_activate_stat(feat, sink, outdeg, d, stat_mode);
        }
        const uint64_t word = found ? (d >> 5) : ~0ull;
        uint32_t bits = found ? (1u << (d & 31)) : 0u;
#pragma unroll
        for (int o = 1; o < GZ_WAVE; o <<= 1) {
            const uint32_t xb = __shfl_down(bits, o, GZ_WAVE);
            const uint64_t xw = (uint64_t)__shfl_down((unsigned long long)word, o, GZ_WAVE);
            if (lane + o < GZ_WAVE && xw == word) bits |= xb;
        }
        const uint64_t prev_word = (uint64_t)__shfl_up((unsigned long long)word, 1, GZ_WAVE);
        if (found && (lane == 0 || prev_word != word)) {
            atomicAnd(reinterpret_cast<unsigned int*>(wants_info) + word, ~bits);
            atomicOr(reinterpret_cast<unsigned int*>(next_frontier) + word, bits);
        }
    }
    const uint64_t total = gz_wave_sum_u64(vote);
    if (lane == 0) partials[gw] = total;
    gz_feature_flush(feat, sink, gw);
}

static gz_slot_source gz_bfs_source(const uint64_t* offsets, uint64_t count)
{
    gz_slot_source src;
    src.offsets = offsets;
    src.list = gz_scratch.active_vertices;
    src.prefix = gz_scratch.active_prefix;
    src.totals = gz_scratch.active_totals;
    src.sweep_count = count;
    return src;
}

/* The bottom-up step's pending destinations (first vector not qualifying,
 * more vectors to test): a bitmap and its compaction, sized for the graph. */
static struct {
    uint64_t* bits;
    gz_compact_out out;
    uint64_t vertices;
} gz_bfs_pend;

static void gz_bfs_pending_reserve(uint64_t V)
{
    if (gz_bfs_pend.vertices >= V && gz_bfs_pend.bits != nullptr) return;
    gz_device_free(gz_bfs_pend.bits);
    gz_device_free(gz_bfs_pend.out.list);
    gz_device_free(gz_bfs_pend.out.prefix);
    gz_device_free(gz_bfs_pend.out.block_sums);
    gz_device_free(gz_bfs_pend.out.totals);
    gz_bfs_pend.bits = (uint64_t*)gz_device_malloc(((V + 63) / 64 + 8) * sizeof(uint64_t));
    gz_bfs_pend.out.list = (uint64_t*)gz_device_malloc((V + 8) * sizeof(uint64_t));
    gz_bfs_pend.out.prefix = (uint64_t*)gz_device_malloc((V + 8) * sizeof(uint64_t));
    gz_bfs_pend.out.block_sums = (uint64_t*)gz_device_malloc((2 * (gz_compact_blocks(V) + 1) + 8) * sizeof(uint64_t));
    gz_bfs_pend.out.totals = (uint64_t*)gz_device_malloc(8 * sizeof(uint64_t));
    gz_bfs_pend.vertices = V;
}

extern "C" void gz_bfs_pending_release(void)
{
    gz_device_free(gz_bfs_pend.bits);
    gz_device_free(gz_bfs_pend.out.list);
    gz_device_free(gz_bfs_pend.out.prefix);
    gz_device_free(gz_bfs_pend.out.block_sums);
    gz_device_free(gz_bfs_pend.out.totals);
    memset(&gz_bfs_pend, 0, sizeof(gz_bfs_pend));
}

/* Frontier statistic of the next push, when the driver knows it (the vote of
 * the level that built the frontier); ~0 when unknown.  Consumed by the next
 * push. */
static uint64_t gz_bfs_next_push_vote = ~0ull;

extern "C" void gz_bfs_push_vote_hint(uint64_t vote) { gz_bfs_next_push_vote = vote; }

static void gz_bfs_retire(void)
{
    const uint64_t words = (gz_graph.num_vertices + 63) / 64;
    uint64_t rblocks = (words + GZ_BLOCK - 1) / GZ_BLOCK;
    if (rblocks > 4096) rblocks = 4096;
    hipLaunchKernelGGL(gz_k_bfs_retire_wants, dim3((unsigned)rblocks), dim3(GZ_BLOCK), 0, gz_stream,
                       gz_graph.frontier_wants_info, (const uint64_t*)gz_graph.vertex_accumulators, words);
    GZ_LAUNCH_CHECK();
}

extern "C" void perform_edge_push_phase_bfs(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    const uint64_t chunk_len = 256;
    const uint64_t hint = gz_bfs_next_push_vote;
    gz_bfs_next_push_vote = ~0ull;
    gz_scratch_reserve(0, gz_graph.num_vertices);
    if (gz_push_sparse_eligible(hint)) {
        /* reserve first: the feature sink points into the per-wave scratch */
        gz_scratch_reserve(gz_push_sparse_waves(hint), gz_graph.num_vertices);
        const gz_bfs_push_op op{gz_graph.frontier_wants_info,
                                (uint64_t*)gz_graph.vertex_props,
                                (uint64_t*)gz_graph.vertex_accumulators,
                                gz_graph.vertex_outdegrees,
                                gz_phase_config.stat_mode,
                                (uint64_t)gz_phase_config.bfs_level,
                                {gz_scratch.features_on ? gz_scratch.features : nullptr, gz_scratch.features_indeg,
                                 gz_scratch.partials_capacity}};
        uint64_t* stamp = gz_scratch.stamp_target;
        gz_scratch.stamp_target = nullptr;
        gz_ktimer_begin(GZ_K_PUSH_BFS);
        gz_launch_push_sparse(op, edge_list, hint, stamp);
        gz_ktimer_end(GZ_K_PUSH_BFS);
        gz_bfs_retire();
        return;
    }
    const bool active = gz_phase_config.push_frontier_skip && gz_graph.push_vertex_offsets != nullptr;
    gz_slot_source src = gz_bfs_source(gz_graph.push_vertex_offsets, edge_list_count);
    uint64_t blocks;
    if (active) {
        gz_frontier_compact(gz_graph.frontier_has_info, gz_graph.push_vertex_offsets, gz_graph.num_vertices,
                            edge_list_count);
        blocks = GZ_PERSISTENT_BLOCKS;
    } else {
        blocks = ((edge_list_count + chunk_len - 1) / chunk_len + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
        if (blocks == 0) blocks = 1;
    }
    const uint64_t waves = blocks * GZ_WAVES_PER_BLOCK;
    gz_scratch_reserve(waves, gz_graph.num_vertices);
    gz_scratch.partials_is_double = 0;
    gz_scratch.partials_count = waves;
    gz_scratch.stats_count = gz_phase_config.collect_stats ? waves : 0;
    const uint64_t level = gz_phase_config.bfs_level;
    const gz_feature_sink sink{gz_scratch.features_on ? gz_scratch.features : nullptr, gz_scratch.features_indeg,
                                 gz_scratch.partials_capacity};
    gz_ktimer_begin(GZ_K_PUSH_BFS);
    if (active)
        hipLaunchKernelGGL(gz_k_edge_push_bfs<true>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list,
                           src, chunk_len, gz_graph.frontier_has_info, gz_graph.frontier_wants_info,
                           (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators,
                           gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, level,
                           (uint64_t*)gz_scratch.partials, gz_scratch.stats, gz_phase_config.collect_stats, sink);
    else
        hipLaunchKernelGGL(gz_k_edge_push_bfs<false>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list,
                           src, chunk_len, gz_graph.frontier_has_info, gz_graph.frontier_wants_info,
                           (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators,
                           gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, level,
                           (uint64_t*)gz_scratch.partials, gz_scratch.stats, gz_phase_config.collect_stats, sink);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PUSH_BFS);
    gz_bfs_retire();
}

extern "C" void perform_edge_pull_phase_bfs(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    const uint64_t chunk_len = 256;
    gz_scratch_reserve(0, gz_graph.num_vertices);
    const bool active = gz_phase_config.push_frontier_skip && gz_graph.pull_vertex_offsets != nullptr;
    gz_slot_source src = gz_bfs_source(gz_graph.pull_vertex_offsets, edge_list_count);
    uint64_t blocks;
    if (active) {
        /* pass A0 and the visit walk the unvisited list; pass A has a list of
         * its own (below), so no vector prefix is needed here */
        gz_frontier_list(gz_graph.frontier_wants_info, gz_graph.pull_vertex_offsets, gz_graph.num_vertices,
                         edge_list_count);
        blocks = GZ_PERSISTENT_BLOCKS;
    } else {
        blocks = ((edge_list_count + chunk_len - 1) / chunk_len + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
        if (blocks == 0) blocks = 1;
    }
    uint64_t visit_blocks = GZ_PERSISTENT_BLOCKS;
    const uint64_t waves = ((blocks > visit_blocks) ? blocks : visit_blocks) * GZ_WAVES_PER_BLOCK;
    /* with the frontier skip, pass A0 (on pass A's grid of `blocks`) and pass
     * B each have a slot per wave: A0's first, then B's */
    const uint64_t vote_slots = active ? (blocks + visit_blocks) * GZ_WAVES_PER_BLOCK : visit_blocks * GZ_WAVES_PER_BLOCK;
    gz_scratch_reserve(waves > vote_slots ? waves : vote_slots, gz_graph.num_vertices);
    gz_scratch.partials_is_double = 0;
    gz_scratch.partials_count = vote_slots;
    gz_scratch.stats_count = gz_phase_config.collect_stats ? blocks * GZ_WAVES_PER_BLOCK : 0;
    const uint64_t level = gz_phase_config.bfs_level;
    const gz_feature_sink sink{gz_scratch.features_on ? gz_scratch.features : nullptr, gz_scratch.features_indeg,
                                 gz_scratch.partials_capacity};
    gz_ktimer_begin(GZ_K_PULL_BFS_FIND);
    gz_slot_source pend_src = src;
    if (active) {
        const uint64_t V = gz_graph.num_vertices;
        gz_bfs_pending_reserve(V);
        gz_dev_fill_u64(gz_bfs_pend.bits, 0, (V + 63) / 64);
        hipLaunchKernelGGL(gz_k_edge_pull_bfs_first, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list,
                           src, gz_graph.frontier_has_info, gz_graph.frontier_wants_info, gz_bfs_pend.bits,
                           (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators,
                           gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, level,
                           (uint64_t*)gz_scratch.partials, sink, gz_scratch.stats, gz_phase_config.collect_stats);
        GZ_LAUNCH_CHECK();
        /* pass A walks the pending destinations' vectors only (it skips each
         * one's first vector, which pass A0 tested); the compaction has its
         * own kernel timer */
        gz_ktimer_end(GZ_K_PULL_BFS_FIND);
        gz_frontier_compact_to(gz_bfs_pend.bits, gz_graph.pull_vertex_offsets, V, edge_list_count, gz_bfs_pend.out);
        gz_ktimer_begin(GZ_K_PULL_BFS_FIND);
        pend_src.list = gz_bfs_pend.out.list;
        pend_src.prefix = gz_bfs_pend.out.prefix;
        pend_src.totals = gz_bfs_pend.out.totals;
    }
    if (active)
        hipLaunchKernelGGL(gz_k_edge_pull_bfs_find<true>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream,
                           edge_list, pend_src, chunk_len, gz_graph.frontier_has_info, gz_graph.frontier_wants_info,
                           gz_scratch.bfs_first_vector, gz_scratch.stats, gz_phase_config.collect_stats);
    else
        hipLaunchKernelGGL(gz_k_edge_pull_bfs_find<false>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream,
                           edge_list, src, chunk_len, gz_graph.frontier_has_info, gz_graph.frontier_wants_info,
                           gz_scratch.bfs_first_vector, gz_scratch.stats, gz_phase_config.collect_stats);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PULL_BFS_FIND);
    gz_ktimer_begin(GZ_K_PULL_BFS_VISIT);
    if (active) {
        /* pass B visits the pending destinations pass A resolved */
        const uint64_t a0_waves = blocks * GZ_WAVES_PER_BLOCK;
        const gz_feature_sink bsink{sink.out ? sink.out + a0_waves : nullptr, sink.indeg, sink.stride};
        hipLaunchKernelGGL(gz_k_edge_pull_bfs_visit<true>, dim3((unsigned)visit_blocks), dim3(GZ_BLOCK), 0, gz_stream,
                           edge_list, gz_bfs_pend.out.list, gz_bfs_pend.out.totals, gz_graph.num_vertices,
                           gz_graph.frontier_has_info, gz_graph.frontier_wants_info, gz_scratch.bfs_first_vector,
                           (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators,
                           gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, level,
                           (uint64_t*)gz_scratch.partials + a0_waves, bsink);
    }
    else
        hipLaunchKernelGGL(gz_k_edge_pull_bfs_visit<false>, dim3((unsigned)visit_blocks), dim3(GZ_BLOCK), 0, gz_stream,
                           edge_list, gz_scratch.active_vertices, gz_scratch.active_totals, gz_graph.num_vertices,
                           gz_graph.frontier_has_info, gz_graph.frontier_wants_info, gz_scratch.bfs_first_vector,
                           (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators,
                           gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, level,
                           (uint64_t*)gz_scratch.partials, sink);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PULL_BFS_VISIT);
}

extern "C" void gz_clear_bitmap_range(uint64_t* bitmap, uint64_t vertex_start, uint64_t vertex_count);

extern "C" void perform_vertex_phase_bfs(uint64_t vertex_start, uint64_t vertex_count, const uint64_t* reduce_buffer)
{
    (void)reduce_buffer;
    gz_clear_bitmap_range((uint64_t*)gz_graph.vertex_accumulators, vertex_start, vertex_count);
}
