/*
 * gz_sparse_push.h -- push over a small frontier without compacting it.
 *
 * The frontier-driven pushes (CC, BFS) normally compact the HasInfo bitmap
 * into a list of active sources first (three kernels).  When the driver knows
 * the frontier statistic that built the frontier and it is small, these
 * kernels read the bitmap words directly instead:
 *
 *  - gz_k_push_sparse: each wavefront reads kWords HasInfo words (64 vertices
 *    each) and numbers their set bits with a scan of the word popcounts.
 *    Rounds of 64 sources follow: each lane selects one source (owner word by
 *    binary search over the scan, then the n-th set bit), and a scan of the
 *    sources' vector counts spreads their vectors over the 64 lanes, each
 *    lane finding its source the same way, in batches of four steps of 64
 *    vectors.  With kQueue a round of more than one batch is cut instead into
 *    pieces of at most GZ_PUSH_PIECE_VECTORS vectors, appended to the overflow
 *    queue, so that no wavefront walks a high-degree source alone; without
 *    kQueue every round is pushed in place.
 *  - gz_k_push_pieces: wavefront gw pushes pieces gw, gw + W, ... of the queue.
 *
 * The per-edge work is the Op's (one struct per algorithm):
 *   uint64_t message(uint64_t s)            the source's message, per vector;
 *   void gather(uint64_t d, uint64_t g[2])  reads for edge (s, d), issued for a
 *                                           whole batch before any update;
 *   void apply(uint64_t d, uint64_t msg, const uint64_t g[2], uint64_t& vote,
 *              gz_feature_acc& feat)        the update;
 *   void finish(uint64_t wave, const gz_feature_acc& feat)  per-wave output
 *                                           besides the vote and statistics.
 */
#ifndef GZ_SPARSE_PUSH_H
#define GZ_SPARSE_PUSH_H

#include "gz_walk.h"

/* Index of the first lane whose (non-decreasing) inclusive count exceeds x;
 * 63 when none does. */
__device__ __forceinline__ int gz_wave_owner(uint64_t incl, uint64_t x)
{
    int o = 0;
#pragma unroll
    for (int step = GZ_WAVE / 2; step >= 1; step >>= 1) {
        const uint64_t y = (uint64_t)__shfl((unsigned long long)incl, o + step - 1, GZ_WAVE);
        if (y <= x) o += step;
    }
    return o;
}

__device__ __forceinline__ uint64_t gz_wave_inclusive_sum(uint64_t x, int lane)
{
#pragma unroll
    for (int o = 1; o < GZ_WAVE; o <<= 1) {
        const uint64_t t = (uint64_t)__shfl_up((unsigned long long)x, o, GZ_WAVE);
        if (lane >= o) x += t;
    }
    return x;
}

/* Push up to kSteps steps of 64 vectors: step u < nsteps holds vector vi[u]
 * of source src[u] on lanes with ok[u].  All vector loads are issued first,
 * then every edge's gathers, so a lane keeps 4 kSteps of them in flight.
 * Lanes without ok[u] read vector 0 and gather for vertex 0, and are
 * ignored. */
template <int kSteps, class Op>
__device__ __forceinline__ void gz_push_steps(const gz_vec_t* __restrict__ vecs, const Op& op, int nsteps,
                                              const bool* ok, const uint64_t* vi, const uint64_t* src,
                                              uint64_t& vote, gz_feature_acc& feat, uint32_t& vectors,
                                              uint32_t& edges)
{
    uint64_t l[kSteps][4];
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        l[u][0] = l[u][1] = l[u][2] = l[u][3] = 0;
        if (u >= nsteps) continue;
        const gz_lanes vec = gz_load_vector(vecs, ok[u] ? vi[u] : 0);
        l[u][0] = vec.l0;
        l[u][1] = vec.l1;
        l[u][2] = vec.l2;
        l[u][3] = vec.l3;
    }
    uint64_t msg[kSteps], g[kSteps][4][2];
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        msg[u] = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) g[u][k][0] = g[u][k][1] = 0;
        if (u >= nsteps) continue;
        msg[u] = op.message(ok[u] ? src[u] : 0);
#pragma unroll
        for (int k = 0; k < 4; ++k) op.gather((ok[u] && gz_lane_valid(l[u][k])) ? gz_lane_id(l[u][k]) : 0, g[u][k]);
    }
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        if (u >= nsteps || !ok[u]) continue;
        vectors += 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!gz_lane_valid(l[u][k])) continue;
            edges += 1;
            op.apply(gz_lane_id(l[u][k]), msg[u], g[u][k], vote, feat);
        }
    }
}

template <class Op>
__device__ __forceinline__ void gz_push_finish(const Op& op, uint64_t gw, uint64_t vote, const gz_feature_acc& feat,
                                               uint32_t vectors, uint32_t edges, uint64_t* __restrict__ partials,
                                               uint64_t* __restrict__ stats, uint32_t collect)
{
    const uint64_t tot = gz_wave_sum_u64(vote);
    if (gz_lane() == 0) partials[gw] = tot;
    op.finish(gw, feat);
    if (collect) {
        const uint64_t tv = gz_wave_sum_u64(vectors), te = gz_wave_sum_u64(edges);
        if (gz_lane() == 0) {
            stats[2 * gw] = tv;
            stats[2 * gw + 1] = te;
        }
    }
}

template <class Op, int kWords, bool kQueue>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_push_sparse(
    Op op, const gz_vec_t* __restrict__ vecs, const uint64_t* __restrict__ offsets,
    const uint64_t* __restrict__ has_info, uint64_t num_vertices, uint64_t* __restrict__ pieces,
    uint64_t* __restrict__ queue, uint64_t queue_cap, uint64_t* __restrict__ partials, uint64_t* __restrict__ stats,
    uint32_t collect, uint64_t* __restrict__ stamp, gz_guard guard)
{
    if (gz_guard_off(guard)) return;
    constexpr int kSteps = GZ_PUSH_PIECE_VECTORS / GZ_WAVE;
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    if (stamp != nullptr && gw == 0 && lane == 0) *stamp = (uint64_t)wall_clock64();
    const uint64_t nwords = (num_vertices + 63) >> 6;
    const uint64_t w = gw * kWords + (uint64_t)lane;
    const bool word_lane = lane < kWords && w < nwords;
    uint64_t bits = word_lane ? has_info[w] : 0;
    if (word_lane && w == nwords - 1 && (num_vertices & 63)) bits &= (1ull << (num_vertices & 63)) - 1;
    const uint64_t pc = (uint64_t)__popcll(bits);
    const uint64_t pincl = gz_wave_inclusive_sum(pc, lane);
    const uint64_t sources = gz_readlane_u64(pincl, GZ_WAVE - 1);
    uint64_t vote = 0;
    gz_feature_acc feat{0, 0, 0};
    uint32_t vectors = 0, edges = 0;
    for (uint64_t r = 0; r < sources; r += GZ_WAVE) {
        const uint64_t idx = r + (uint64_t)lane;
        const int ow = gz_wave_owner(pincl, idx);
        const uint64_t wbits = (uint64_t)__shfl((unsigned long long)bits, ow, GZ_WAVE);
        const uint64_t wex = (uint64_t)__shfl((unsigned long long)(pincl - pc), ow, GZ_WAVE);
        uint64_t v = 0, first = 0, n = 0;
        if (idx < sources) {
            v = ((gw * kWords + (uint64_t)ow) << 6) + gz_nth_set_bit(wbits, (uint32_t)(idx - wex));
            first = offsets[v];
            n = offsets[v + 1] - first;
        }
        const uint64_t incl = gz_wave_inclusive_sum(n, lane);
        const uint64_t total = gz_readlane_u64(incl, GZ_WAVE - 1);
        if (kQueue && total > (uint64_t)GZ_PUSH_PIECE_VECTORS) {
            /* one queue reservation for the round's pieces */
            const uint64_t np = (n + GZ_PUSH_PIECE_VECTORS - 1) / GZ_PUSH_PIECE_VECTORS;
            const uint64_t npincl = gz_wave_inclusive_sum(np, lane);
            const uint64_t npall = gz_readlane_u64(npincl, GZ_WAVE - 1);
            uint64_t at = 0;
            if (lane == 0) at = (uint64_t)atomicAdd((unsigned long long*)queue, (unsigned long long)npall);
            at = gz_readlane_u64(at, 0) + (npincl - np);
            for (uint64_t k = 0; k < np; ++k) {
                const uint64_t pf = first + k * GZ_PUSH_PIECE_VECTORS;
                const uint64_t rest = n - k * GZ_PUSH_PIECE_VECTORS;
                const uint64_t pl = rest < GZ_PUSH_PIECE_VECTORS ? rest : GZ_PUSH_PIECE_VECTORS;
                const uint64_t slot = at + k;
                if (slot < queue_cap) {
                    pieces[3 * slot] = v;
                    pieces[3 * slot + 1] = pf;
                    pieces[3 * slot + 2] = pl;
                } else {
                    /* queue full (more pieces than a frontier within the
                     * sparse bound can make): this lane pushes the piece */
                    const uint64_t m = op.message(v);
                    for (uint64_t j = 0; j < pl; ++j) {
                        const gz_lanes vec = gz_load_vector(vecs, pf + j);
                        const uint64_t lj[4] = {vec.l0, vec.l1, vec.l2, vec.l3};
                        vectors += 1;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            if (!gz_lane_valid(lj[q])) continue;
                            edges += 1;
                            uint64_t g[2];
                            op.gather(gz_lane_id(lj[q]), g);
                            op.apply(gz_lane_id(lj[q]), m, g, vote, feat);
                        }
                    }
                }
            }
            continue;
        }
        const uint64_t excl = incl - n;
        for (uint64_t base = 0; base < total; base += (uint64_t)kSteps * GZ_WAVE) {
            /* wave-uniform count of the steps this batch holds */
            const uint64_t left = (total - base + GZ_WAVE - 1) / GZ_WAVE;
            const int nsteps = left < (uint64_t)kSteps ? (int)left : kSteps;
            bool ok[kSteps];
            uint64_t vi[kSteps], src[kSteps];
#pragma unroll
            for (int u = 0; u < kSteps; ++u) {
                ok[u] = false;
                vi[u] = src[u] = 0;
                if (u >= nsteps) continue;
                const uint64_t slot = base + (uint64_t)(u * GZ_WAVE + lane);
                const int o = gz_wave_owner(incl, slot);
                src[u] = (uint64_t)__shfl((unsigned long long)v, o, GZ_WAVE);
                const uint64_t f = (uint64_t)__shfl((unsigned long long)first, o, GZ_WAVE);
                const uint64_t e0 = (uint64_t)__shfl((unsigned long long)excl, o, GZ_WAVE);
                ok[u] = slot < total;
                vi[u] = f + (slot - e0);
            }
            gz_push_steps<kSteps>(vecs, op, nsteps, ok, vi, src, vote, feat, vectors, edges);
        }
    }
    gz_push_finish(op, gw, vote, feat, vectors, edges, partials, stats, collect);
}

/* The overflow queue of gz_k_push_sparse: wavefront gw pushes pieces gw,
 * gw + W, ... of the queue[parity] appended there.  It also clears the other
 * counter, which the next sparse push appends to (the kernel that read that
 * counter last has finished: launches on one stream run in order).  Its
 * per-wave outputs (vote partial, statistics, the Op's features) go to slot
 * wave_base + gw, after the wavefronts of gz_k_push_sparse. */
template <class Op>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_push_pieces(Op op, const gz_vec_t* __restrict__ vecs,
                                                             const uint64_t* __restrict__ pieces,
                                                             uint64_t* __restrict__ queue, uint64_t queue_cap,
                                                             uint32_t parity, uint64_t wave_base,
                                                             uint64_t* __restrict__ partials,
                                                             uint64_t* __restrict__ stats, uint32_t collect,
                                                             gz_guard guard)
{
    if (gz_guard_off(guard)) return;
    constexpr int kSteps = GZ_PUSH_PIECE_VECTORS / GZ_WAVE;
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t W = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    if (gw == 0 && lane == 0) queue[parity ^ 1u] = 0;
    uint64_t count = queue[parity];
    if (count > queue_cap) count = queue_cap;
    uint64_t vote = 0;
    gz_feature_acc feat{0, 0, 0};
    uint32_t vectors = 0, edges = 0;
    for (uint64_t p = gw; p < count; p += W) {
        const uint64_t s = pieces[3 * p], f = pieces[3 * p + 1], len = pieces[3 * p + 2];
        const int nsteps = (int)((len + GZ_WAVE - 1) / GZ_WAVE);
        bool ok[kSteps];
        uint64_t vi[kSteps], src[kSteps];
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
            const uint64_t slot = (uint64_t)(u * GZ_WAVE + lane);
            ok[u] = slot < len;
            vi[u] = f + slot;
            src[u] = s;
        }
        gz_push_steps<kSteps>(vecs, op, nsteps, ok, vi, src, vote, feat, vectors, edges);
    }
    gz_push_finish(op, wave_base + gw, vote, feat, vectors, edges, partials, stats, collect);
}

/* Whether a push over the frontier that a vote of `hint` built may skip the
 * compaction: out-degrees must be in the vote, so that it bounds the
 * frontier's vectors (every frontier source with vectors added at least its
 * out-degree), and the push must be frontier-driven. */
static inline bool gz_push_sparse_eligible(uint64_t hint)
{
    return gz_phase_config.push_frontier_skip && gz_graph.push_vertex_offsets != nullptr &&
           gz_phase_config.stat_mode != GZ_STAT_COUNT_ONLY && hint <= GZ_SPARSE_PUSH_VOTE &&
           gz_graph.num_vertices > 0;
}

/* Words of the bitmap one wavefront of the sparse push reads for a vote of
 * `hint`: a tiny frontier pays mostly for the grid; a larger one for the
 * wavefront that holds the most sources, so it gets fewer words each and the
 * overflow queue. */
static inline uint64_t gz_push_sparse_words(uint64_t hint) { return hint <= GZ_SPARSE_PUSH_WIDE_VOTE ? 64 : 16; }

/* Per-wave partials the sparse push writes for a vote of `hint`. */
static inline uint64_t gz_push_sparse_waves(uint64_t hint)
{
    const uint64_t words = gz_push_sparse_words(hint);
    const uint64_t waves = (((gz_graph.num_vertices + 63) >> 6) + words - 1) / words;
    const uint64_t blocks = (waves + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    const uint64_t qblocks = (words == 64) ? 0 : GZ_PERSISTENT_BLOCKS / 4;
    return (blocks + qblocks) * GZ_WAVES_PER_BLOCK;
}

/* Launch the sparse push for a vote of `hint` (gz_push_sparse_eligible).  Up
 * to GZ_SPARSE_PUSH_WIDE_VOTE one wavefront may walk every vector itself;
 * up to GZ_SPARSE_PUSH_VOTE the pieces (at most one per source plus one
 * per GZ_PUSH_PIECE_VECTORS vectors) fit the overflow queue.  Sets the
 * per-wave partials and statistics counts; `stamp`, when set, receives the
 * device clock as the push starts. */
template <class Op>
static void gz_launch_push_sparse(const Op& op, const gz_vec_t* edge_list, uint64_t hint, uint64_t* stamp)
{
    const uint64_t V = gz_graph.num_vertices;
    const uint64_t words = gz_push_sparse_words(hint);
    const bool wide = words == 64;
    const uint64_t waves = (((V + 63) >> 6) + words - 1) / words;
    const uint64_t blocks = (waves + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    const uint64_t qblocks = wide ? 0 : GZ_PERSISTENT_BLOCKS / 4;
    const uint64_t sw = blocks * GZ_WAVES_PER_BLOCK, all = sw + qblocks * GZ_WAVES_PER_BLOCK;
    gz_scratch_reserve(all, V);
    gz_scratch.partials_is_double = 0;
    gz_scratch.partials_count = all;
    gz_scratch.stats_count = gz_phase_config.collect_stats ? all : 0;
    uint64_t* partials = (uint64_t*)gz_scratch.partials;
    const uint32_t parity = gz_scratch.push_queue_parity;
#define GZ_PUSH_SPARSE_ARGS                                                                                   \
    dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, op, edge_list, gz_graph.push_vertex_offsets,           \
        gz_graph.frontier_has_info, V, gz_scratch.push_pieces, gz_scratch.push_queue + parity,                    \
        (uint64_t)gz_push_queue_cap, partials, gz_scratch.stats, gz_phase_config.collect_stats, stamp,             \
        gz_guard{nullptr, 0}
    if (wide) {
        hipLaunchKernelGGL((gz_k_push_sparse<Op, 64, false>), GZ_PUSH_SPARSE_ARGS);
        GZ_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL((gz_k_push_sparse<Op, 16, true>), GZ_PUSH_SPARSE_ARGS);
        GZ_LAUNCH_CHECK();
        hipLaunchKernelGGL((gz_k_push_pieces<Op>), dim3((unsigned)qblocks), dim3(GZ_BLOCK), 0, gz_stream, op,
                           edge_list, gz_scratch.push_pieces, gz_scratch.push_queue, (uint64_t)gz_push_queue_cap,
                           parity, sw, partials,
                           gz_scratch.stats, gz_phase_config.collect_stats, gz_guard{nullptr, 0});
        GZ_LAUNCH_CHECK();
        gz_scratch.push_queue_parity = parity ^ 1u;
    }
#undef GZ_PUSH_SPARSE_ARGS
}

/* Wavefronts of the small-frontier push at each width for the loaded graph:
 * 64 words per wavefront, and 16 words per wavefront plus the queue's. */
static inline void gz_push_sparse_loop_waves(uint64_t* wide_waves, uint64_t* narrow_waves)
{
    const uint64_t words = (gz_graph.num_vertices + 63) >> 6;
    const uint64_t wblocks = ((words + 63) / 64 + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    const uint64_t nblocks = ((words + 15) / 16 + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    *wide_waves = (wblocks ? wblocks : 1) * GZ_WAVES_PER_BLOCK;
    *narrow_waves = ((nblocks ? nblocks : 1) + GZ_PERSISTENT_BLOCKS / 4) * GZ_WAVES_PER_BLOCK;
}

/* Both widths of the small-frontier push, each under its own guard (for a
 * loop that picks the width on the device from the vote).  The queue
 * counters must be zero when a narrow push starts: the loop's control kernel
 * clears them after every iteration. */
template <class Op>
static void gz_launch_push_sparse_guarded(const Op& op, const gz_vec_t* edge_list, const uint64_t* has_info,
                                          gz_guard wide, gz_guard narrow)
{
    const uint64_t V = gz_graph.num_vertices;
    uint64_t ww, nw;
    gz_push_sparse_loop_waves(&ww, &nw);
    uint64_t* partials = (uint64_t*)gz_scratch.partials;
    hipLaunchKernelGGL((gz_k_push_sparse<Op, 64, false>), dim3((unsigned)(ww / GZ_WAVES_PER_BLOCK)), dim3(GZ_BLOCK), 0,
                       gz_stream, op, edge_list, gz_graph.push_vertex_offsets, has_info, V, gz_scratch.push_pieces,
                       gz_scratch.push_queue, (uint64_t)gz_push_queue_cap, partials, gz_scratch.stats, 0u,
                       (uint64_t*)nullptr, wide);
    GZ_LAUNCH_CHECK();
    const uint64_t qblocks = GZ_PERSISTENT_BLOCKS / 4;
    const uint64_t sblocks = nw / GZ_WAVES_PER_BLOCK - qblocks;
    hipLaunchKernelGGL((gz_k_push_sparse<Op, 16, true>), dim3((unsigned)sblocks), dim3(GZ_BLOCK), 0, gz_stream, op,
                       edge_list, gz_graph.push_vertex_offsets, has_info, V, gz_scratch.push_pieces,
                       gz_scratch.push_queue, (uint64_t)gz_push_queue_cap, partials, gz_scratch.stats, 0u,
                       (uint64_t*)nullptr, narrow);
    GZ_LAUNCH_CHECK();
    hipLaunchKernelGGL((gz_k_push_pieces<Op>), dim3((unsigned)qblocks), dim3(GZ_BLOCK), 0, gz_stream, op, edge_list,
                       gz_scratch.push_pieces, gz_scratch.push_queue, (uint64_t)gz_push_queue_cap, 0u,
                       sblocks * GZ_WAVES_PER_BLOCK, partials, gz_scratch.stats, 0u, narrow);
    GZ_LAUNCH_CHECK();
}

#endif /* GZ_SPARSE_PUSH_H */
