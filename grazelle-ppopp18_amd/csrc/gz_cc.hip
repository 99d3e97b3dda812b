/*
 * gz_cc.hip -- Connected Components phases on gfx950 (min-label propagation
 * along directed edges, uint64 labels).
 *
 *   perform_edge_pull_phase_cc  phase_edge_pull_cc.asm:398-529 (step :50-325, final write :357-392)
 *   perform_edge_push_phase_cc  phase_edge_push_cc.asm:579-673 (step :51-573)
 *   perform_vertex_phase_cc     phase_vertex_cc.asm:64-170
 *
 * Pull: the same wavefront chunking and segmented scan as the PageRank pull,
 * with min in place of sum.  Every run of a destination (including the runs
 * cut by chunk boundaries) is applied with a 64-bit atomic min, the GPU form
 * of the asm's lock cmpxchg at each destination change; a label that
 * actually drops sets the destination's HasInfo* bit and adds
 * phase_helper_add_frontier_stat to the wavefront's vote.  Gathers read the
 * labels in place, so updates made earlier in the sweep propagate within it,
 * as in the reference.
 *
 * Push: only sources in HasInfo send (the strong frontier check).  The
 * frontier is compacted first and the kernel walks the active sources'
 * vectors with all lanes busy; each valid lane applies an atomic min of the
 * source's label (read once per vector, as the asm reads props[src] once).
 * When the driver reports a small frontier statistic, one kernel reads the
 * HasInfo words directly instead (gz_k_edge_push_cc_sparse), with the same
 * per-edge semantics.
 */
#include "gz_slots.h"
#include "gz_walk.h"

__device__ __forceinline__ void gz_cc_apply(uint64_t* __restrict__ labels, uint64_t* __restrict__ next_frontier,
                                            const double* __restrict__ outdeg, uint32_t stat_mode, uint64_t d,
                                            uint64_t m, uint64_t& vote)
{
    if (m < labels[d]) {
        const unsigned long long old = atomicMin((unsigned long long*)&labels[d], (unsigned long long)m);
        if (old > m) {
            gz_bit_set(next_frontier, d);
            vote += gz_frontier_stat(outdeg, d, stat_mode);
        }
    }
}

struct gz_cc_pull_op {
    uint64_t* __restrict__ labels;
    uint64_t* __restrict__ next_frontier;
    const double* __restrict__ outdeg;
    uint32_t stat_mode;
    uint64_t vote;
    __device__ __forceinline__ uint64_t identity() const { return ~0ull; }
    __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) const { return a < b ? a : b; }
    __device__ __forceinline__ uint64_t gather(const gz_lanes& v, bool in_range, uint32_t& edges) const
    {
        /* branch-free gathers: invalid lanes read element 0 and are ignored */
        const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
        uint64_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = labels[(in_range && gz_lane_valid(l[k])) ? gz_lane_id(l[k]) : 0];
        uint64_t m = ~0ull;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ok = in_range && gz_lane_valid(l[k]);
            m = (ok && x[k] < m) ? x[k] : m;
            edges += ok ? 1u : 0u;
        }
        return m;
    }
    __device__ __forceinline__ void store(uint64_t d, uint64_t m)
    {
        gz_cc_apply(labels, next_frontier, outdeg, stat_mode, d, m, vote);
    }
};

template <int VPL>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_cc(const gz_vec_t* __restrict__ vecs, uint64_t nvec,
                                                              uint64_t chunk_len, uint64_t nchunks,
                                                              uint64_t* __restrict__ labels,
                                                              uint64_t* __restrict__ next_frontier,
                                                              const double* __restrict__ outdeg, uint32_t stat_mode,
                                                              uint64_t* __restrict__ partials,
                                                              uint64_t* __restrict__ stats, uint32_t collect)
{
    const int lane = gz_lane();
    const uint64_t chunk = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    if (chunk >= nchunks) return;
    const uint64_t b = chunk * chunk_len;
    const uint64_t e = (b + chunk_len < nvec) ? b + chunk_len : nvec;
    gz_cc_pull_op op{labels, next_frontier, outdeg, stat_mode, 0};
    uint64_t first_d, carry_d, carry_m;
    uint32_t edges = 0;
    gz_walk_chunk<VPL, uint64_t>(vecs, b, e, op, first_d, carry_d, carry_m, edges);
    if (lane == 0) op.store(carry_d, carry_m);
    const uint64_t total = gz_wave_sum_u64(op.vote);
    if (lane == 0) partials[chunk] = total;
    if (collect) {
        const uint64_t tot_edges = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * chunk] = e - b;
            stats[2 * chunk + 1] = tot_edges;
        }
    }
}

extern "C" void perform_edge_pull_phase_cc(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    if (gz_phase_config.pull_variant == GZ_PULL_SOURCE_SLICED && edge_list_count > 0) {
        gz_pull_cc_sliced(edge_list, edge_list_count);
        return;
    }
    const uint64_t nchunks = gz_chunks_for(edge_list_count);
    gz_scratch_reserve(nchunks, gz_graph.num_vertices);
    gz_scratch.partials_is_double = 0;
    gz_scratch.partials_count = nchunks;
    gz_scratch.stats_count = gz_phase_config.collect_stats ? nchunks : 0;
    if (nchunks == 0) return;
    const uint64_t chunk_len = gz_chunk_len();
    const uint64_t blocks = (nchunks + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    const uint32_t vpl = gz_vectors_per_lane();
#define GZ_PULL_CC_ARGS                                                                                       \
    dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list, edge_list_count, chunk_len, nchunks,      \
        (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators, gz_graph.vertex_outdegrees, \
        gz_phase_config.stat_mode, (uint64_t*)gz_scratch.partials, gz_scratch.stats, gz_phase_config.collect_stats
    gz_ktimer_begin(GZ_K_PULL_CC);
    if (vpl == 4)
        hipLaunchKernelGGL(gz_k_edge_pull_cc<4>, GZ_PULL_CC_ARGS);
    else if (vpl == 2)
        hipLaunchKernelGGL(gz_k_edge_pull_cc<2>, GZ_PULL_CC_ARGS);
    else
        hipLaunchKernelGGL(gz_k_edge_pull_cc<1>, GZ_PULL_CC_ARGS);
#undef GZ_PULL_CC_ARGS
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PULL_CC);
}

template <bool kActive>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_push_cc(const gz_vec_t* __restrict__ vecs, gz_slot_source src,
                                                              uint64_t chunk_len, const uint64_t* __restrict__ has_info,
                                                              uint64_t* __restrict__ labels,
                                                              uint64_t* __restrict__ next_frontier,
                                                              const double* __restrict__ outdeg, uint32_t stat_mode,
                                                              uint64_t* __restrict__ partials,
                                                              uint64_t* __restrict__ stats, uint32_t collect)
{
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t W = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    const uint64_t T = gz_slot_total<kActive>(src);
    uint64_t vote = 0;
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
            const uint64_t msg = labels[s];
            const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (gz_lane_valid(l[k])) {
                    edges += 1;
                    gz_cc_apply(labels, next_frontier, outdeg, stat_mode, gz_lane_id(l[k]), msg, vote);
                }
            }
        }
    }
    const uint64_t total = gz_wave_sum_u64(vote);
    if (lane == 0) partials[gw] = total;
    if (collect) {
        const uint64_t tv = gz_wave_sum_u64(vectors), te = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * gw] = tv;
            stats[2 * gw + 1] = te;
        }
    }
}

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
 * then all label reads, so a lane keeps 4 kSteps label reads in flight.
 * Lanes without ok[u] read vector 0 and label 0 and are ignored. */
template <int kSteps>
__device__ __forceinline__ void gz_cc_push_steps(const gz_vec_t* __restrict__ vecs, uint64_t* __restrict__ labels,
                                                 uint64_t* __restrict__ next_frontier,
                                                 const double* __restrict__ outdeg, uint32_t stat_mode, int nsteps,
                                                 const bool* ok, const uint64_t* vi, const uint64_t* src,
                                                 uint64_t& vote, uint32_t& vectors, uint32_t& edges)
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
    uint64_t msg[kSteps], x[kSteps][4];
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        msg[u] = ~0ull;
        x[u][0] = x[u][1] = x[u][2] = x[u][3] = 0;
        if (u >= nsteps) continue;
        msg[u] = labels[ok[u] ? src[u] : 0];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[u][k] = labels[(ok[u] && gz_lane_valid(l[u][k])) ? gz_lane_id(l[u][k]) : 0];
    }
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        if (u >= nsteps || !ok[u]) continue;
        vectors += 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!gz_lane_valid(l[u][k])) continue;
            edges += 1;
            const uint64_t d = gz_lane_id(l[u][k]);
            if (msg[u] < x[u][k]) {
                const unsigned long long old = atomicMin((unsigned long long*)&labels[d], (unsigned long long)msg[u]);
                if (old > msg[u]) {
                    gz_bit_set(next_frontier, d);
                    vote += gz_frontier_stat(outdeg, d, stat_mode);
                }
            }
        }
    }
}

/* Push for a small frontier, with no compaction pass.  Each wavefront reads
 * kWords HasInfo words (64 vertices each) and numbers their set bits with a
 * scan of the word popcounts.  Rounds of 64 sources follow: each lane selects
 * one source (owner word by binary search over the scan, then the n-th set
 * bit), and a scan of the sources' vector counts spreads their vectors over
 * the 64 lanes, each lane finding its source the same way, in batches of
 * four steps of 64 vectors.  With kQueue a round of more than one batch is
 * cut instead into pieces of at most GZ_PUSH_PIECE_VECTORS vectors,
 * appended to the overflow queue (counter *queue) for
 * gz_k_edge_push_cc_pieces, so that no wavefront walks a high-degree source
 * alone; without kQueue every round is pushed here. */
template <int kWords, bool kQueue>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_push_cc_sparse(
    const gz_vec_t* __restrict__ vecs, const uint64_t* __restrict__ offsets, const uint64_t* __restrict__ has_info,
    uint64_t num_vertices, uint64_t* __restrict__ labels, uint64_t* __restrict__ next_frontier,
    const double* __restrict__ outdeg, uint32_t stat_mode, uint64_t* __restrict__ pieces, uint64_t* __restrict__ queue,
    uint64_t* __restrict__ partials, uint64_t* __restrict__ stats, uint32_t collect)
{
    constexpr int kSteps = GZ_PUSH_PIECE_VECTORS / GZ_WAVE;
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t nwords = (num_vertices + 63) >> 6;
    const uint64_t w = gw * kWords + (uint64_t)lane;
    const bool word_lane = lane < kWords && w < nwords;
    uint64_t bits = word_lane ? has_info[w] : 0;
    if (word_lane && w == nwords - 1 && (num_vertices & 63)) bits &= (1ull << (num_vertices & 63)) - 1;
    const uint64_t pc = (uint64_t)__popcll(bits);
    const uint64_t pincl = gz_wave_inclusive_sum(pc, lane);
    const uint64_t sources = gz_readlane_u64(pincl, GZ_WAVE - 1);
    uint64_t vote = 0;
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
                if (slot < GZ_PUSH_QUEUE_PIECES) {
                    pieces[3 * slot] = v;
                    pieces[3 * slot + 1] = pf;
                    pieces[3 * slot + 2] = pl;
                } else {
                    /* queue full (more pieces than a frontier within the
                     * sparse bound can make): this lane pushes the piece */
                    const uint64_t m = labels[v];
                    for (uint64_t j = 0; j < pl; ++j) {
                        const gz_lanes vec = gz_load_vector(vecs, pf + j);
                        const uint64_t lj[4] = {vec.l0, vec.l1, vec.l2, vec.l3};
                        vectors += 1;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            if (!gz_lane_valid(lj[q])) continue;
                            edges += 1;
                            gz_cc_apply(labels, next_frontier, outdeg, stat_mode, gz_lane_id(lj[q]), m, vote);
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
            gz_cc_push_steps<kSteps>(vecs, labels, next_frontier, outdeg, stat_mode, nsteps, ok, vi, src, vote,
                                     vectors, edges);
        }
    }
    const uint64_t tot = gz_wave_sum_u64(vote);
    if (lane == 0) partials[gw] = tot;
    if (collect) {
        const uint64_t tv = gz_wave_sum_u64(vectors), te = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * gw] = tv;
            stats[2 * gw + 1] = te;
        }
    }
}

/* The overflow queue of gz_k_edge_push_cc_sparse: wavefront gw pushes pieces
 * gw, gw + W, ... of the queue[parity] appended there.  It also clears the
 * other counter, which the next sparse push appends to (the kernel that read
 * that counter last has finished: launches on one stream run in order). */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_push_cc_pieces(
    const gz_vec_t* __restrict__ vecs, const uint64_t* __restrict__ pieces, uint64_t* __restrict__ queue,
    uint32_t parity, uint64_t* __restrict__ labels, uint64_t* __restrict__ next_frontier,
    const double* __restrict__ outdeg, uint32_t stat_mode, uint64_t* __restrict__ partials,
    uint64_t* __restrict__ stats, uint32_t collect)
{
    constexpr int kSteps = GZ_PUSH_PIECE_VECTORS / GZ_WAVE;
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t W = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    if (gw == 0 && lane == 0) queue[parity ^ 1u] = 0;
    uint64_t count = queue[parity];
    if (count > GZ_PUSH_QUEUE_PIECES) count = GZ_PUSH_QUEUE_PIECES;
    uint64_t vote = 0;
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
        gz_cc_push_steps<kSteps>(vecs, labels, next_frontier, outdeg, stat_mode, nsteps, ok, vi, src, vote, vectors,
                                 edges);
    }
    const uint64_t tot = gz_wave_sum_u64(vote);
    if (lane == 0) partials[gw] = tot;
    if (collect) {
        const uint64_t tv = gz_wave_sum_u64(vectors), te = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * gw] = tv;
            stats[2 * gw + 1] = te;
        }
    }
}

/* Frontier statistic of the next push, when the driver knows it (the vote of
 * the iteration that built the frontier, which counts every activated source
 * at least once); ~0 when unknown.  Consumed by the next push. */
static uint64_t gz_cc_next_push_vote = ~0ull;

extern "C" void gz_cc_push_vote_hint(uint64_t vote) { gz_cc_next_push_vote = vote; }

/* Launch a frontier-driven push over the push list for sources in has_info. */
static void gz_cc_launch_push(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    const uint64_t hint = gz_cc_next_push_vote;
    gz_cc_next_push_vote = ~0ull;
    /* With out-degrees in the vote, the vote bounds the frontier's vectors
     * (every frontier source with vectors added at least its out-degree).  Up
     * to GZ_CC_SPARSE_PUSH_WIDE_VOTE one wavefront may walk them all; up to
     * GZ_CC_SPARSE_PUSH_VOTE their pieces (at most one per source plus one per
     * GZ_PUSH_PIECE_VECTORS vectors) fit the overflow queue. */
    if (gz_phase_config.push_frontier_skip && gz_graph.push_vertex_offsets != nullptr &&
        gz_phase_config.stat_mode != GZ_STAT_COUNT_ONLY && hint <= GZ_CC_SPARSE_PUSH_VOTE &&
        gz_graph.num_vertices > 0) {
        const uint64_t V = gz_graph.num_vertices;
        /* a tiny frontier pays mostly for the grid; a larger one for the
         * wavefront that holds the most sources, so it gets fewer words each
         * and the overflow queue */
        const bool wide = hint <= GZ_CC_SPARSE_PUSH_WIDE_VOTE;
        const uint64_t words = wide ? 64 : 16;
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
        gz_ktimer_begin(GZ_K_PUSH_CC);
#define GZ_PUSH_SPARSE_ARGS                                                                                     \
    dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list, gz_graph.push_vertex_offsets,               \
        gz_graph.frontier_has_info, V, (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators, \
        gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, gz_scratch.push_pieces,                             \
        gz_scratch.push_queue + parity, partials, gz_scratch.stats, gz_phase_config.collect_stats
        if (wide) {
            hipLaunchKernelGGL((gz_k_edge_push_cc_sparse<64, false>), GZ_PUSH_SPARSE_ARGS);
        } else {
            hipLaunchKernelGGL((gz_k_edge_push_cc_sparse<16, true>), GZ_PUSH_SPARSE_ARGS);
            GZ_LAUNCH_CHECK();
            hipLaunchKernelGGL(gz_k_edge_push_cc_pieces, dim3((unsigned)qblocks), dim3(GZ_BLOCK), 0, gz_stream,
                               edge_list, gz_scratch.push_pieces, gz_scratch.push_queue, parity,
                               (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators,
                               gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, partials + sw,
                               gz_scratch.stats + 2 * sw, gz_phase_config.collect_stats);
            gz_scratch.push_queue_parity = parity ^ 1u;
        }
#undef GZ_PUSH_SPARSE_ARGS
        GZ_LAUNCH_CHECK();
        gz_ktimer_end(GZ_K_PUSH_CC);
        return;
    }
    const uint64_t chunk_len = 256;
    gz_slot_source src;
    src.offsets = gz_graph.push_vertex_offsets;
    src.list = gz_scratch.active_vertices;
    src.prefix = gz_scratch.active_prefix;
    src.totals = gz_scratch.active_totals;
    src.sweep_count = edge_list_count;
    const bool active = gz_phase_config.push_frontier_skip && gz_graph.push_vertex_offsets != nullptr;
    uint64_t blocks;
    if (active) {
        gz_frontier_compact(gz_graph.frontier_has_info, gz_graph.push_vertex_offsets, gz_graph.num_vertices);
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
    gz_ktimer_begin(GZ_K_PUSH_CC);
    if (active)
        hipLaunchKernelGGL(gz_k_edge_push_cc<true>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list, src,
                           chunk_len, gz_graph.frontier_has_info, (uint64_t*)gz_graph.vertex_props,
                           (uint64_t*)gz_graph.vertex_accumulators, gz_graph.vertex_outdegrees,
                           gz_phase_config.stat_mode, (uint64_t*)gz_scratch.partials, gz_scratch.stats,
                           gz_phase_config.collect_stats);
    else
        hipLaunchKernelGGL(gz_k_edge_push_cc<false>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list,
                           src, chunk_len, gz_graph.frontier_has_info, (uint64_t*)gz_graph.vertex_props,
                           (uint64_t*)gz_graph.vertex_accumulators, gz_graph.vertex_outdegrees,
                           gz_phase_config.stat_mode, (uint64_t*)gz_scratch.partials, gz_scratch.stats,
                           gz_phase_config.collect_stats);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PUSH_CC);
}

extern "C" void perform_edge_push_phase_cc(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    gz_scratch_reserve(0, gz_graph.num_vertices);
    gz_cc_launch_push(edge_list, edge_list_count);
}

/* Clear the HasInfo* words covering [vertex_start, vertex_start+count)
 * (phase_vertex_cc.asm:64-170 zeroes 512 vertices per step). */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_clear_words(uint64_t* __restrict__ words, uint64_t first, uint64_t end)
{
    for (uint64_t w = first + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < end; w += (uint64_t)gridDim.x * blockDim.x)
        words[w] = 0;
}

extern "C" void gz_clear_bitmap_range(uint64_t* bitmap, uint64_t vertex_start, uint64_t vertex_count)
{
    if (vertex_count == 0) return;
    const uint64_t first = vertex_start >> 6;
    const uint64_t end = (vertex_start + vertex_count + 63) >> 6;
    uint64_t blocks = (end - first + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gz_k_clear_words, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, bitmap, first, end);
    GZ_LAUNCH_CHECK();
}

extern "C" void perform_vertex_phase_cc(uint64_t vertex_start, uint64_t vertex_count, const uint64_t* reduce_buffer)
{
    (void)reduce_buffer;
    gz_clear_bitmap_range((uint64_t*)gz_graph.vertex_accumulators, vertex_start, vertex_count);
}
