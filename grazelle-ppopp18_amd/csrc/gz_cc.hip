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
 * With the merge-buffer variant (GZ_PULL_MERGE_BUFFER) complete runs use
 * plain stores and each chunk's last run goes through its merge record,
 * merged with min after the sweep, as EXPERIMENT_EDGE_PULL_FORCE_MERGE does.
 *
 * Push: only sources in HasInfo send (the strong frontier check).  The
 * frontier is compacted first and the kernel walks the active sources'
 * vectors with all lanes busy; each valid lane applies an atomic min of the
 * source's label (read once per vector, as the asm reads props[src] once).
 * When the driver reports a small frontier statistic, one kernel reads the
 * HasInfo words directly instead (gz_sparse_push.h), with the same per-edge
 * semantics.
 */
#include "grazelle_host.h"
#include "gz_host_internal.h"
#include "gz_slots.h"
#include "gz_sparse_push.h"
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

/* ---- the merge-buffer pull (EXPERIMENT_EDGE_PULL_FORCE_MERGE) ----------------
 * phase_edge_pull_cc.asm:136-181 and :328-354, execution_cc.c:165-181: a
 * run that is complete inside its unit lowers the label with a plain store
 * (no other unit holds that destination); the unit's last run goes to its
 * merge-buffer record as min(run, current label), setting HasInfo* and
 * counting the frontier statistic when it is lower; afterwards the records
 * are merged with min (edge_pull_op_merge_with_merge_buffer with
 * SCALAR_REDUCE_OP = min), including the label the next unit stored
 * directly when it starts on the same destination. */
struct gz_cc_merge_op {
    uint64_t* __restrict__ labels;
    uint64_t* __restrict__ next_frontier;
    const double* __restrict__ outdeg;
    uint32_t stat_mode;
    uint64_t vote;
    __device__ __forceinline__ uint64_t identity() const { return ~0ull; }
    __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) const { return a < b ? a : b; }
    __device__ __forceinline__ uint64_t gather(const gz_lanes& v, bool in_range, uint32_t& edges) const
    {
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
        if (m < labels[d]) {
            labels[d] = m;
            gz_bit_set(next_frontier, d);
            vote += gz_frontier_stat(outdeg, d, stat_mode);
        }
    }
};

template <int VPL>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_cc_merge(const gz_vec_t* __restrict__ vecs, uint64_t nvec,
                                                                    uint64_t chunk_len, uint64_t nchunks,
                                                                    uint64_t* __restrict__ labels,
                                                                    uint64_t* __restrict__ next_frontier,
                                                                    const double* __restrict__ outdeg,
                                                                    uint32_t stat_mode, mergeaccum_t* __restrict__ merge,
                                                                    uint64_t* __restrict__ partials,
                                                                    uint64_t* __restrict__ stats, uint32_t collect)
{
    const int lane = gz_lane();
    const uint64_t chunk = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    if (chunk >= nchunks) return;
    const uint64_t b = chunk * chunk_len;
    const uint64_t e = (b + chunk_len < nvec) ? b + chunk_len : nvec;
    gz_cc_merge_op op{labels, next_frontier, outdeg, stat_mode, 0};
    uint64_t first_d, carry_d, carry_m;
    uint32_t edges = 0;
    gz_walk_chunk<VPL, uint64_t>(vecs, b, e, op, first_d, carry_d, carry_m, edges);
    if (lane == 0) {
        const uint64_t cur = labels[carry_d];
        const uint64_t m = carry_m < cur ? carry_m : cur;
        if (m < cur) {
            gz_bit_set(next_frontier, carry_d);
            op.vote += gz_frontier_stat(outdeg, carry_d, stat_mode);
        }
        mergeaccum_t r;
        r.initial_vertex_id = first_d;
        r.final_vertex_id = carry_d;
        r.final_partial_value = __longlong_as_double((long long)m);
        r.__padding__ = 0;
        merge[chunk] = r;
    }
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

/* The records' merge with min (phases.c:41-70), one thread per run head. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_merge_min(const mergeaccum_t* __restrict__ mb, uint64_t count,
                                                           uint64_t* __restrict__ labels)
{
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= count) return;
    const mergeaccum_t r = mb[u];
    if (u > 0 && mb[u - 1].final_vertex_id == r.final_vertex_id) return;
    uint64_t proposed = (uint64_t)__double_as_longlong(r.final_partial_value);
    uint64_t j = u + 1;
    while (j < count && mb[j].final_vertex_id == r.final_vertex_id) {
        const uint64_t x = (uint64_t)__double_as_longlong(mb[j].final_partial_value);
        proposed = x < proposed ? x : proposed;
        ++j;
    }
    if (j < count && mb[j].initial_vertex_id == r.final_vertex_id) {
        const uint64_t x = labels[mb[j].initial_vertex_id];
        proposed = x < proposed ? x : proposed;
    }
    labels[r.final_vertex_id] = proposed;
}

static void gz_pull_cc_merge(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    const uint64_t nchunks = gz_chunks_for(edge_list_count);
    if (gz_merge_buffer_count < nchunks) {
        fprintf(stderr, "grazelle: merge buffer holds %llu units, the pull phase needs %llu\n",
                (unsigned long long)gz_merge_buffer_count, (unsigned long long)nchunks);
        exit(72);
    }
    const uint64_t chunk_len = gz_chunk_len();
    const uint64_t blocks = (nchunks + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    const uint32_t vpl = gz_vectors_per_lane();
    uint64_t* labels = (uint64_t*)gz_graph.vertex_props;
#define GZ_PULL_CC_MERGE_ARGS                                                                                 \
    dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list, edge_list_count, chunk_len, nchunks,      \
        labels, (uint64_t*)gz_graph.vertex_accumulators, gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, \
        gz_merge_buffer, (uint64_t*)gz_scratch.partials, gz_scratch.stats, gz_phase_config.collect_stats
    gz_ktimer_begin(GZ_K_PULL_CC);
    if (vpl == 4)
        hipLaunchKernelGGL(gz_k_edge_pull_cc_merge<4>, GZ_PULL_CC_MERGE_ARGS);
    else if (vpl == 2)
        hipLaunchKernelGGL(gz_k_edge_pull_cc_merge<2>, GZ_PULL_CC_MERGE_ARGS);
    else
        hipLaunchKernelGGL(gz_k_edge_pull_cc_merge<1>, GZ_PULL_CC_MERGE_ARGS);
#undef GZ_PULL_CC_MERGE_ARGS
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PULL_CC);
    gz_ktimer_begin(GZ_K_MERGE);
    hipLaunchKernelGGL(gz_k_merge_min, dim3((unsigned)((nchunks + GZ_BLOCK - 1) / GZ_BLOCK)), dim3(GZ_BLOCK), 0,
                       gz_stream, gz_merge_buffer, nchunks, labels);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_MERGE);
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
    if (gz_phase_config.pull_variant == GZ_PULL_MERGE_BUFFER) {
        gz_pull_cc_merge(edge_list, edge_list_count);
        return;
    }
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
                                                              uint64_t* __restrict__ stats, uint32_t collect,
                                                              gz_guard guard)
{
    if (gz_guard_off(guard)) return;
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

/* The CC push's per-edge work for the sparse push (gz_sparse_push.h): the
 * source's label, the destination's label read ahead, and the atomic min. */
struct gz_cc_push_op {
    uint64_t* labels;
    uint64_t* next_frontier;
    const double* outdeg;
    uint32_t stat_mode;
    __device__ __forceinline__ uint64_t message(uint64_t s) const { return labels[s]; }
    __device__ __forceinline__ void gather(uint64_t d, uint64_t* g) const { g[0] = labels[d]; }
    __device__ __forceinline__ void apply(uint64_t d, uint64_t m, const uint64_t* g, uint64_t& vote,
                                          gz_feature_acc&) const
    {
        if (m < g[0]) {
            const unsigned long long old = atomicMin((unsigned long long*)&labels[d], (unsigned long long)m);
            if (old > m) {
                gz_bit_set(next_frontier, d);
                vote += gz_frontier_stat(outdeg, d, stat_mode);
            }
        }
    }
    __device__ __forceinline__ void finish(uint64_t, const gz_feature_acc&) const {}
};

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
    if (gz_push_sparse_eligible(hint)) {
        const gz_cc_push_op op{(uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators,
                               gz_graph.vertex_outdegrees, gz_phase_config.stat_mode};
        gz_ktimer_begin(GZ_K_PUSH_CC);
        gz_launch_push_sparse(op, edge_list, hint, nullptr);
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
    gz_ktimer_begin(GZ_K_PUSH_CC);
    if (active)
        hipLaunchKernelGGL(gz_k_edge_push_cc<true>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list, src,
                           chunk_len, gz_graph.frontier_has_info, (uint64_t*)gz_graph.vertex_props,
                           (uint64_t*)gz_graph.vertex_accumulators, gz_graph.vertex_outdegrees,
                           gz_phase_config.stat_mode, (uint64_t*)gz_scratch.partials, gz_scratch.stats,
                           gz_phase_config.collect_stats, gz_guard{nullptr, 0});
    else
        hipLaunchKernelGGL(gz_k_edge_push_cc<false>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list,
                           src, chunk_len, gz_graph.frontier_has_info, (uint64_t*)gz_graph.vertex_props,
                           (uint64_t*)gz_graph.vertex_accumulators, gz_graph.vertex_outdegrees,
                           gz_phase_config.stat_mode, (uint64_t*)gz_scratch.partials, gz_scratch.stats,
                           gz_phase_config.collect_stats, gz_guard{nullptr, 0});
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PUSH_CC);
}

extern "C" void perform_edge_push_phase_cc(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    gz_scratch_reserve(0, gz_graph.num_vertices);
    gz_cc_launch_push(edge_list, edge_list_count);
}

/* ---- push tail: small-frontier iterations in one workgroup -------------------
 * The last CC iterations of a graph with long paths (planted trees, R-MAT's
 * sparse fringe) push tiny frontiers, and each pays its launches and the
 * host's read of the vote more than its edges.  While the frontier statistic
 * stays at most `vote_cap` (<= GZ_CC_TAIL_CAP), gz_k_cc_push_tail runs these
 * push iterations ("rounds") back to back in one workgroup: round 1 pushes
 * the compacted HasInfo list, and each later round pushes the vertices whose
 * labels the round before lowered, listed in LDS.  Every edge gets the update
 * of the push (gz_cc_apply: atomic min, HasInfo* bit, frontier statistic per
 * lowering), so labels reach the same fixpoint as one launch per iteration.
 *
 * A round's list cannot outgrow the LDS: its vertices are targets of edges of
 * the round before, at most that frontier's out-degrees, which its statistic
 * counts and vote_cap bounds (the list check is kept as a guard).
 *
 * Between rounds the HasInfo* bitmap is also the set of vertices listed for
 * the next round: a round first clears its own vertices' bits, so a vertex
 * lowered again after its label was read is listed once more.  A source's
 * label is read with a device-scope atomic load after that clear: another
 * wavefront may have lowered it in L2 since this compute unit's L1 cached
 * the line, and the value sent must be the latest.  (The destination's label
 * is read plainly: a stale, higher value only costs one more atomic min.)
 *
 * The workgroup stops after a round that activates nothing, whose statistic
 * exceeds vote_cap (with vote_cap at most the hybrid threshold, the next
 * iteration would pull), or after max_rounds, and leaves what one push
 * iteration leaves: HasInfo* holds the last round's activations, out[0] its
 * statistic (the vote) and out[7] the rounds run. */
#define GZ_CC_TAIL_THREADS 1024
#define GZ_CC_TAIL_CAP 16384
#define GZ_CC_TAIL_SLOTS 1
/* Default entry bound.  Measured on CC over R-MAT s22 with planted trees
 * (MI355X): bounds of 1024 and 4096 give the same time to solution and 16384
 * a slower one, since a round of a few thousand vectors is faster on the
 * whole grid than in one workgroup, whose rounds are latency-bound. */
#define GZ_CC_TAIL_VOTE 4096

/* Words the tail's no-op atomics go to, GZ_CC_TAIL_SLOTS * 4 per thread.
 * Nothing changes them: an atomic min of ~0 and an atomic or of 0 leave a
 * word as it is, and their returned values are never used. */
__device__ unsigned long long gz_cc_tail_sink[GZ_CC_TAIL_THREADS * GZ_CC_TAIL_SLOTS * 4];

__global__ __launch_bounds__(GZ_CC_TAIL_THREADS) void gz_k_cc_push_tail(
    uint64_t* __restrict__ labels, uint64_t* __restrict__ next_frontier, const gz_vec_t* __restrict__ vecs,
    const uint64_t* __restrict__ offsets, const double* __restrict__ outdeg, uint32_t stat_mode,
    const uint64_t* __restrict__ first_list, const uint64_t* __restrict__ first_count, uint64_t vote_cap,
    uint64_t max_rounds, uint64_t* __restrict__ out, uint64_t* __restrict__ mirror)
{
    constexpr uint32_t T = GZ_CC_TAIL_THREADS;
    constexpr int kWaves = GZ_CC_TAIL_THREADS / GZ_WAVE;
    __shared__ uint32_t lists[2][GZ_CC_TAIL_CAP];
    __shared__ uint32_t incl[T];   /* inclusive vector counts of the chunk's sources */
    __shared__ uint32_t first[T];  /* their first push vector */
    __shared__ uint64_t msg[T];    /* their labels */
    __shared__ uint32_t srcv[T];   /* the sources */
    __shared__ uint64_t wave_tot[kWaves];
    __shared__ uint32_t next_count;
    __shared__ unsigned long long round_vote;
    const uint32_t tid = threadIdx.x;
    const int lane = gz_lane(), wave = (int)(tid / GZ_WAVE);
    uint64_t n = *first_count;
    uint32_t cur = 0;
    uint64_t rounds = 0, vote = 0;
    for (;;) {
        if (rounds > 0)
            for (uint64_t i = tid; i < n; i += T) gz_bit_clear(next_frontier, lists[cur][i]);
        if (tid == 0) {
            next_count = 0;
            round_vote = 0;
        }
        /* the clears are done in L2 before any wavefront sets a bit this round */
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        uint64_t my_vote = 0;
        for (uint64_t base = 0; base < n; base += T) {
            const uint64_t i = base + tid;
            uint64_t c = 0, f = 0, m = 0, v = 0;
            if (i < n) {
                v = rounds == 0 ? first_list[i] : (uint64_t)lists[cur][i];
                f = offsets[v];
                c = offsets[v + 1] - f;
                m = __hip_atomic_load(&labels[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const uint64_t wincl = gz_wave_inclusive_sum(c, lane);
            if (lane == GZ_WAVE - 1) wave_tot[wave] = wincl;
            __syncthreads();
            uint64_t before = 0, total = 0;
            for (int w = 0; w < kWaves; ++w) {
                const uint64_t x = wave_tot[w];
                if (w < wave) before += x;
                total += x;
            }
            incl[tid] = (uint32_t)(before + wincl);
            first[tid] = (uint32_t)f;
            msg[tid] = m;
            srcv[tid] = (uint32_t)v;
            __syncthreads();
            const uint32_t nsrc = (uint32_t)(n - base < T ? n - base : T);
            /* kSlots vectors per thread at a time.  The edge steps are
             * branch-free, so each step's loads and atomics are in flight
             * together: an edge whose update does not apply (or a lane
             * without an edge) sends its atomic min (of ~0) and atomic or
             * (of 0) to its own word of gz_cc_tail_sink, which no other lane
             * touches, instead of to a label that many edges share. */
            constexpr int kSlots = GZ_CC_TAIL_SLOTS;
            unsigned long long* sink = gz_cc_tail_sink + (uint64_t)tid * (kSlots * 4);
            for (uint64_t s0 = tid; s0 < total; s0 += (uint64_t)kSlots * T) {
                uint64_t l[kSlots][4], mm[kSlots];
                uint32_t sv[kSlots];
#pragma unroll
                for (int b = 0; b < kSlots; ++b) {
                    const uint64_t slot = s0 + (uint64_t)b * T;
                    l[b][0] = l[b][1] = l[b][2] = l[b][3] = 0;
                    mm[b] = ~0ull;
                    sv[b] = srcv[0];
                    if (slot >= total) continue;
                    /* the first source whose inclusive count exceeds the slot */
                    uint32_t lo = 0, hi = nsrc - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if ((uint64_t)incl[mid] > slot)
                            hi = mid;
                        else
                            lo = mid + 1;
                    }
                    const uint64_t excl = lo ? incl[lo - 1] : 0;
                    const gz_lanes vec = gz_load_vector(vecs, (uint64_t)first[lo] + (slot - excl));
                    l[b][0] = vec.l0;
                    l[b][1] = vec.l1;
                    l[b][2] = vec.l2;
                    l[b][3] = vec.l3;
                    mm[b] = msg[lo];
                    sv[b] = srcv[lo];
                }
                uint32_t d[kSlots][4]; /* vertex ids fit 32 bits here */
                uint64_t g[kSlots][4];
#pragma unroll
                for (int b = 0; b < kSlots; ++b)
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        d[b][k] = (uint32_t)(gz_lane_valid(l[b][k]) ? gz_lane_id(l[b][k]) : sv[b]);
                        g[b][k] = labels[d[b][k]];
                    }
                /* gz_frontier_stat for the modes with out-degrees (the
                 * only ones the tail runs with), its loads issued before
                 * the atomics (memory results return in order) */
                const uint64_t with_count = stat_mode != GZ_STAT_OUTDEG_ONLY ? 1 : 0;
                double od[kSlots][4];
#pragma unroll
                for (int b = 0; b < kSlots; ++b)
#pragma unroll
                    for (int k = 0; k < 4; ++k) od[b][k] = outdeg[d[b][k]];
                uint64_t old[kSlots][4];
                bool pass[kSlots][4];
#pragma unroll
                for (int b = 0; b < kSlots; ++b)
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        pass[b][k] = gz_lane_valid(l[b][k]) && mm[b] < g[b][k];
                        unsigned long long* at = pass[b][k] ? (unsigned long long*)&labels[d[b][k]] : sink + 4 * b + k;
                        old[b][k] = atomicMin(at, pass[b][k] ? (unsigned long long)mm[b] : ~0ull);
                    }
                unsigned int prev[kSlots][4];
                bool low[kSlots][4];
#pragma unroll
                for (int b = 0; b < kSlots; ++b)
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        low[b][k] = pass[b][k] && old[b][k] > mm[b];
                        const uint64_t st = with_count + (uint64_t)llrint(od[b][k]);
                        my_vote += low[b][k] ? st : 0;
                        unsigned int* w = low[b][k] ? reinterpret_cast<unsigned int*>(next_frontier) + (d[b][k] >> 5)
                                                    : reinterpret_cast<unsigned int*>(sink + 4 * b + k);
                        prev[b][k] = atomicOr(w, low[b][k] ? 1u << (d[b][k] & 31) : 0u);
                    }
#pragma unroll
                for (int b = 0; b < kSlots; ++b)
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        if (!low[b][k] || ((prev[b][k] >> (d[b][k] & 31)) & 1u)) continue;
                        const uint32_t at = atomicAdd(&next_count, 1u);
                        if (at < GZ_CC_TAIL_CAP) lists[cur ^ 1u][at] = (uint32_t)d[b][k];
                    }
            }
            __syncthreads(); /* the next chunk rewrites incl, first, msg, wave_tot */
        }
        const uint64_t wv = gz_wave_sum_u64(my_vote);
        if (lane == 0 && wv) atomicAdd(&round_vote, (unsigned long long)wv);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        rounds += 1;
        vote = round_vote;
        const uint64_t listed = next_count;
        if (vote == 0 || vote > vote_cap || listed > GZ_CC_TAIL_CAP || rounds >= max_rounds) break;
        n = listed;
        cur ^= 1u;
        __syncthreads(); /* every wavefront has read round_vote and next_count */
    }
    if (tid == 0) {
        out[0] = vote;
        out[7] = rounds;
        gz_mirror_words(mirror, out);
    }
}

static uint64_t gz_cc_tail_vote = GZ_CC_TAIL_VOTE;

extern "C" void gz_cc_push_tail(uint64_t max_vote)
{
    gz_cc_tail_vote = max_vote < GZ_CC_TAIL_CAP ? max_vote : GZ_CC_TAIL_CAP;
}

extern "C" int gz_cc_push_tail_applies(uint64_t vote)
{
    return gz_cc_tail_vote > 0 && vote <= gz_cc_tail_vote && gz_phase_config.push_frontier_skip &&
           gz_graph.push_vertex_offsets != nullptr && gz_graph.push_vectors != nullptr &&
           gz_phase_config.stat_mode != GZ_STAT_COUNT_ONLY && gz_graph.num_vertices > 0 &&
           gz_graph.num_vertices < (1ull << 32) && gz_graph.push_vector_count < (1ull << 32);
}

extern "C" void gz_cc_launch_push_tail(uint64_t vote_cap, uint64_t max_rounds, uint64_t* reduce_buffer)
{
    const uint64_t V = gz_graph.num_vertices;
    if (vote_cap > gz_cc_tail_vote) vote_cap = gz_cc_tail_vote;
    if (max_rounds == 0) max_rounds = 1;
    gz_scratch_reserve(0, V);
    gz_frontier_list(gz_graph.frontier_has_info, gz_graph.push_vertex_offsets, V, gz_graph.push_vector_count);
    uint64_t* mirror = gz_reduce_mirror_target(reduce_buffer);
    gz_ktimer_begin(GZ_K_PUSH_CC);
    hipLaunchKernelGGL(gz_k_cc_push_tail, dim3(1), dim3(GZ_CC_TAIL_THREADS), 0, gz_stream,
                       (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators, gz_graph.push_vectors,
                       gz_graph.push_vertex_offsets, gz_graph.vertex_outdegrees, gz_phase_config.stat_mode,
                       gz_scratch.active_vertices, gz_scratch.active_totals, vote_cap, max_rounds, reduce_buffer,
                       mirror);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PUSH_CC);
}

/* Clear the HasInfo* words covering [vertex_start, vertex_start+count)
 * (phase_vertex_cc.asm:64-170 zeroes 512 vertices per step). */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_clear_words(uint64_t* __restrict__ words, uint64_t first, uint64_t end,
                                                             gz_guard guard)
{
    if (gz_guard_off(guard)) return;
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
    hipLaunchKernelGGL(gz_k_clear_words, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, bitmap, first, end,
                       gz_guard{nullptr, 0});
    GZ_LAUNCH_CHECK();
}

extern "C" void perform_vertex_phase_cc(uint64_t vertex_start, uint64_t vertex_count, const uint64_t* reduce_buffer)
{
    (void)reduce_buffer;
    gz_clear_bitmap_range((uint64_t*)gz_graph.vertex_accumulators, vertex_start, vertex_count);
}

/* ---- the CC iteration loop with the engine chosen on the device -------------
 * (execution_cc.c:124-142 decides on the host after reading the vote.)  Each
 * iteration's launches read the loop's control words instead: the pull,
 * the frontier compaction and the compacted push each run only when the
 * engine word names them, a one-block control kernel sums the vote partials
 * of the engine that ran, logs (engine, vote) and picks the next engine by
 * the same rule, 0 when the vote is 0 or the iteration cap is reached, and
 * the clear of the frontier that becomes the next HasInfo* runs while the
 * loop goes on.  No launch depends on a host value that changes between
 * iterations, so a run of iterations can be captured into a HIP graph (in
 * pairs, the frontier bitmaps changing places as execution_cc.c:245-250
 * swaps them) and replayed until the engine word is 0. */

/* The pull over a persistent grid whose waves take chunks from the loop's
 * counter, one vote partial per wave. */
template <int VPL>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_cc_loop(const gz_vec_t* __restrict__ vecs, uint64_t nvec,
                                                                   uint64_t chunk_len, uint64_t nchunks,
                                                                   uint64_t* __restrict__ labels,
                                                                   uint64_t* __restrict__ next_frontier,
                                                                   const double* __restrict__ outdeg,
                                                                   uint32_t stat_mode, uint64_t* __restrict__ partials,
                                                                   uint64_t* __restrict__ counter, gz_guard guard)
{
    if (gz_guard_off(guard)) return;
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    gz_cc_pull_op op{labels, next_frontier, outdeg, stat_mode, 0};
    while (true) {
        uint64_t chunk = 0;
        if (gz_lane() == 0) chunk = atomicAdd((unsigned long long*)counter, 1ull);
        chunk = gz_readlane_u64(chunk, 0);
        if (chunk >= nchunks) break;
        const uint64_t b = chunk * chunk_len;
        const uint64_t e = (b + chunk_len < nvec) ? b + chunk_len : nvec;
        uint64_t first_d, carry_d, carry_m;
        uint32_t edges = 0;
        gz_walk_chunk<VPL, uint64_t>(vecs, b, e, op, first_d, carry_d, carry_m, edges);
        if (gz_lane() == 0) op.store(carry_d, carry_m);
    }
    const uint64_t total = gz_wave_sum_u64(op.vote);
    if (gz_lane() == 0) partials[gw] = total;
}

struct gz_loop_waves {
    uint64_t n[5]; /* vote partials each engine writes, by engine code */
};

/* One block: the vote of the engine that ran, the log entry, the next
 * engine; the pull counter and the overflow queue's counters are cleared for
 * the next iteration. */
__global__ __launch_bounds__(1024) void gz_k_cc_loop_control(uint64_t* __restrict__ ctrl,
                                                             const uint64_t* __restrict__ partials, gz_loop_waves waves,
                                                             uint64_t* __restrict__ queue)
{
    __shared__ unsigned long long lds[1024 / GZ_WAVE];
    const uint64_t engine = ctrl[GZ_LOOP_ENGINE];
    if (engine == 0 || engine > 4) return;
    const uint64_t n = waves.n[engine];
    unsigned long long x = 0;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) x += partials[i];
    x = gz_wave_sum_u64(x);
    if (gz_lane() == 0) lds[threadIdx.x / GZ_WAVE] = x;
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint64_t vote = 0;
    for (uint32_t w = 0; w < blockDim.x / GZ_WAVE; ++w) vote += lds[w];
    const uint64_t it = ctrl[GZ_LOOP_ITERATIONS];
    if (it < ctrl[GZ_LOOP_LOG_CAP]) {
        ctrl[GZ_LOOP_LOG + 2 * it] = engine == 1 ? 1 : 2;
        ctrl[GZ_LOOP_LOG + 2 * it + 1] = ctrl[GZ_LOOP_VOTE];
    }
    ctrl[GZ_LOOP_ITERATIONS] = it + 1;
    ctrl[engine == 1 ? GZ_LOOP_PULLS : GZ_LOOP_PUSHES] += 1;
    ctrl[GZ_LOOP_VOTE] = vote;
    ctrl[GZ_LOOP_PULL_NEXT] = 0;
    queue[0] = 0;
    queue[1] = 0;
    const uint64_t cap = ctrl[GZ_LOOP_MAX_ITERATIONS];
    uint64_t next = 0;
    if (vote != 0 && (cap == 0 || it + 1 < cap)) {
        const uint64_t mode = ctrl[GZ_LOOP_MODE];
        const int pull = mode == GZ_ENGINE_PULL ? 1 : mode == GZ_ENGINE_PUSH ? 0 : (vote > ctrl[GZ_LOOP_THRESHOLD]);
        next = gz_loop_engine(vote, pull, ctrl[GZ_LOOP_SPARSE_OK]);
    }
    ctrl[GZ_LOOP_ENGINE] = next;
}

static gz_loop_waves gz_cc_loop_wave_counts()
{
    gz_loop_waves w;
    uint64_t pull_waves, push_waves, wide, narrow;
    gz_cc_loop_waves(&pull_waves, &push_waves);
    gz_push_sparse_loop_waves(&wide, &narrow);
    w.n[0] = 0;
    w.n[1] = pull_waves;
    w.n[2] = push_waves;
    w.n[3] = wide;
    w.n[4] = narrow;
    return w;
}

extern "C" uint64_t gz_cc_loop_waves(uint64_t* pull_waves, uint64_t* push_waves)
{
    const uint64_t nchunks = gz_chunks_for(gz_graph.pull_vector_count);
    uint64_t pw = nchunks;
    if (pw > (uint64_t)GZ_PERSISTENT_BLOCKS * GZ_WAVES_PER_BLOCK) pw = (uint64_t)GZ_PERSISTENT_BLOCKS * GZ_WAVES_PER_BLOCK;
    pw = (pw + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK * GZ_WAVES_PER_BLOCK;
    if (pw == 0) pw = GZ_WAVES_PER_BLOCK;
    *pull_waves = pw;
    *push_waves = (uint64_t)GZ_PERSISTENT_BLOCKS * GZ_WAVES_PER_BLOCK;
    uint64_t wide, narrow;
    gz_push_sparse_loop_waves(&wide, &narrow);
    uint64_t m = pw > *push_waves ? pw : *push_waves;
    if (wide > m) m = wide;
    if (narrow > m) m = narrow;
    return m;
}

/* One iteration's launches, over the current HasInfo / HasInfo* roles. */
extern "C" void gz_cc_loop_enqueue(uint64_t* ctrl)
{
    const uint64_t V = gz_graph.num_vertices;
    const gz_loop_waves waves = gz_cc_loop_wave_counts();
    uint64_t* labels = (uint64_t*)gz_graph.vertex_props;
    uint64_t* next = (uint64_t*)gz_graph.vertex_accumulators;
    uint64_t* partials = (uint64_t*)gz_scratch.partials;
    const uint64_t* engine = ctrl + GZ_LOOP_ENGINE;
    const gz_guard pull_only{engine, 1ull << 1}, compacted{engine, 1ull << 2}, wide{engine, 1ull << 3},
        narrow{engine, 1ull << 4}, going_on{engine, 0x1eull};
    const uint64_t chunk_len = gz_chunk_len();
    const uint64_t nchunks = gz_chunks_for(gz_graph.pull_vector_count);
    const uint32_t vpl = gz_vectors_per_lane();
#define GZ_PULL_LOOP_ARGS                                                                                      \
    dim3((unsigned)(waves.n[1] / GZ_WAVES_PER_BLOCK)), dim3(GZ_BLOCK), 0, gz_stream, gz_graph.pull_vectors,     \
        gz_graph.pull_vector_count, chunk_len, nchunks, labels, next, gz_graph.vertex_outdegrees,               \
        gz_phase_config.stat_mode, partials, ctrl + GZ_LOOP_PULL_NEXT, pull_only
    if (vpl == 4)
        hipLaunchKernelGGL(gz_k_edge_pull_cc_loop<4>, GZ_PULL_LOOP_ARGS);
    else if (vpl == 2)
        hipLaunchKernelGGL(gz_k_edge_pull_cc_loop<2>, GZ_PULL_LOOP_ARGS);
    else
        hipLaunchKernelGGL(gz_k_edge_pull_cc_loop<1>, GZ_PULL_LOOP_ARGS);
#undef GZ_PULL_LOOP_ARGS
    GZ_LAUNCH_CHECK();
    const gz_cc_push_op op{labels, next, gz_graph.vertex_outdegrees, gz_phase_config.stat_mode};
    gz_launch_push_sparse_guarded(op, gz_graph.push_vectors, gz_graph.frontier_has_info, wide, narrow);
    gz_frontier_compact_guarded(gz_graph.frontier_has_info, gz_graph.push_vertex_offsets, V,
                                gz_graph.push_vector_count, compacted, GZ_PERSISTENT_BLOCKS);
    gz_slot_source src;
    src.offsets = gz_graph.push_vertex_offsets;
    src.list = gz_scratch.active_vertices;
    src.prefix = gz_scratch.active_prefix;
    src.totals = gz_scratch.active_totals;
    src.sweep_count = gz_graph.push_vector_count;
    hipLaunchKernelGGL(gz_k_edge_push_cc<true>, dim3(GZ_PERSISTENT_BLOCKS), dim3(GZ_BLOCK), 0, gz_stream,
                       gz_graph.push_vectors, src, (uint64_t)256, gz_graph.frontier_has_info, labels, next,
                       gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, partials, gz_scratch.stats, 0u, compacted);
    GZ_LAUNCH_CHECK();
    hipLaunchKernelGGL(gz_k_cc_loop_control, dim3(1), dim3(1024), 0, gz_stream, ctrl, partials, waves,
                       gz_scratch.push_queue);
    GZ_LAUNCH_CHECK();
    /* the frontier just pushed from becomes the next HasInfo*: cleared
     * while the loop goes on (perform_vertex_phase_cc after the swap) */
    const uint64_t words = (V + 63) >> 6;
    uint64_t blocks = (words + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gz_k_clear_words, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream,
                       gz_graph.frontier_has_info, (uint64_t)0, words, going_on);
    GZ_LAUNCH_CHECK();
}
