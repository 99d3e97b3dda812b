/*
 * gz_pr.hip -- PageRank phases on gfx950.
 *
 *   perform_edge_pull_phase_pr  phase_edge_pull_pr.asm:189-307 (step :45-171, merge entry :174-183)
 *   perform_edge_push_phase_pr  phase_edge_push_pr.asm:360-425 (traditional atomic variant)
 *   perform_vertex_phase_pr     phase_vertex_pr.asm:190-285   (init :27-72, step :75-134)
 *
 * Pull: a scheduling unit is one wavefront's chunk of `chunk_vectors`
 * consecutive vectors (the reference's unit of ~Nv/(32*threads) vectors,
 * scheduler_pull.inc:94-138).  The wavefront streams its chunk 64 vectors
 * per step, one vector (two 128-bit loads) per lane, gathers up to four FP64
 * source properties, and combines lanes that share a destination with a
 * segmented inclusive scan.  Destinations whose run ends inside the chunk are
 * stored directly (plain store, as the asm does when the destination
 * changes); the chunk's last run goes to its merge-buffer record, merged by
 * edge_pull_op_merge_with_merge_buffer.  The global PageRank sum S is the sum
 * of all gathered values; each wavefront writes its share to the partials.
 * No floating-point atomics, so results are reproducible run to run.
 */
#include "gz_pr_ops.h"
#include "gz_walk.h"

#include <type_traits>

/* Set by the edge phases: the scheduler-aware pull overwrites every
 * accumulator it owns, the traditional pull and the push accumulate with
 * atomics, so the following vertex phase must zero them
 * (phase_vertex_pr.asm:101-113). */
static uint32_t gz_pr_acc_needs_zero = 0;
/* Set by the source-sliced pull: the vertex phase sums the slices'
 * accumulators instead of reading the accumulator array. */
static uint32_t gz_pr_acc_in_slices = 0;
/* Set by the push: the vertex phase reads its fixed-point accumulators. */
static uint32_t gz_pr_acc_is_fixed = 0;

template <bool kSchedAware, int VPL>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_pr(const gz_vec_t* __restrict__ vecs, uint64_t nvec,
                                                              uint64_t chunk_len, uint64_t nchunks,
                                                              const double* __restrict__ props,
                                                              double* __restrict__ acc,
                                                              mergeaccum_t* __restrict__ merge,
                                                              double* __restrict__ partials,
                                                              uint64_t* __restrict__ stats, uint32_t collect)
{
    const int lane = gz_lane();
    const uint64_t chunk = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    if (chunk >= nchunks) return;
    const uint64_t b = chunk * chunk_len;
    const uint64_t e = (b + chunk_len < nvec) ? b + chunk_len : nvec;
    gz_pr_pull_op op{props, acc, kSchedAware, 0.0};
    uint64_t first_d, carry_d;
    double carry_v;
    uint32_t edges = 0;
    gz_walk_chunk<VPL, double>(vecs, b, e, op, first_d, carry_d, carry_v, edges);
    const double S = gz_wave_sum(op.sum);
    if (lane == 0) {
        if (kSchedAware) {
            mergeaccum_t r;
            r.initial_vertex_id = first_d;
            r.final_vertex_id = carry_d;
            r.final_partial_value = carry_v;
            r.__padding__ = 0;
            merge[chunk] = r;
        } else {
            unsafeAtomicAdd(&acc[carry_d], carry_v);
        }
        partials[chunk] = S;
    }
    if (collect) {
        const uint64_t tot_edges = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * chunk] = e - b;
            stats[2 * chunk + 1] = tot_edges;
        }
    }
}

template <bool kSchedAware>
static void gz_launch_pull_pr(uint64_t blocks, const gz_vec_t* edge_list, uint64_t count, uint64_t chunk_len,
                              uint64_t nchunks)
{
    const uint32_t vpl = gz_vectors_per_lane();
#define GZ_PULL_PR_ARGS                                                                                        \
    dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list, count, chunk_len, nchunks,                 \
        (const double*)gz_graph.vertex_props, (double*)gz_graph.vertex_accumulators, gz_merge_buffer,          \
        (double*)gz_scratch.partials, gz_scratch.stats, gz_phase_config.collect_stats
    if (vpl == 4)
        hipLaunchKernelGGL((gz_k_edge_pull_pr<kSchedAware, 4>), GZ_PULL_PR_ARGS);
    else if (vpl == 2)
        hipLaunchKernelGGL((gz_k_edge_pull_pr<kSchedAware, 2>), GZ_PULL_PR_ARGS);
    else
        hipLaunchKernelGGL((gz_k_edge_pull_pr<kSchedAware, 1>), GZ_PULL_PR_ARGS);
#undef GZ_PULL_PR_ARGS
}

extern "C" void perform_edge_pull_phase_pr(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    const uint64_t nchunks = gz_chunks_for(edge_list_count);
    gz_scratch_reserve(nchunks, gz_graph.num_vertices);
    gz_scratch.partials_is_double = 1;
    gz_scratch.partials_count = nchunks;
    gz_scratch.stats_count = gz_phase_config.collect_stats ? nchunks : 0;
    if (nchunks == 0) return;
    if (gz_phase_config.pull_variant == GZ_PULL_SOURCE_SLICED) {
        gz_pr_acc_needs_zero = 0;
        gz_pr_acc_in_slices = 1;
        gz_pr_acc_is_fixed = 0;
        gz_pull_pr_sliced(edge_list, edge_list_count);
        return;
    }
    gz_pr_acc_in_slices = 0;
    gz_pr_acc_is_fixed = 0;
    /* PageRank always merges a unit's boundary runs; the merge-buffer variant
     * is the scheduler-aware pull here. */
    const bool sa = gz_phase_config.pull_variant == GZ_PULL_SCHED_AWARE ||
                    gz_phase_config.pull_variant == GZ_PULL_MERGE_BUFFER;
    gz_pr_acc_needs_zero = sa ? 0u : 1u;
    if (sa && gz_merge_buffer_count < nchunks) {
        fprintf(stderr, "grazelle: merge buffer holds %llu units, the pull phase needs %llu\n",
                (unsigned long long)gz_merge_buffer_count, (unsigned long long)nchunks);
        exit(72);
    }
    const uint64_t chunk_len = gz_chunk_len();
    const uint64_t blocks = (nchunks + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    gz_ktimer_begin(GZ_K_PULL_PR);
    if (sa)
        gz_launch_pull_pr<true>(blocks, edge_list, edge_list_count, chunk_len, nchunks);
    else
        gz_launch_pull_pr<false>(blocks, edge_list, edge_list_count, chunk_len, nchunks);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PULL_PR);
}

/* The scheduler-aware pull over explicit arrays (the allocation-free
 * PageRank iteration): the same kernel, merge records and partials into the
 * caller's buffers, no statistics. */
void gz_launch_pr_pull_explicit(const gz_vec_t* vecs, uint64_t nvec, uint64_t chunk_len, uint64_t nchunks,
                                const double* props, double* acc, mergeaccum_t* merge, double* partials)
{
    if (nchunks == 0) return;
    const uint64_t blocks = (nchunks + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    const uint32_t vpl = gz_vectors_per_lane();
#define GZ_PULL_PR_EXPLICIT_ARGS                                                                               \
    dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, vecs, nvec, chunk_len, nchunks, props, acc, merge,     \
        partials, (uint64_t*)nullptr, 0u
    if (vpl == 4)
        hipLaunchKernelGGL((gz_k_edge_pull_pr<true, 4>), GZ_PULL_PR_EXPLICIT_ARGS);
    else if (vpl == 2)
        hipLaunchKernelGGL((gz_k_edge_pull_pr<true, 2>), GZ_PULL_PR_EXPLICIT_ARGS);
    else
        hipLaunchKernelGGL((gz_k_edge_pull_pr<true, 1>), GZ_PULL_PR_EXPLICIT_ARGS);
#undef GZ_PULL_PR_EXPLICIT_ARGS
    GZ_LAUNCH_CHECK();
}

/* ---- the traditional push with an order-independent sum -----------------
 * The reference's push (phase_edge_push_pr.asm:360-425) adds each edge's
 * message to its destination with a CAS loop, so the sum of a destination
 * with n in-edges is rounded in whatever order the threads reach it.  Here
 * each message is added as an unsigned 128-bit fixed-point integer with
 * GZ_PR_FIXED_BITS fraction bits, by two 64-bit integer atomics (the low word,
 * then the high word plus the carry the low add produced).  Integer addition
 * is associative, so every order gives the same bits: the exact sum of the
 * FP64 messages (rank / outdeg' >= 0.15 / V^2, at least 2^-67 for V <= 2^32,
 * so no message loses a bit; sums stay below 2^8, since the messages of all
 * edges add up to at most 1).  The vertex phase converts it to FP64 (two
 * conversions and one add: within two units in the last place of the exact
 * sum) and clears it. */
#define GZ_PR_FIXED_BITS 120

/* x >= 0 as a 128-bit fixed-point integer (lo, hi), truncating below 2^-120. */
__device__ __forceinline__ void gz_pr_to_fixed(double x, uint64_t& lo, uint64_t& hi)
{
    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    const int e = (int)((bits >> 52) & 0x7ff);
    lo = hi = 0;
    if (e == 0) return; /* zero (and denormals, far below 2^-120) */
    const uint64_t m = (bits & 0xfffffffffffffull) | (1ull << 52);
    const int sh = e - 1075 + GZ_PR_FIXED_BITS; /* value = m * 2^(e - 1075) */
    if (sh <= -64) return;
    if (sh < 0) {
        lo = m >> (-sh);
    } else if (sh < 64) {
        lo = m << sh;
        hi = sh ? (m >> (64 - sh)) : 0;
    } else if (sh < 128) {
        hi = m << (sh - 64);
    }
}

__device__ __forceinline__ double gz_pr_from_fixed(uint64_t lo, uint64_t hi)
{
    return ldexp((double)hi, 64 - GZ_PR_FIXED_BITS) + ldexp((double)lo, -GZ_PR_FIXED_BITS);
}

__device__ __forceinline__ void gz_pr_fixed_add(uint64_t* __restrict__ a, uint64_t lo, uint64_t hi)
{
    const unsigned long long old = atomicAdd((unsigned long long*)a, (unsigned long long)lo);
    const uint64_t carry = ((uint64_t)old + lo < (uint64_t)old) ? 1u : 0u;
    if (hi + carry) atomicAdd((unsigned long long*)(a + 1), (unsigned long long)(hi + carry));
}

/* Accumulators of the push: two words per vertex. */
static uint64_t* gz_pr_fixed_acc = nullptr;
static uint64_t gz_pr_fixed_vertices = 0;

extern "C" void gz_pr_release(void)
{
    gz_device_free(gz_pr_fixed_acc);
    gz_pr_fixed_acc = nullptr;
    gz_pr_fixed_vertices = 0;
    gz_pr_acc_is_fixed = 0;
}

/* A zeroed fixed-point accumulator pair per vertex of the loaded graph. */
static uint64_t* gz_pr_fixed_reserve(uint64_t num_vertices)
{
    if (gz_pr_fixed_acc == nullptr || gz_pr_fixed_vertices != num_vertices) {
        gz_device_free(gz_pr_fixed_acc);
        gz_pr_fixed_acc = (uint64_t*)gz_device_malloc((num_vertices + 8) * 2 * sizeof(uint64_t));
        gz_pr_fixed_vertices = num_vertices;
    }
    return gz_pr_fixed_acc;
}

extern "C" void gz_pr_clear_fixed(void)
{
    if (gz_pr_fixed_acc != nullptr) gz_device_zero(gz_pr_fixed_acc, (gz_pr_fixed_vertices + 8) * 2 * sizeof(uint64_t));
}

/* Traditional push: every valid lane adds the source's outbound message to
 * its destination's accumulator with integer atomics (above), and S gains
 * the message once per valid lane. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_push_pr(const gz_vec_t* __restrict__ vecs, uint64_t nvec,
                                                              uint64_t chunk_len, uint64_t nchunks,
                                                              const double* __restrict__ props,
                                                              uint64_t* __restrict__ acc, double* __restrict__ partials,
                                                              uint64_t* __restrict__ stats, uint32_t collect)
{
    const int lane = gz_lane();
    const uint64_t chunk = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    if (chunk >= nchunks) return;
    const uint64_t b = chunk * chunk_len;
    const uint64_t e = (b + chunk_len < nvec) ? b + chunk_len : nvec;
    double s_lane = 0.0;
    uint32_t edges = 0;
    for (uint64_t i = b + (uint64_t)lane; i < e; i += GZ_WAVE) {
        const gz_lanes v = gz_load_vector(vecs, i);
        const double msg = props[gz_shared_vertex(v)];
        uint64_t mlo, mhi;
        gz_pr_to_fixed(msg, mlo, mhi);
        const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (gz_lane_valid(l[k])) {
                gz_pr_fixed_add(acc + 2 * gz_lane_id(l[k]), mlo, mhi);
                s_lane += msg;
                edges += 1;
            }
        }
    }
    const double S = gz_wave_sum(s_lane);
    if (lane == 0) partials[chunk] = S;
    if (collect) {
        const uint64_t tot_edges = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * chunk] = e - b;
            stats[2 * chunk + 1] = tot_edges;
        }
    }
}

extern "C" void perform_edge_push_phase_pr(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    gz_pr_acc_needs_zero = 1;
    gz_pr_acc_in_slices = 0;
    gz_pr_acc_is_fixed = 1;
    uint64_t* fixed = gz_pr_fixed_reserve(gz_graph.num_vertices);
    const uint64_t nchunks = gz_chunks_for(edge_list_count);
    gz_scratch_reserve(nchunks, gz_graph.num_vertices);
    gz_scratch.partials_is_double = 1;
    gz_scratch.partials_count = nchunks;
    gz_scratch.stats_count = gz_phase_config.collect_stats ? nchunks : 0;
    if (nchunks == 0) return;
    const uint64_t chunk_len = gz_chunk_len();
    const uint64_t blocks = (nchunks + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    gz_ktimer_begin(GZ_K_PUSH_PR);
    hipLaunchKernelGGL(gz_k_edge_push_pr, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, edge_list,
                       edge_list_count, chunk_len, nchunks, (const double*)gz_graph.vertex_props, fixed,
                       (double*)gz_scratch.partials, gz_scratch.stats, gz_phase_config.collect_stats);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_PUSH_PR);
}

/* props[v] = ((acc[v] + (1-S)/V) * d + (1-d)/V) / outdeg'(v), outdeg' = V
 * for sinks, as separate IEEE mul/add/div (this file is built with
 * -ffp-contract=off, matching the asm's vaddpd/vmulpd/vaddpd/vdivpd). */
__device__ __forceinline__ double gz_pr_vertex_value(double a, double od, double c, double k, double damping,
                                                     double num_vertices)
{
    if (od == 0.0) od = num_vertices;
    double x = a + c;
    x = x * damping;
    x = x + k;
    return x / od;
}

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_vertex_pr(double* __restrict__ props, double* __restrict__ acc,
                                                           const double* __restrict__ outdeg, uint64_t start,
                                                           uint64_t end, double num_vertices,
                                                           const uint64_t* __restrict__ reduce_buffer,
                                                           double damping, uint32_t zero_acc)
{
    const uint64_t sbits = reduce_buffer[0];
    const double S = __longlong_as_double((long long)sbits);
    const double c = (1.0 - S) / num_vertices;
    const double k = (1.0 - damping) / num_vertices;
    for (uint64_t v = start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < end;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const double a = acc[v];
        if (zero_acc) acc[v] = 0.0;
        props[v] = gz_pr_vertex_value(a, outdeg[v], c, k, damping, num_vertices);
    }
}

/* The vertex phase after a source-sliced pull: acc[v] is the sum of the
 * slices holding v (mask bit x), added in slice order, 0 without in-edges.
 * Slice x keeps its sums densely by local rank; one wavefront takes 64
 * consecutive vertices, and v's accumulator in slice x is the block's first
 * one there (rank_base[block][x], an index into the concatenated runs) plus
 * the lanes below v that are also in slice x, so each slice's sums are read
 * as one contiguous run.  Every slice is added, 0 where v is not in it: x +
 * 0.0 is x, so the sum is the in-order sum of v's own slices.  A level of
 * eight slices that no vertex of the block is in costs one ballot.  Masks are
 * 32-bit words for up to 32 slices, else 64-bit. */
template <int NS>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_vertex_pr_sliced(
    double* __restrict__ props, const double* __restrict__ outdeg, uint64_t start, uint64_t end,
    uint64_t num_vertices, const uint64_t* __restrict__ reduce_buffer, double damping,
    const void* __restrict__ mask_words, const double* __restrict__ parts, const uint32_t* __restrict__ rank_base)
{
    typedef typename std::conditional<(NS <= 32), uint32_t, uint64_t>::type mask_t;
    const mask_t* __restrict__ mask = (const mask_t*)mask_words;
    const double nv = (double)num_vertices;
    const double S = __longlong_as_double((long long)reduce_buffer[0]);
    const double c = (1.0 - S) / nv;
    const double k = (1.0 - damping) / nv;
    const int lane = gz_lane();
    const uint64_t last = (end + GZ_WAVE - 1) / GZ_WAVE;
    const uint64_t waves = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    const uint32_t wib = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / GZ_WAVE));
    uint64_t blk = start / GZ_WAVE + (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + wib;
    /* the next block's mask and accumulator bases (lane x < NS holds slice
     * x's) are loaded one iteration ahead, so that the slice sums (whose
     * addresses depend on them) are the only wait of an iteration */
    mask_t m_next = 0;
    uint32_t rb_next = 0;
    if (blk < last) {
        const uint64_t v0 = blk * GZ_WAVE + (uint64_t)lane;
        m_next = (v0 < num_vertices) ? mask[v0] : (mask_t)0;
        rb_next = (lane < NS) ? rank_base[blk * (uint64_t)NS + lane] : 0u;
    }
    for (; blk < last; blk += waves) {
        const uint64_t v = blk * GZ_WAVE + (uint64_t)lane;
        const mask_t m = m_next;
        const uint32_t rb = rb_next;
        const uint64_t nb = blk + waves;
        if (nb < last) {
            const uint64_t vn = nb * GZ_WAVE + (uint64_t)lane;
            m_next = (vn < num_vertices) ? mask[vn] : (mask_t)0;
            rb_next = (lane < NS) ? rank_base[nb * (uint64_t)NS + lane] : 0u;
        }
        const bool mine = v >= start && v < end;
        const double od = mine ? outdeg[v] : 1.0;
        double a = 0.0;
#pragma unroll
        for (int g = 0; g < NS; g += GZ_XCDS) {
            if (__ballot(((m >> g) & (mask_t)0xff) != 0) == 0) continue; /* wave-uniform */
            double p[GZ_XCDS];
#pragma unroll
            for (int j = 0; j < GZ_XCDS; ++j) {
                p[j] = 0.0;
                if (g + j < NS) {
                    const bool has = (m >> (g + j)) & 1u;
                    const uint64_t bal = __ballot(has);
                    if (bal != 0) { /* wave-uniform: no block vertex in slice g + j, no load */
                        const double* run = parts + (uint32_t)__builtin_amdgcn_readlane((int)rb, g + j);
                        if (has) p[j] = run[__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u))];
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < GZ_XCDS; ++j)
                if (g + j < NS) a = a + p[j];
        }
        if (mine) props[v] = gz_pr_vertex_value(a, od, c, k, damping, nv);
    }
}

/* The vertex phase after the push: the fixed-point sums rounded once to
 * FP64, then cleared for the next push. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_vertex_pr_fixed(double* __restrict__ props,
                                                                 uint64_t* __restrict__ fixed,
                                                                 const double* __restrict__ outdeg, uint64_t start,
                                                                 uint64_t end, double num_vertices,
                                                                 const uint64_t* __restrict__ reduce_buffer,
                                                                 double damping)
{
    const double S = __longlong_as_double((long long)reduce_buffer[0]);
    const double c = (1.0 - S) / num_vertices;
    const double k = (1.0 - damping) / num_vertices;
    for (uint64_t v = start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < end;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const gz_u64x2 w = *reinterpret_cast<const gz_u64x2*>(fixed + 2 * v);
        *reinterpret_cast<gz_u64x2*>(fixed + 2 * v) = gz_u64x2{0, 0};
        props[v] = gz_pr_vertex_value(gz_pr_from_fixed(w.x, w.y), outdeg[v], c, k, damping, num_vertices);
    }
}

/* The vertex phase over explicit arrays (the allocation-free PageRank
 * iteration), S read from sum_word on the device. */
void gz_launch_pr_vertex_explicit(double* props, double* acc, const double* outdeg, uint64_t num_vertices,
                                  const uint64_t* sum_word, double damping)
{
    if (num_vertices == 0) return;
    uint64_t blocks = (num_vertices + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gz_k_vertex_pr, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, props, acc, outdeg,
                       (uint64_t)0, num_vertices, (double)num_vertices, sum_word, damping, 0u);
    GZ_LAUNCH_CHECK();
}

extern "C" void perform_vertex_phase_pr(uint64_t vertex_start, uint64_t vertex_count, const uint64_t* reduce_buffer)
{
    if (vertex_count == 0) return;
    const uint64_t end = vertex_start + vertex_count;
    uint64_t blocks = (vertex_count + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks > 4096) blocks = 4096;
    gz_ktimer_begin(GZ_K_VERTEX_PR);
    if (gz_pr_acc_in_slices && gz_slices.mask != nullptr) {
        /* wavefronts stride over the 64-vertex blocks */
        const uint64_t vblocks = (end + GZ_WAVE - 1) / GZ_WAVE - vertex_start / GZ_WAVE;
        uint64_t sblocks = (vblocks + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
        /* many more wavefronts than fit at once (one to three blocks each):
         * a wavefront that finishes early is replaced at once, so the tail
         * is short.  On the 41.7 M-vertex graph with 57 slices, 4096
         * workgroups took 1.03-1.06 ms and 65536 took 0.93-0.95 ms. */
        if (sblocks > 65536) sblocks = 65536;
#define GZ_VERTEX_SLICED_ARGS                                                                                  \
    dim3((unsigned)sblocks), dim3(GZ_BLOCK), 0, gz_stream, (double*)gz_graph.vertex_props,                        \
        gz_graph.vertex_outdegrees, vertex_start, end, gz_graph.num_vertices, reduce_buffer,                     \
        gz_phase_config.damping, (const void*)gz_slices.mask, (const double*)gz_slices.parts,                    \
        (const uint32_t*)gz_slices.rank_base
#define GZ_VERTEX_SLICED_CASE(L)                                                                                \
    case GZ_XCDS * (L) + 1:                                                                                      \
        hipLaunchKernelGGL(gz_k_vertex_pr_sliced<GZ_XCDS * (L) + 1>, GZ_VERTEX_SLICED_ARGS);                       \
        break;
        switch (gz_slices.nslices) {
            GZ_VERTEX_SLICED_CASE(1)
            GZ_VERTEX_SLICED_CASE(2)
            GZ_VERTEX_SLICED_CASE(3)
            GZ_VERTEX_SLICED_CASE(4)
            GZ_VERTEX_SLICED_CASE(5)
            GZ_VERTEX_SLICED_CASE(6)
            GZ_VERTEX_SLICED_CASE(7)
        default:
            fprintf(stderr, "grazelle: no vertex phase for %u slices\n", gz_slices.nslices);
            exit(72);
        }
#undef GZ_VERTEX_SLICED_CASE
#undef GZ_VERTEX_SLICED_ARGS
    } else if (gz_pr_acc_is_fixed && gz_pr_fixed_acc != nullptr && gz_pr_fixed_vertices == gz_graph.num_vertices) {
        hipLaunchKernelGGL(gz_k_vertex_pr_fixed, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream,
                           (double*)gz_graph.vertex_props, gz_pr_fixed_acc, gz_graph.vertex_outdegrees, vertex_start,
                           end, (double)gz_graph.num_vertices, reduce_buffer, gz_phase_config.damping);
    } else {
        hipLaunchKernelGGL(gz_k_vertex_pr, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream,
                           (double*)gz_graph.vertex_props, (double*)gz_graph.vertex_accumulators,
                           gz_graph.vertex_outdegrees, vertex_start, end, (double)gz_graph.num_vertices,
                           reduce_buffer, gz_phase_config.damping, gz_pr_acc_needs_zero);
    }
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_VERTEX_PR);
}
