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
/* The one-launch PageRank's per-vertex run heads (gz_k_pr_persistent). */
static uint32_t* gz_pr_run_head = nullptr;
static uint64_t gz_pr_run_head_cap = 0;

/* One wavefront's pull over chunk `chunk`: complete runs stored, the last
 * run to the chunk's merge record (or added atomically), the chunk's share
 * of S to partials[chunk].  Returns the valid lanes the lane examined. */
template <bool kSchedAware, int VPL>
__device__ __forceinline__ uint32_t gz_pr_pull_chunk(const gz_vec_t* __restrict__ vecs, uint64_t nvec,
                                                     uint64_t chunk_len, uint64_t chunk,
                                                     const double* __restrict__ props, double* __restrict__ acc,
                                                     mergeaccum_t* __restrict__ merge, double* __restrict__ partials)
{
    const int lane = gz_lane();
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
    return edges;
}

template <bool kSchedAware, int VPL>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_pr(const gz_vec_t* __restrict__ vecs, uint64_t nvec,
                                                              uint64_t chunk_len, uint64_t nchunks,
                                                              const double* __restrict__ props,
                                                              double* __restrict__ acc,
                                                              mergeaccum_t* __restrict__ merge,
                                                              double* __restrict__ partials,
                                                              uint64_t* __restrict__ stats, uint32_t collect)
{
    const uint64_t chunk = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    if (chunk >= nchunks) return;
    const uint32_t edges =
        gz_pr_pull_chunk<kSchedAware, VPL>(vecs, nvec, chunk_len, chunk, props, acc, merge, partials);
    if (collect) {
        const uint64_t b = chunk * chunk_len;
        const uint64_t e = (b + chunk_len < nvec) ? b + chunk_len : nvec;
        const uint64_t tot_edges = gz_wave_sum_u64(edges);
        if (gz_lane() == 0) {
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
    gz_device_free(gz_pr_run_head);
    gz_pr_run_head = nullptr;
    gz_pr_run_head_cap = 0;
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

/* ---- every iteration of a small graph in one launch -------------------------
 * On a graph whose iteration takes a few tens of microseconds (R-MAT s16:
 * about 21 us over four launches), the kernels' own start and drain cost
 * more than their work.  gz_k_pr_persistent runs all iterations in one
 * cooperative launch (every workgroup resident), with two grid barriers per
 * iteration:
 *   pull        the chunks of gz_k_edge_pull_pr<true>, strided over the grid
 *   S, vertex   every workgroup sums the chunks' partials itself in the two
 *               stages' order of gz_k_reduce (up to 8192 chunks; beyond, the
 *               first stage's slices are spread over the workgroups and a
 *               third barrier precedes the second stage), then the vertices
 *               of gz_k_vertex_pr, where a vertex that ends a run of merge
 *               records first adds them as gz_merge_sum_record does
 * The same arithmetic in the same order on the same data, so its results are
 * bit-identical to the launches it replaces.
 *
 * The barrier: every wavefront waits until its stores have reached the L2;
 * the workgroup's thread 0 then releases at agent scope (the XCDs' L2 caches
 * are not coherent with each other, so the L2 is written back), arrives on
 * one of eight group counters (the last of a group on the global one), polls
 * the generation word with relaxed coherent loads until it moves, and
 * acquires at agent scope (L1 and L2 invalidated) before the workgroup goes
 * on.  A wait gives up after `spin_limit`
 * polls: it raises the abort word, which every waiting workgroup sees, so
 * every workgroup leaves the kernel; the host reports it and the driver
 * restarts the run with the per-phase launches. */
#define GZ_BAR_GROUPS 8
struct gz_grid_bar_t {
    uint32_t group[GZ_BAR_GROUPS][16]; /* arrivals per group, one 64-byte line each */
    uint32_t count;                    /* groups arrived */
    uint32_t gen;
    uint32_t abort;
    uint32_t pad[13];
};

/* Workgroup w arrives on group w % 8; the last of a group arrives on the
 * global count, whose last arrival moves the generation word. */
__device__ __forceinline__ bool gz_grid_sync(gz_grid_bar_t* bar, uint32_t spin_limit)
{
    __shared__ uint32_t s_ok;
    /* every wavefront's stores have reached the XCD's L2 */
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        const uint32_t grp = blockIdx.x % GZ_BAR_GROUPS;
        const uint32_t members = gridDim.x / GZ_BAR_GROUPS + (grp < gridDim.x % GZ_BAR_GROUPS ? 1u : 0u);
        const uint32_t groups = gridDim.x < GZ_BAR_GROUPS ? gridDim.x : GZ_BAR_GROUPS;
        const uint32_t g = __hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        /* a release arrival: the XCD's L2 is written back first, so the other
         * XCDs read what this workgroup wrote */
        const uint32_t a =
            __hip_atomic_fetch_add(&bar->group[grp][0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        bool last = false;
        if (a == members) {
            __hip_atomic_store(&bar->group[grp][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t c =
                __hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
            if (c == groups) {
                last = true;
                __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&bar->gen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (!last) {
            /* relaxed polls read the coherent word without touching the
             * caches; the acquire below follows the last one */
            uint32_t polls = 0;
            while (__hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                    ++polls > spin_limit) {
                    __hip_atomic_store(&bar->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) ok = 0;
        /* acquire: the CU's L1 and the XCD's L2 drop what other XCDs may
         * have rewritten; the wait orders it before the workgroup's loads */
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

/* S from the chunks' partials in the order of gz_k_reduce: slice b of the
 * first stage is ceil(n / blocks) consecutive partials summed by 256 threads
 * with stride 256 and gz_block_sum, the second stage the same over the slice
 * sums.  Every thread of the workgroup calls it; all get the result. */
__device__ __forceinline__ double gz_pr_sum_partials(const double* partials, uint64_t n, uint32_t blocks,
                                                     double* lds, double* slot_lds, double* total_lds)
{
    const uint64_t per = (n + blocks - 1) / blocks;
    for (uint32_t b = 0; b < blocks; ++b) {
        const uint64_t lo = (uint64_t)b * per;
        const uint64_t hi = (lo + per < n) ? lo + per : n;
        double x = 0.0;
        for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) x += partials[i];
        const double r = gz_block_sum<double>(x, lds);
        if (threadIdx.x == 0) slot_lds[b] = r;
    }
    __syncthreads();
    double y = 0.0;
    for (uint32_t i = threadIdx.x; i < blocks; i += blockDim.x) y += slot_lds[i];
    const double total = gz_block_sum<double>(y, lds);
    if (threadIdx.x == 0) *total_lds = total;
    __syncthreads();
    return *total_lds;
}

/* The chunks whose partials every workgroup sums itself (S in LDS, no grid
 * barrier for the first stage's slice sums). */
#define GZ_PR_ONE_LAUNCH_LOCAL_SUM_CHUNKS 8192u

/* acc[v] after the merge of the pull's records (gz_merge_sum_record), for a
 * vertex phase that merges as it goes: a vertex that ends a run of records
 * (run_head[v] = the run's first record + 1) adds the run's partials in
 * record order, then the accumulator the next unit stored directly when it
 * starts on v, and keeps the result in acc[v]; any other vertex reads
 * acc[v].  The same additions in the same order as the merge kernel. */
__device__ __forceinline__ double gz_pr_merged_acc(uint64_t v, const uint32_t* __restrict__ run_head,
                                                   const mergeaccum_t* __restrict__ merge, uint64_t count,
                                                   double* __restrict__ acc)
{
    const uint32_t h = run_head[v];
    if (h == 0u) return acc[v];
    uint64_t j = h - 1u;
    double a = merge[j].final_partial_value;
    ++j;
    while (j < count && merge[j].final_vertex_id == v) {
        a = a + merge[j].final_partial_value;
        ++j;
    }
    if (j < count && merge[j].initial_vertex_id == v) a = a + acc[v];
    acc[v] = a;
    return a;
}

/* run_head[final vertex of a run of records] = first record of the run + 1
 * (the array is cleared first).  The chunking and the graph fix the records'
 * vertex ids, so one pull's records serve every later iteration. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_pr_run_heads(const mergeaccum_t* __restrict__ merge, uint64_t count,
                                                              uint32_t* __restrict__ run_head)
{
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= count) return;
    const uint64_t f = merge[u].final_vertex_id;
    if (u == 0 || merge[u - 1].final_vertex_id != f) run_head[f] = (uint32_t)(u + 1);
}

/* The merge, the sum of S and the vertex phase in one launch, after a
 * scheduler-aware pull of at most GZ_PR_FUSED_TAIL_CHUNKS chunks: every
 * workgroup sums the chunks' partials itself in gz_k_reduce's order, then
 * its vertices merge as they go (gz_pr_merged_acc).  Bit-identical to
 * gz_k_merge_sum + gz_k_reduce + gz_k_vertex_pr. */
#define GZ_PR_FUSED_TAIL_CHUNKS 4096u
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_vertex_pr_merged(double* __restrict__ props, double* __restrict__ acc,
                                                                  const double* __restrict__ outdeg,
                                                                  uint64_t num_vertices,
                                                                  const mergeaccum_t* __restrict__ merge,
                                                                  uint64_t nchunks, const double* __restrict__ partials,
                                                                  uint32_t reduce_blocks,
                                                                  const uint32_t* __restrict__ run_head,
                                                                  uint64_t* __restrict__ reduce_buffer, double damping)
{
    __shared__ double lds[GZ_WAVES_PER_BLOCK];
    __shared__ double slot_lds[2];
    __shared__ double s_total;
    const double S = gz_pr_sum_partials(partials, nchunks, reduce_blocks, lds, slot_lds, &s_total);
    if (blockIdx.x == 0 && threadIdx.x == 0) reduce_buffer[0] = (uint64_t)__double_as_longlong(S);
    const double nv = (double)num_vertices;
    const double c = (1.0 - S) / nv;
    const double k = (1.0 - damping) / nv;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < num_vertices;
         v += (uint64_t)gridDim.x * blockDim.x)
        props[v] = gz_pr_vertex_value(gz_pr_merged_acc(v, run_head, merge, nchunks, acc), outdeg[v], c, k, damping,
                                      nv);
}

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_pr_clear_u32(uint32_t* __restrict__ p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

static int gz_pr_fused_tail_on = 1;

extern "C" void gz_pr_fused_tail(int on) { gz_pr_fused_tail_on = on ? 1 : 0; }

static uint32_t* gz_pr_run_heads_reserve(uint64_t V)
{
    if (gz_pr_run_head_cap < V) {
        gz_device_free(gz_pr_run_head);
        gz_pr_run_head = (uint32_t*)gz_device_malloc(V * sizeof(uint32_t));
        gz_pr_run_head_cap = V;
    }
    return gz_pr_run_head;
}

extern "C" int gz_pr_fused_tail_applies(void)
{
    const uint64_t nvec = gz_graph.pull_vector_count;
    if (!gz_pr_fused_tail_on || gz_graph.pull_vectors == nullptr || nvec == 0 || gz_graph.num_vertices == 0 ||
        gz_graph.num_vertices >= 0xffffffffull)
        return 0;
    const uint64_t nchunks = gz_chunks_for(nvec);
    return nchunks <= GZ_PR_FUSED_TAIL_CHUNKS && gz_merge_buffer_count >= nchunks;
}

extern "C" void gz_pr_run_heads_build(void)
{
    const uint64_t V = gz_graph.num_vertices, nchunks = gz_chunks_for(gz_graph.pull_vector_count);
    uint32_t* rh = gz_pr_run_heads_reserve(V);
    uint64_t cb = (V + GZ_BLOCK - 1) / GZ_BLOCK;
    if (cb > 4096) cb = 4096;
    hipLaunchKernelGGL(gz_k_pr_clear_u32, dim3((unsigned)cb), dim3(GZ_BLOCK), 0, gz_stream, rh, V);
    GZ_LAUNCH_CHECK();
    hipLaunchKernelGGL(gz_k_pr_run_heads, dim3((unsigned)((nchunks + GZ_BLOCK - 1) / GZ_BLOCK)), dim3(GZ_BLOCK), 0,
                       gz_stream, gz_merge_buffer, nchunks, rh);
    GZ_LAUNCH_CHECK();
}

extern "C" void gz_pr_vertex_merged(uint64_t* reduce_buffer)
{
    gz_reduce_mirror_stale();
    const uint64_t V = gz_graph.num_vertices, nchunks = gz_chunks_for(gz_graph.pull_vector_count);
    uint64_t blocks = (V + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks > 1024) blocks = 1024;
    gz_ktimer_begin(GZ_K_VERTEX_PR);
    hipLaunchKernelGGL(gz_k_vertex_pr_merged, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream,
                       (double*)gz_graph.vertex_props, (double*)gz_graph.vertex_accumulators,
                       gz_graph.vertex_outdegrees, V, gz_merge_buffer, nchunks, (const double*)gz_scratch.partials,
                       (uint32_t)gz_reduce_blocks(nchunks), gz_pr_run_head, reduce_buffer, gz_phase_config.damping);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_VERTEX_PR);
}

template <int VPL>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_pr_persistent(
    const gz_vec_t* __restrict__ vecs, uint64_t nvec, uint64_t chunk_len, uint64_t nchunks, double* props,
    double* acc, const double* __restrict__ outdeg, uint64_t num_vertices, mergeaccum_t* merge, double* partials,
    double* slots, uint32_t reduce_blocks, uint32_t* run_head, uint64_t* __restrict__ reduce_buffer,
    double damping, uint64_t iterations, gz_grid_bar_t* bar, uint32_t spin_limit)
{
    __shared__ double lds[GZ_WAVES_PER_BLOCK];
    __shared__ double slot_lds[2];
    __shared__ double s_total;
    const uint64_t waves = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t threads = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double nv = (double)num_vertices;
    const double k = (1.0 - damping) / nv;
    const bool local_sum = nchunks <= GZ_PR_ONE_LAUNCH_LOCAL_SUM_CHUNKS;
    const uint64_t per = (nchunks + reduce_blocks - 1) / reduce_blocks;
    for (uint64_t it = 0; it < iterations; ++it) {
        for (uint64_t c = gw; c < nchunks; c += waves)
            gz_pr_pull_chunk<true, VPL>(vecs, nvec, chunk_len, c, props, acc, merge, partials);
        if (!gz_grid_sync(bar, spin_limit)) return;
        if (it == 0) {
            /* run_head[final vertex of a run of records] = first record of
             * the run + 1 (zero elsewhere, cleared before the launch); the
             * chunking and the graph fix the records' ids for every
             * iteration */
            for (uint64_t u = gt; u < nchunks; u += threads) {
                const uint64_t f = merge[u].final_vertex_id;
                if (u == 0 || merge[u - 1].final_vertex_id != f) run_head[f] = (uint32_t)(u + 1);
            }
            if (!gz_grid_sync(bar, spin_limit)) return;
        }
        double S;
        if (local_sum) {
            /* at most two slices: every workgroup sums all partials itself */
            S = gz_pr_sum_partials(partials, nchunks, reduce_blocks, lds, slot_lds, &s_total);
        } else {
            for (uint32_t b = blockIdx.x; b < reduce_blocks; b += gridDim.x) {
                const uint64_t lo = (uint64_t)b * per;
                const uint64_t hi = (lo + per < nchunks) ? lo + per : nchunks;
                double x = 0.0;
                for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) x += partials[i];
                const double r = gz_block_sum<double>(x, lds);
                if (threadIdx.x == 0) slots[b] = r;
            }
            if (!gz_grid_sync(bar, spin_limit)) return;
            double y = 0.0;
            for (uint64_t i = threadIdx.x; i < reduce_blocks; i += blockDim.x) y += slots[i];
            const double total = gz_block_sum<double>(y, lds);
            if (threadIdx.x == 0) s_total = total;
            __syncthreads();
            S = s_total;
        }
        if (gt == 0) reduce_buffer[0] = (uint64_t)__double_as_longlong(S);
        const double c = (1.0 - S) / nv;
        /* the vertex phase with the merge folded in: a vertex that ends a
         * chunk's last run (run_head[v] = first record of its run + 1) adds
         * the run's records as gz_merge_sum_record does */
        for (uint64_t v = gt; v < num_vertices; v += threads)
            props[v] = gz_pr_vertex_value(gz_pr_merged_acc(v, run_head, merge, nchunks, acc), outdeg[v], c, k,
                                          damping, nv);
        if (!gz_grid_sync(bar, spin_limit)) return;
    }
}

/* 0: off (the default: see gz_pr_persistent); 1: graphs whose per-iteration
 * bytes fit 128 MiB; 2: any graph. */
static int gz_pr_persistent_mode = 0;
static uint32_t gz_pr_persistent_per_cu = 2;
static uint32_t gz_pr_persistent_spin = 1u << 22;
static uint64_t gz_pr_persistent_count = 0;
static uint64_t gz_pr_persistent_aborts = 0;
static int gz_pr_persistent_unavailable = 0;
static gz_grid_bar_t* gz_pr_bar = nullptr;
static uint32_t* gz_pr_bar_host = nullptr;

extern "C" int gz_pr_persistent(int mode, uint32_t workgroups_per_cu)
{
    if (mode < 0 || mode > 2 || workgroups_per_cu < 1 || workgroups_per_cu > 8) return -1;
    gz_pr_persistent_mode = mode;
    gz_pr_persistent_per_cu = workgroups_per_cu;
    return 0;
}

extern "C" uint64_t gz_pr_persistent_runs(void) { return gz_pr_persistent_count; }

extern "C" uint64_t gz_pr_persistent_gave_up(void) { return gz_pr_persistent_aborts; }

extern "C" void gz_pr_persistent_spin_limit(uint32_t polls) { gz_pr_persistent_spin = polls; }

extern "C" int gz_pr_run_persistent(uint64_t iterations, uint64_t* reduce_buffer)
{
    gz_reduce_mirror_stale();
    const uint64_t nvec = gz_graph.pull_vector_count, V = gz_graph.num_vertices;
    if (gz_pr_persistent_mode == 0 || gz_pr_persistent_unavailable || iterations == 0 ||
        gz_graph.pull_vectors == nullptr || nvec == 0 || V == 0)
        return 0;
    if (gz_pr_persistent_mode == 1 && 32 * nvec + 24 * V > (128ull << 20)) return 0;
    int dev = 0, coop = 0, cus = 0;
    GZ_CHECK(hipGetDevice(&dev));
    GZ_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    GZ_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (!coop || cus <= 0) return 0;
    const uint32_t vpl = gz_vectors_per_lane();
    const void* fn = vpl == 4 ? (const void*)gz_k_pr_persistent<4>
                   : vpl == 2 ? (const void*)gz_k_pr_persistent<2>
                              : (const void*)gz_k_pr_persistent<1>;
    int resident = 0;
    GZ_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, fn, GZ_BLOCK, 0));
    if (resident <= 0) return 0;
    const uint32_t per_cu = (uint32_t)resident < gz_pr_persistent_per_cu ? (uint32_t)resident : gz_pr_persistent_per_cu;
    const uint32_t grid = per_cu * (uint32_t)cus;

    const uint64_t nchunks = gz_chunks_for(nvec);
    if (nchunks >= 0xffffffffull) return 0; /* run heads are 32-bit */
    if (gz_merge_buffer_count < nchunks) {
        fprintf(stderr, "grazelle: merge buffer holds %llu units, the pull phase needs %llu\n",
                (unsigned long long)gz_merge_buffer_count, (unsigned long long)nchunks);
        exit(72);
    }
    gz_scratch_reserve(nchunks, V);
    gz_scratch.partials_is_double = 1;
    gz_scratch.partials_count = nchunks;
    gz_scratch.stats_count = 0;
    gz_pr_acc_needs_zero = 0;
    gz_pr_acc_in_slices = 0;
    gz_pr_acc_is_fixed = 0;
    if (gz_pr_bar == nullptr) {
        gz_pr_bar = (gz_grid_bar_t*)gz_device_malloc(sizeof(gz_grid_bar_t));
        GZ_CHECK(hipHostMalloc((void**)&gz_pr_bar_host, sizeof(uint32_t)));
    }
    GZ_CHECK(hipMemsetAsync(gz_pr_bar, 0, sizeof(gz_grid_bar_t), gz_stream));
    GZ_CHECK(hipMemsetAsync(gz_pr_run_heads_reserve(V), 0, V * sizeof(uint32_t), gz_stream));

    const gz_vec_t* vecs = gz_graph.pull_vectors;
    uint64_t chunk_len = gz_chunk_len();
    double* props = (double*)gz_graph.vertex_props;
    double* acc = (double*)gz_graph.vertex_accumulators;
    const double* outdeg = gz_graph.vertex_outdegrees;
    mergeaccum_t* merge = gz_merge_buffer;
    double* partials = (double*)gz_scratch.partials;
    double* slots = (double*)gz_scratch.reduce_tmp;
    uint32_t reduce_blocks = (uint32_t)gz_reduce_blocks(nchunks);
    uint32_t* run_head = gz_pr_run_head;
    double damping = gz_phase_config.damping;
    gz_grid_bar_t* bar = gz_pr_bar;
    uint32_t spin = gz_pr_persistent_spin;
    uint64_t n_vec = nvec, n_chunks = nchunks, n_v = V, iters = iterations;
    void* args[] = {&vecs, &n_vec, &chunk_len, &n_chunks, &props, &acc, &outdeg, &n_v, &merge, &partials,
                    &slots, &reduce_blocks, &run_head, &reduce_buffer, &damping, &iters, &bar, &spin};
    /* A cooperative launch can be refused while another process holds the
     * device's cooperative resources; the per-phase launches then run, for
     * the rest of this process. */
    const hipError_t err = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(GZ_BLOCK), args, 0, gz_stream);
    if (err != hipSuccess) {
        (void)hipGetLastError();
        gz_pr_persistent_unavailable = 1;
        fprintf(stderr, "grazelle: one-launch PageRank unavailable (%s); per-phase launches from now on\n",
                hipGetErrorString(err));
        return 0;
    }
    GZ_CHECK(hipMemcpyAsync(gz_pr_bar_host, &gz_pr_bar->abort, sizeof(uint32_t), hipMemcpyDeviceToHost, gz_stream));
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    if (*gz_pr_bar_host != 0) {
        gz_pr_persistent_aborts += 1;
        fprintf(stderr, "grazelle: a grid barrier of the one-launch PageRank gave up waiting; "
                        "the run restarts with per-phase launches\n");
        return -1;
    }
    gz_pr_persistent_count += 1;
    return 1;
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
    double* __restrict__ props, const uint32_t* __restrict__ outdeg32, uint64_t start, uint64_t end,
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
        const double od = mine ? (double)outdeg32[v] : 1.0;
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
        (const uint32_t*)gz_slices.outdeg32, vertex_start, end, gz_graph.num_vertices, reduce_buffer,            \
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
