/*
 * gz_slices.hip -- PageRank pull over the source-sliced layout.
 *
 * The scheduler-aware pull (phase_edge_pull_pr.asm:189-307) gathers one FP64
 * property per edge from anywhere in the vertex array.  On a power-law graph
 * renumbered by degree, the few million hottest sources carry most edges, but
 * their properties exceed one XCD's L2, and every XCD caches the same hot
 * lines.  Here the pull list is split by source into levels of eight slices
 * (tiers of S-vertex windows, and LDS blocks carved from the tier-0 windows)
 * and a cold rest; the chunk table hands XCD x (workgroups w with w % 8 == x,
 * the dispatcher's round robin) only the chunks of its slices x, 8 + x, ...,
 * level after level, plus an even share of the cold slice, so each XCD
 * gathers from one window at a time, and from LDS alone in a block slice.
 *
 * Each slice is an ordinary pull list whose destinations are numbered
 * densely inside the slice, stored packed (24-bit window offsets in the
 * window slices, 16-bit block offsets in the block slices, 32-bit ids in the
 * last; destinations implied by flags: 12, 8 or 16 bytes per vector instead
 * of 32), walked exactly as in gz_pr.hip
 * into its own accumulator run (complete runs stored directly, the chunk's last run into
 * its merge record).  The merge then runs per slice
 * (edge_pull_op_merge_with_merge_buffer's rule, phases.c:41-70), and the
 * vertex phase takes acc[v] = the sum over the slices holding v, in slice
 * order (gz_k_vertex_pr_sliced reads the slices' accumulators directly, so the
 * accumulator array itself is not written).  No atomics, so the result is
 * reproducible run to run; the per-vertex sum is regrouped by slice, which
 * moves it by rounding only.
 */
#include "grazelle_host.h"
#include "gz_host_internal.h"
#include "gz_pr_ops.h"
#include "gz_walk.h"

#include <string.h>

gz_source_slices_t gz_slices;
/* the per-XCD work counters sit 256 bytes apart, each in a line of its own */
#define GZ_COUNTER_STRIDE 64
static uint32_t gz_slices_cold_spread = 1;
static uint32_t gz_slices_cold_stream = 0;
static uint32_t gz_slices_hot_cache = 0;

/* ---- chunk table (host) ---------------------------------------------------- */

extern "C" uint64_t gz_slices_chunk_table(const uint64_t* vectors, uint32_t levels, uint64_t chunk_len,
                                          uint32_t cold_spread, gz_slice_chunk_t* out, uint64_t capacity)
{
    if (chunk_len == 0 || levels < 1 || levels > GZ_MAX_LEVELS) return 0;
    const int ns = GZ_XCDS * (int)levels + 1, cold_slice = ns - 1;
    uint64_t base[GZ_MAX_SLICES + 1], chunks[GZ_MAX_SLICES], rec[GZ_MAX_SLICES + 1];
    base[0] = 0;
    rec[0] = 0;
    for (int x = 0; x < ns; ++x) {
        chunks[x] = (vectors[x] + chunk_len - 1) / chunk_len;
        base[x + 1] = base[x] + vectors[x];
        rec[x + 1] = rec[x] + chunks[x];
    }
    if (rec[ns] >= (1ull << GZ_SLICE_RECORD_BITS)) {
        fprintf(stderr, "grazelle: %llu chunks exceed the source-sliced record field\n", (unsigned long long)rec[ns]);
        exit(72);
    }
    /* an XCD's own chunks: its slice in every level */
    uint64_t own[GZ_XCDS];
    for (int x = 0; x < GZ_XCDS; ++x) {
        own[x] = 0;
        for (uint32_t t = 0; t < levels; ++t) own[x] += chunks[GZ_XCDS * t + x];
    }
    /* chunks per XCD: an even share of all chunks, at least its own */
    uint64_t target = (rec[ns] + GZ_XCDS - 1) / GZ_XCDS;
    for (int x = 0; x < GZ_XCDS; ++x)
        if (own[x] > target) target = own[x];
    uint64_t give[GZ_XCDS], offered = 0;
    for (int x = 0; x < GZ_XCDS; ++x) {
        give[x] = target - own[x];
        offered += give[x];
    }
    /* take back what the cold slice cannot fill, one chunk per XCD in turn
     * from the XCDs that were offered the most */
    uint64_t excess = offered - chunks[cold_slice];
    while (excess > 0) {
        uint64_t most = 0;
        for (int x = 0; x < GZ_XCDS; ++x)
            if (give[x] > most) most = give[x];
        for (int x = 0; x < GZ_XCDS && excess > 0; ++x)
            if (give[x] == most) {
                give[x] -= 1;
                excess -= 1;
            }
    }
    uint64_t len[GZ_XCDS], blocks = 0;
    for (int x = 0; x < GZ_XCDS; ++x) {
        len[x] = own[x] + give[x];
        const uint64_t bx = (len[x] + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
        if (bx > blocks) blocks = bx;
    }
    const uint64_t slots = blocks * GZ_XCDS * GZ_WAVES_PER_BLOCK;
    if (out == nullptr) return slots;
    if (capacity < slots) {
        fprintf(stderr, "grazelle: chunk table needs %llu slots, %llu given\n", (unsigned long long)slots,
                (unsigned long long)capacity);
        exit(72);
    }
    memset(out, 0, slots * sizeof(gz_slice_chunk_t));
    uint64_t cold_next = 0;
    for (int x = 0; x < GZ_XCDS; ++x) {
        uint32_t level = 0;
        uint64_t hot_next = 0;
        const uint64_t n = len[x], g = give[x];
        for (uint64_t i = 0; i < n; ++i) {
            /* cold positions: where floor(i * g / n) steps (spread), or the
             * last g positions */
            const bool cold = cold_spread ? ((i + 1) * g / n != i * g / n) : (i >= n - g);
            int sl;
            uint64_t k;
            if (cold) {
                sl = cold_slice;
                k = cold_next++;
            } else {
                while (hot_next >= chunks[GZ_XCDS * level + x]) {
                    hot_next = 0;
                    ++level;
                }
                sl = GZ_XCDS * (int)level + x;
                k = hot_next++;
            }
            gz_slice_chunk_t c;
            c.begin = base[sl] + k * chunk_len;
            const uint64_t left = vectors[sl] - k * chunk_len;
            c.count = (uint32_t)(left < chunk_len ? left : chunk_len);
            c.record = ((uint32_t)sl << GZ_SLICE_RECORD_BITS) | (uint32_t)(rec[sl] + k);
            /* the XCD's i-th chunk: workgroup 8 * (i / 4) + x, wave i % 4 */
            const uint64_t wg = (uint64_t)GZ_XCDS * (i / GZ_WAVES_PER_BLOCK) + (uint64_t)x;
            out[wg * GZ_WAVES_PER_BLOCK + i % GZ_WAVES_PER_BLOCK] = c;
        }
    }
    return slots;
}

extern "C" void gz_slices_set_cold_spread(uint32_t cold_spread) { gz_slices_cold_spread = cold_spread ? 1u : 0u; }
extern "C" void gz_slices_set_cold_stream(uint32_t cold_stream) { gz_slices_cold_stream = cold_stream ? 1u : 0u; }

extern "C" int gz_slices_set_hot_cache(uint32_t vertices)
{
    if (vertices != 0 && vertices != 4096 && vertices != 8192 && vertices != 16384 && vertices != 20480) return -1;
    gz_slices_hot_cache = vertices;
    return 0;
}

extern "C" void gz_slices_settings(uint32_t* out3)
{
    out3[0] = gz_slices_hot_cache;
    out3[1] = gz_slices_cold_spread;
    out3[2] = gz_slices_cold_stream;
}

extern "C" void gz_source_slices_release(void)
{
    gz_device_free(gz_slices.nbr24);
    gz_device_free(gz_slices.nbr16);
    gz_device_free(gz_slices.nbr);
    gz_device_free(gz_slices.shared);
    gz_device_free(gz_slices.mask);
    gz_device_free(gz_slices.parts);
    gz_device_free(gz_slices.rank_base);
    gz_device_free(gz_slices.outdeg32);
    gz_device_free(gz_slices.gid);
    gz_device_free(gz_slices.table);
    gz_device_free(gz_slices.merge);
    gz_device_free(gz_slices.counters);
    memset(&gz_slices, 0, sizeof(gz_slices));
}

extern "C" int gz_graph_source_slices_info(uint64_t* vectors, uint64_t* slice_vertices, uint32_t* tiers,
                                           uint32_t* blocks, uint64_t* block_vertices)
{
    if (gz_slices.nbr == nullptr) return 0;
    if (vectors)
        for (uint32_t x = 0; x < gz_slices.nslices; ++x) vectors[x] = gz_slices.begin[x + 1] - gz_slices.begin[x];
    if (slice_vertices) *slice_vertices = gz_slices.slice_vertices;
    if (tiers) *tiers = gz_slices.tiers;
    if (blocks) *blocks = gz_slices.blocks;
    if (block_vertices) *block_vertices = gz_slices.block_vertices;
    return 1;
}

extern "C" int gz_graph_source_slices_download(gz_vec_t* vectors, uint64_t* mask)
{
    if (gz_slices.nbr == nullptr) return 0;
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    if (vectors) {
        /* back to Vector-Sparse form: the packed copy keeps the neighbour ids
         * and the shared vertex, which is all a vector holds */
        const uint64_t n = gz_slices.begin[gz_slices.nslices];
        const uint64_t hot = gz_slices.begin[gz_slices.nslices - 1];
        const uint64_t span = gz_slices.block_span, bfirst = gz_slices.begin[GZ_XCDS];
        const uint64_t windows = hot - span;
        uint32_t* w24 = (uint32_t*)malloc((windows ? 3 * windows : 1) * sizeof(uint32_t));
        uint64_t* w16 = (uint64_t*)malloc((span ? span : 1) * sizeof(uint64_t));
        uint32_t* nbr = (uint32_t*)malloc((n - hot ? n - hot : 1) * 4 * sizeof(uint32_t));
        uint32_t* shared = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
        gz_copy_to_host(w24, gz_slices.nbr24, 3 * windows * sizeof(uint32_t));
        gz_copy_to_host(w16, gz_slices.nbr16, span * sizeof(uint64_t));
        gz_copy_to_host(nbr, gz_slices.nbr, (n - hot) * 4 * sizeof(uint32_t));
        gz_copy_to_host(shared, gz_slices.shared, n * sizeof(uint32_t));
        uint32_t x = 0;
        for (uint64_t j = 0; j < n; ++j) {
            while (j >= gz_slices.begin[x + 1]) ++x;
            const uint32_t level = x / GZ_XCDS;
            uint64_t ids[4];
            unsigned count = 0;
            if (j < hot && level >= 1 && level <= gz_slices.blocks) {
                const uint64_t q = w16[j - bfirst];
                const uint32_t f[4] = {(uint32_t)q & 0x7fffu, (uint32_t)(q >> 16) & 0xffffu,
                                       (uint32_t)(q >> 32) & 0xffffu, (uint32_t)(q >> 48)};
                const uint64_t window = gz_slice_window(x, gz_slices.slice_vertices, gz_slices.blocks) +
                                        (uint64_t)level * gz_slices.block_vertices;
                while (count < 4 && (count == 0 || f[count] != GZ_PACKED16_NONE)) {
                    ids[count] = window + f[count];
                    ++count;
                }
            } else if (j < hot) {
                const uint32_t* q = w24 + 3 * (j - (level > gz_slices.blocks ? span : 0));
                const uint32_t f[4] = {q[0] & 0x7fffffu, (q[0] >> 24) | ((q[1] & 0xffffu) << 8),
                                       (q[1] >> 16) | ((q[2] & 0xffu) << 16), q[2] >> 8};
                const uint64_t window = gz_slice_window(x, gz_slices.slice_vertices, gz_slices.blocks);
                while (count < 4 && (count == 0 || f[count] != 0xffffffu)) {
                    ids[count] = window + f[count];
                    ++count;
                }
            } else {
                const uint32_t* q = nbr + 4 * (j - hot);
                while (count < 4 && (count == 0 || q[count] != GZ_PACKED_NONE)) {
                    ids[count] = q[count] & (count == 0 ? 0x7fffffffu : 0xffffffffu);
                    ++count;
                }
            }
            gz_vec_encode(&vectors[j], shared[j], ids, count);
        }
        free(w24);
        free(w16);
        free(nbr);
        free(shared);
    }
    if (mask) {
        const uint64_t nv = gz_graph.num_vertices;
        if (gz_slices.nslices > 32) {
            gz_copy_to_host(mask, gz_slices.mask, nv * sizeof(uint64_t));
        } else {
            uint32_t* m = (uint32_t*)malloc((nv ? nv : 1) * sizeof(uint32_t));
            gz_copy_to_host(m, gz_slices.mask, nv * sizeof(uint32_t));
            for (uint64_t v = 0; v < nv; ++v) mask[v] = m[v];
            free(m);
        }
    }
    return 1;
}

/* Build (or rebuild for a new chunk length) the chunk table and the merge
 * records it points into. */
static void gz_slices_prepare_table(uint64_t chunk_len)
{
    if (gz_slices.table != nullptr && gz_slices.table_chunk_len == chunk_len &&
        gz_slices.table_cold_spread == gz_slices_cold_spread)
        return;
    const uint32_t ns = gz_slices.nslices;
    uint64_t vectors[GZ_MAX_SLICES];
    for (uint32_t x = 0; x < ns; ++x) vectors[x] = gz_slices.begin[x + 1] - gz_slices.begin[x];
    const uint32_t levels = gz_slices.tiers + gz_slices.blocks;
    const uint64_t slots = gz_slices_chunk_table(vectors, levels, chunk_len, gz_slices_cold_spread, nullptr, 0);
    gz_slice_chunk_t* h = (gz_slice_chunk_t*)malloc((slots ? slots : 1) * sizeof(gz_slice_chunk_t));
    gz_slices_chunk_table(vectors, levels, chunk_len, gz_slices_cold_spread, h, slots);
    uint64_t records = 0;
    for (uint32_t x = 0; x < ns; ++x) records += (vectors[x] + chunk_len - 1) / chunk_len;
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_device_free(gz_slices.table);
    gz_device_free(gz_slices.merge);
    gz_slices.table = (gz_slice_chunk_t*)gz_device_malloc((slots ? slots : 1) * sizeof(gz_slice_chunk_t));
    gz_copy_to_device(gz_slices.table, h, slots * sizeof(gz_slice_chunk_t));
    gz_slices.merge = (mergeaccum_t*)gz_device_malloc((records ? records : 1) * sizeof(mergeaccum_t));
    gz_slices.records = records;
    for (int x = 0; x < GZ_XCDS; ++x) gz_slices.xcd_chunks[x] = 0;
    for (uint64_t i = 0; i < slots; ++i)
        if (h[i].count) gz_slices.xcd_chunks[(i / GZ_WAVES_PER_BLOCK) % GZ_XCDS] += 1;
    if (gz_slices.counters == nullptr)
        gz_slices.counters = (uint32_t*)gz_device_malloc(GZ_XCDS * GZ_COUNTER_STRIDE * sizeof(uint32_t));
    free(h);
    gz_slices.slots = slots;
    gz_slices.table_chunk_len = chunk_len;
    gz_slices.table_cold_spread = gz_slices_cold_spread;
}

/* ---- kernels ---------------------------------------------------------------- */

struct gz_slice_bounds {
    uint64_t rec[GZ_MAX_SLICES + 1];
};

/* The packed slices as kernel arguments: the window slices in 24-bit form,
 * the block slices in 16-bit form, the last slice in 32-bit form. */
struct gz_packed_lists {
    const uint32_t* nbr24;
    const uint64_t* nbr16;
    const gz_u32x4* nbr;
    const uint32_t* shared;
    uint64_t cold_first;
    uint64_t block_first; /* first vector of the block slices */
    uint64_t block_span;  /* their vector count */
    uint64_t slice_vertices;
    uint64_t block_vertices;
    uint32_t cold_slice;
    uint32_t blocks;
    __device__ __forceinline__ bool is_block(uint32_t sl) const
    {
        const uint32_t level = sl / GZ_XCDS;
        return sl != cold_slice && level >= 1 && level <= blocks;
    }
    __device__ __forceinline__ gz_packed_list cold() const { return gz_packed_list{{shared}, nbr, cold_first}; }
    __device__ __forceinline__ gz_packed24_list window(uint32_t sl) const
    {
        return gz_packed24_list{{shared}, nbr24, (uint32_t)gz_slice_window(sl, slice_vertices, blocks),
                                (sl / GZ_XCDS > blocks) ? block_span : 0ull};
    }
    __device__ __forceinline__ gz_packed16_list block(uint32_t sl) const
    {
        return gz_packed16_list{
            {shared}, nbr16,
            (uint32_t)(gz_slice_window(sl, slice_vertices, blocks) + (uint64_t)(sl / GZ_XCDS) * block_vertices),
            block_first};
    }
};

static gz_packed_lists gz_slices_lists()
{
    gz_packed_lists p;
    p.nbr24 = gz_slices.nbr24;
    p.nbr = gz_slices.nbr;
    p.shared = gz_slices.shared;
    p.cold_first = gz_slices.begin[gz_slices.nslices - 1];
    p.slice_vertices = gz_slices.slice_vertices;
    p.cold_slice = gz_slices.nslices - 1;
    p.blocks = gz_slices.blocks;
    p.nbr16 = gz_slices.nbr16;
    p.block_first = gz_slices.begin[GZ_XCDS];
    p.block_span = gz_slices.block_span;
    p.block_vertices = gz_slices.block_vertices;
    return p;
}

/* Walk one chunk into slice accumulators `acc`; returns the wave's share of S. */
template <int VPL, bool kStream, typename L>
__device__ __forceinline__ double gz_sliced_walk(const L& list, uint64_t b, uint64_t e,
                                                 const double* __restrict__ props, double* __restrict__ acc,
                                                 mergeaccum_t* __restrict__ record, uint32_t& edges)
{
    gz_pr_pull_op_t<kStream> op{props, acc, true, 0.0};
    uint64_t first_d, carry_d;
    double carry_v;
    gz_walk_list<VPL, double, false>(list, b, e, op, first_d, carry_d, carry_v, edges);
    if (gz_lane() == 0) {
        mergeaccum_t r;
        r.initial_vertex_id = first_d;
        r.final_vertex_id = carry_d;
        r.final_partial_value = carry_v;
        r.__padding__ = 0;
        *record = r;
    }
    return gz_wave_sum(op.sum);
}

/* One wavefront per chunk slot.  kColdStream: the cold slice's gathers, which
 * mostly miss, are nontemporal. */
template <int VPL, bool kColdStream>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_edge_pull_pr_sliced(gz_packed_lists lists,
                                                                     const gz_slice_chunk_t* __restrict__ table,
                                                                     const double* __restrict__ props,
                                                                     double* __restrict__ parts, gz_slice_offsets part_base,
                                                                     uint32_t cold_slice,
                                                                     mergeaccum_t* __restrict__ merge,
                                                                     double* __restrict__ partials,
                                                                     uint64_t* __restrict__ stats, uint32_t collect)
{
    const int lane = gz_lane();
    const uint64_t slot = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const gz_slice_chunk_t c = table[slot];
    if (c.count == 0) {
        if (lane == 0) {
            partials[slot] = 0.0;
            if (collect) {
                stats[2 * slot] = 0;
                stats[2 * slot + 1] = 0;
            }
        }
        return;
    }
    const uint32_t x = c.record >> GZ_SLICE_RECORD_BITS;
    const uint64_t rec = c.record & ((1u << GZ_SLICE_RECORD_BITS) - 1u);
    const uint64_t b = c.begin, e = c.begin + c.count;
    double* acc = parts + part_base.off[x];
    uint32_t edges = 0;
    double S;
    if (x == cold_slice)
        S = kColdStream ? gz_sliced_walk<VPL, true>(lists.cold(), b, e, props, acc, merge + rec, edges)
                        : gz_sliced_walk<VPL, false>(lists.cold(), b, e, props, acc, merge + rec, edges);
    else if (lists.is_block(x))
        S = gz_sliced_walk<VPL, false>(lists.block(x), b, e, props, acc, merge + rec, edges);
    else
        S = gz_sliced_walk<VPL, false>(lists.window(x), b, e, props, acc, merge + rec, edges);
    if (lane == 0) partials[slot] = S;
    if (collect) {
        const uint64_t tot_edges = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * slot] = c.count;
            stats[2 * slot + 1] = tot_edges;
        }
    }
}

/* The merge of gz_k_merge_sum applied to each slice's own records and
 * accumulators: runs never cross a slice boundary. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_merge_sliced(const mergeaccum_t* __restrict__ mb,
                                                              gz_slice_bounds bounds, uint32_t nslices,
                                                              double* __restrict__ parts, gz_slice_offsets part_base)
{
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= bounds.rec[nslices]) return;
    uint32_t x = 0;
    while (u >= bounds.rec[x + 1]) ++x;
    const uint64_t lo = bounds.rec[x], hi = bounds.rec[x + 1];
    double* acc = parts + part_base.off[x];
    const mergeaccum_t r = mb[u];
    if (u > lo && mb[u - 1].final_vertex_id == r.final_vertex_id) return;
    double proposed = r.final_partial_value;
    uint64_t j = u + 1;
    while (j < hi && mb[j].final_vertex_id == r.final_vertex_id) {
        proposed = proposed + mb[j].final_partial_value;
        ++j;
    }
    if (j < hi && mb[j].initial_vertex_id == r.final_vertex_id) proposed = proposed + acc[r.final_vertex_id];
    acc[r.final_vertex_id] = proposed;
}

/* ---- persistent pull with the hottest sources of each XCD in LDS -------------
 * As many workgroups as fit on the device at once (LDS and registers
 * permitting) each copy the first `hot` properties of their XCD's tier-0
 * window into LDS; their wavefronts then take that XCD's chunks from a
 * per-XCD counter.  Gathers that fall in the cached range read LDS instead of
 * issuing an L2 request.  When the layout has LDS blocks no larger than the
 * cache, the workgroup moves its LDS on to block b once its wavefronts reach
 * block b's chunks (the queue holds them level by level), and a block
 * slice's gathers then all read LDS.  Workgroups have 1024 threads, so that
 * 16 wavefronts share one cache; a cache of at most 64 KiB leaves room for
 * two workgroups per CU, and the kernel is then held to the registers that
 * allow both. */

/* Move a workgroup's LDS copy on to block `target` of its XCD's tier-0
 * window, block by block: a barrier (every wavefront is done with the block
 * held so far), the copy of the next block's H values, a barrier (the copy is
 * complete).  Every wavefront of the workgroup steps through the same blocks
 * (when it takes a chunk of a later level, or at its exit), so all of them
 * reach each barrier; `cb` is wave-uniform. */
template <typename T>
__device__ __forceinline__ void gz_lds_advance(T* hot, const T* __restrict__ src, uint64_t lo, uint64_t H, uint32_t& cb,
                                               uint32_t target)
{
    while (cb < target) {
        __syncthreads();
        ++cb;
        const uint64_t base = lo + (uint64_t)cb * H;
        for (uint64_t i = threadIdx.x; i < H; i += blockDim.x) hot[i] = src[base + i];
        __syncthreads();
    }
}

template <int HOT>
struct gz_hot_wg {
    static constexpr int value = 1024;
    /* wavefronts per SIMD the registers must allow: two workgroups' worth
     * when two caches fit in the CU's LDS */
    static constexpr int waves_per_simd = HOT <= 8192 ? 8 : 4;
};

struct gz_xcd_lens {
    uint64_t len[GZ_XCDS];
};

/* Pointers with their address space spelled out: an LDS read and a global
 * read stay two instructions (each under its own lane mask) instead of being
 * folded into one generic (flat) load through a selected pointer. */
typedef __attribute__((address_space(3))) const double gz_lds_f64;
typedef __attribute__((address_space(1))) const double gz_global_f64;

/* The four sources of one packed vector (ids s[k], GZ_PACKED_NONE for an
 * empty lane), each read from the LDS copy `hot` when it lies in
 * [wlo, wlo + wn) and from memory otherwise; `empty` for an empty or
 * out-of-range lane.  The memory reads are all issued before any LDS read and
 * land in registers of their own.  (Written as one if/else per source, both
 * branches share the result register, so each LDS read must wait for the
 * memory read issued before it and a wavefront holds one gather in flight
 * instead of four.) */
template <bool kStream, bool kAllCached, typename T, typename GP, typename LP>
__device__ __forceinline__ void gz_fetch4_hot(GP mem, LP hot, uint32_t wlo, uint32_t wn, const gz_packed_vec& v,
                                              bool in_range, T empty, T (&out)[4], uint32_t& edges)
{
    const uint32_t s[4] = {v.s0, v.s1, v.s2, v.s3};
    if constexpr (kAllCached) {
        /* a block slice with its block in LDS: every source is in
         * [wlo, wlo + wn), and nothing is read from memory */
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ok = in_range && s[k] != GZ_PACKED_NONE;
            const T h = hot[ok ? s[k] - wlo : 0u];
            out[k] = ok ? h : empty;
            edges += (uint32_t)ok;
        }
        return;
    }
    bool ok[4], cached[4];
    uint32_t r[4];
    T m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ok[k] = in_range && s[k] != GZ_PACKED_NONE;
        r[k] = s[k] - wlo;
        cached[k] = ok[k] && r[k] < wn;
        m[k] = empty;
        edges += (uint32_t)ok[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (ok[k] && !cached[k]) m[k] = kStream ? __builtin_nontemporal_load(mem + s[k]) : mem[s[k]];
    if (wn != 0) {
        T h[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) h[k] = hot[cached[k] ? r[k] : 0u];
        /* every lane reads LDS (source 0 of the window where the source is
         * not cached): left to itself the compiler sinks each read into a
         * branch on cached[k] that first waits for the memory reads */
        asm volatile("" ::"v"(h[0]), "v"(h[1]), "v"(h[2]), "v"(h[3]));
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = cached[k] ? h[k] : m[k];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = m[k];
    }
}

template <bool kStream, bool kAllCached = false>
struct gz_pr_pull_op_hot {
    gz_global_f64* __restrict__ props;
    gz_lds_f64* hot;
    uint32_t wlo, wn;
    double* __restrict__ acc;
    double sum;
    __device__ __forceinline__ double identity() const { return 0.0; }
    __device__ __forceinline__ double combine(double a, double b) const { return a + b; }
    /* gz_pr_packed_sum with the cache in front */
    __device__ __forceinline__ double gather(const gz_packed_vec& v, bool in_range, uint32_t& edges)
    {
        double g[4];
        gz_fetch4_hot<kStream, kAllCached>(props, hot, wlo, wn, v, in_range, 0.0, g, edges);
        const double t = (g[2] + g[3]) + (g[0] + g[1]);
        sum += t;
        return t;
    }
    __device__ __forceinline__ void store(uint64_t d, double s) { acc[d] = s; }
};

template <int HOT, int VPL>
__global__ __launch_bounds__(gz_hot_wg<HOT>::value, gz_hot_wg<HOT>::waves_per_simd) void gz_k_edge_pull_pr_sliced_hot(
    gz_packed_lists lists, const gz_slice_chunk_t* __restrict__ table, gz_xcd_lens lens,
    const double* __restrict__ props, double* __restrict__ parts, gz_slice_offsets part_base,
    mergeaccum_t* __restrict__ merge, double* __restrict__ partials, uint64_t* __restrict__ stats, uint32_t collect,
    uint32_t* __restrict__ counters, uint64_t slice_vertices, uint64_t hot_count, uint32_t cold_slice,
    uint32_t cold_stream, uint64_t block_vertices)
{
    __shared__ double hot[HOT];
    const uint32_t x = blockIdx.x % GZ_XCDS;
    const uint64_t lo = (uint64_t)x * slice_vertices;
    for (uint64_t i = threadIdx.x; i < hot_count; i += blockDim.x) hot[i] = props[lo + i];
    __syncthreads();
    /* LDS blocks, when they fit the cache; cb = the block the cache holds */
    const uint32_t G = (block_vertices != 0 && block_vertices <= (uint64_t)HOT) ? lists.blocks : 0u;
    uint32_t cb = 0;
    const int lane = gz_lane();
    const uint64_t n = lens.len[x];
    while (true) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(&counters[x * GZ_COUNTER_STRIDE], 1u);
        /* lane 0's ticket in every lane, and known uniform, so that the chunk's
         * table entry and bounds are scalar */
        idx = (uint32_t)__builtin_amdgcn_readfirstlane((int)idx);
        if ((uint64_t)idx >= n) break;
        const uint64_t slot =
            ((uint64_t)GZ_XCDS * (idx / GZ_WAVES_PER_BLOCK) + x) * GZ_WAVES_PER_BLOCK + idx % GZ_WAVES_PER_BLOCK;
        const gz_slice_chunk_t c = table[slot];
        const uint32_t sl = c.record >> GZ_SLICE_RECORD_BITS;
        const uint64_t rec = c.record & ((1u << GZ_SLICE_RECORD_BITS) - 1u);
        uint64_t first_d, carry_d;
        double carry_v;
        uint32_t edges = 0;
        double sum;
        gz_global_f64* gprops = (gz_global_f64*)props;
        gz_lds_f64* lds = (gz_lds_f64*)hot;
        if (sl == cold_slice) {
            if (cold_stream) {
                gz_pr_pull_op_hot<true> op{gprops, lds, 0, 0, parts + part_base.off[sl], 0.0};
                gz_walk_list<VPL, double, true, true>(lists.cold(), c.begin, c.begin + c.count, op, first_d, carry_d,
                                                carry_v, edges);
                sum = op.sum;
            } else {
                gz_pr_pull_op_hot<false> op{gprops, lds, 0, 0, parts + part_base.off[sl], 0.0};
                gz_walk_list<VPL, double, true, true>(lists.cold(), c.begin, c.begin + c.count, op, first_d, carry_d,
                                                carry_v, edges);
                sum = op.sum;
            }
        } else {
            const uint32_t level = sl / GZ_XCDS;
            if (level > cb && cb < G) gz_lds_advance(hot, props, lo, block_vertices, cb, level < G ? level : G);
            /* the cached range: block 0 for the window, the block itself for a
             * block slice (every source then hits LDS) */
            double* acc = parts + part_base.off[sl];
            if (lists.is_block(sl) && level == cb) {
                gz_pr_pull_op_hot<false, true> op{gprops, lds, (uint32_t)(lo + (uint64_t)level * block_vertices),
                                                  (uint32_t)block_vertices, acc, 0.0};
                gz_walk_list<VPL, double, true, true>(lists.block(sl), c.begin, c.begin + c.count, op, first_d,
                                                      carry_d, carry_v, edges);
                sum = op.sum;
            } else if (lists.is_block(sl)) { /* blocks larger than the cache */
                gz_pr_pull_op_hot<false> op{gprops, lds, 0u, 0u, acc, 0.0};
                gz_walk_list<VPL, double, true, true>(lists.block(sl), c.begin, c.begin + c.count, op, first_d,
                                                      carry_d, carry_v, edges);
                sum = op.sum;
            } else {
                const bool cached = level == 0 && cb == 0;
                gz_pr_pull_op_hot<false> op{gprops, lds, cached ? (uint32_t)lo : 0u,
                                            cached ? (uint32_t)hot_count : 0u, acc, 0.0};
                gz_walk_list<VPL, double, true, true>(lists.window(sl), c.begin, c.begin + c.count, op, first_d,
                                                      carry_d, carry_v, edges);
                sum = op.sum;
            }
        }
        const double S = gz_wave_sum(sum);
        if (lane == 0) {
            mergeaccum_t r;
            r.initial_vertex_id = first_d;
            r.final_vertex_id = carry_d;
            r.final_partial_value = carry_v;
            r.__padding__ = 0;
            merge[rec] = r;
            partials[rec] = S;
        }
        if (collect) {
            const uint64_t tot_edges = gz_wave_sum_u64(edges);
            if (lane == 0) {
                stats[2 * rec] = c.count;
                stats[2 * rec + 1] = tot_edges;
            }
        }
    }
    gz_lds_advance(hot, props, lo, block_vertices, cb, G);
}

/* One workgroup per XCD and resident slot: the grid is exactly what fits. */
template <int HOT, int VPL>
static void gz_launch_hot_vpl(const gz_xcd_lens& lens, uint64_t hot_count)
{
    constexpr int WG = gz_hot_wg<HOT>::value;
    int per_cu = 0;
    GZ_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gz_k_edge_pull_pr_sliced_hot<HOT, VPL>, WG, 0));
    if (per_cu < 1) per_cu = 1;
    uint64_t per_xcd = gz_cu_count / GZ_XCDS;
    if (per_xcd == 0) per_xcd = 1;
    const uint64_t blocks = (uint64_t)GZ_XCDS * per_xcd * (uint64_t)per_cu;
    hipLaunchKernelGGL((gz_k_edge_pull_pr_sliced_hot<HOT, VPL>), dim3((unsigned)blocks), dim3(WG), 0, gz_stream,
                       gz_slices_lists(), gz_slices.table, lens, (const double*)gz_graph.vertex_props, gz_slices.parts,
                       gz_slices.part_base, gz_slices.merge, (double*)gz_scratch.partials, gz_scratch.stats,
                       gz_phase_config.collect_stats, gz_slices.counters, gz_slices.slice_vertices, hot_count,
                       gz_slices.nslices - 1, gz_slices_cold_stream, gz_slices.block_vertices);
    GZ_LAUNCH_CHECK();
}

template <int HOT>
static void gz_launch_hot(const gz_xcd_lens& lens, uint64_t hot_count)
{
    const uint32_t vpl = gz_vectors_per_lane();
    if (vpl == 4)
        gz_launch_hot_vpl<HOT, 4>(lens, hot_count);
    else if (vpl == 2)
        gz_launch_hot_vpl<HOT, 2>(lens, hot_count);
    else
        gz_launch_hot_vpl<HOT, 1>(lens, hot_count);
}

/* ---- Connected Components over the same layout ---------------------------------
 * The min walk of gz_k_edge_pull_cc (phase_edge_pull_cc.asm:50-325) over the
 * packed slices: each complete run of a slice-local destination applies its
 * minimum to the destination's global label with an atomic min, as every run
 * of the plain sweep does, setting HasInfo* and adding to the vote when the
 * label drops.  The LDS cache holds the window's labels as they were when the
 * workgroup started: a label lowered later in the sweep reaches these gathers
 * one sweep later, which changes the number of sweeps, never the labels (a
 * sweep that lowers no label found every label already final). */
typedef __attribute__((address_space(3))) const uint64_t gz_lds_u64;
typedef __attribute__((address_space(1))) const uint64_t gz_global_u64;

template <bool kAllCached = false>
struct gz_cc_pull_op_hot {
    gz_global_u64* __restrict__ in;      /* labels, read */
    gz_lds_u64* hot;
    uint32_t wlo, wn;
    uint64_t* __restrict__ labels;       /* labels, atomic min */
    uint64_t* __restrict__ next_frontier;
    const double* __restrict__ outdeg;
    const uint32_t* __restrict__ gid;    /* this slice's local rank -> vertex */
    uint32_t stat_mode;
    uint64_t vote;
    __device__ __forceinline__ uint64_t identity() const { return ~0ull; }
    __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) const { return a < b ? a : b; }
    __device__ __forceinline__ uint64_t gather(const gz_packed_vec& v, bool in_range, uint32_t& edges) const
    {
        uint64_t x[4];
        gz_fetch4_hot<false, kAllCached>(in, hot, wlo, wn, v, in_range, ~(uint64_t)0, x, edges);
        const uint64_t a = x[0] < x[1] ? x[0] : x[1], b = x[2] < x[3] ? x[2] : x[3];
        return a < b ? a : b;
    }
    __device__ __forceinline__ void store(uint64_t d, uint64_t m)
    {
        const uint64_t g = gid[d];
        if (m < labels[g]) {
            const unsigned long long old = atomicMin((unsigned long long*)&labels[g], (unsigned long long)m);
            if (old > m) {
                gz_bit_set(next_frontier, g);
                vote += gz_frontier_stat(outdeg, g, stat_mode);
            }
        }
    }
};

/* One chunk of the CC pull with operator `op` (by value; its runs apply
 * their minima); the chunk's last run is applied here.  Returns the vote. */
template <bool kAll, typename L>
__device__ __forceinline__ uint64_t gz_cc_walk(const L& list, const gz_slice_chunk_t& c, gz_cc_pull_op_hot<kAll> op,
                                               uint32_t& edges)
{
    uint64_t first_d, carry_d, carry_m;
    gz_walk_list<1, uint64_t, true, true>(list, c.begin, c.begin + c.count, op, first_d, carry_d, carry_m, edges);
    if (gz_lane() == 0) op.store(carry_d, carry_m);
    return op.vote;
}

template <int HOT>
__global__ __launch_bounds__(gz_hot_wg<HOT>::value, gz_hot_wg<HOT>::waves_per_simd) void gz_k_edge_pull_cc_sliced_hot(
    gz_packed_lists lists, const gz_slice_chunk_t* __restrict__ table, gz_xcd_lens lens, uint64_t* __restrict__ labels,
    uint64_t* __restrict__ next_frontier, const double* __restrict__ outdeg, uint32_t stat_mode,
    const uint32_t* __restrict__ gid, gz_slice_offsets part_base, uint64_t* __restrict__ partials,
    uint64_t* __restrict__ stats, uint32_t collect, uint32_t* __restrict__ counters, uint64_t slice_vertices,
    uint64_t hot_count, uint64_t block_vertices)
{
    __shared__ uint64_t hot[HOT];
    const uint32_t x = blockIdx.x % GZ_XCDS;
    const uint64_t lo = (uint64_t)x * slice_vertices;
    for (uint64_t i = threadIdx.x; i < hot_count; i += blockDim.x) hot[i] = labels[lo + i];
    __syncthreads();
    const uint32_t G = (block_vertices != 0 && block_vertices <= (uint64_t)HOT) ? lists.blocks : 0u;
    uint32_t cb = 0;
    const int lane = gz_lane();
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / GZ_WAVE) + threadIdx.x / GZ_WAVE;
    const uint64_t n = lens.len[x];
    uint64_t vote = 0;
    uint32_t edges = 0, vectors = 0;
    while (true) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(&counters[x * GZ_COUNTER_STRIDE], 1u);
        /* lane 0's ticket in every lane, and known uniform, so that the chunk's
         * table entry and bounds are scalar */
        idx = (uint32_t)__builtin_amdgcn_readfirstlane((int)idx);
        if ((uint64_t)idx >= n) break;
        const uint64_t slot =
            ((uint64_t)GZ_XCDS * (idx / GZ_WAVES_PER_BLOCK) + x) * GZ_WAVES_PER_BLOCK + idx % GZ_WAVES_PER_BLOCK;
        const gz_slice_chunk_t c = table[slot];
        const uint32_t sl = c.record >> GZ_SLICE_RECORD_BITS;
        const uint32_t level = sl / GZ_XCDS;
        uint32_t wlo = 0, wn = 0;
        if (sl != lists.cold_slice) {
            if (level > cb && cb < G) gz_lds_advance(hot, labels, lo, block_vertices, cb, level < G ? level : G);
            if (level == 0 && cb == 0) {
                wlo = (uint32_t)lo;
                wn = (uint32_t)hot_count;
            } else if (level >= 1 && level == cb) {
                wlo = (uint32_t)(lo + (uint64_t)level * block_vertices);
                wn = (uint32_t)block_vertices;
            }
        }
        const gz_cc_pull_op_hot<false> op{(gz_global_u64*)labels, (gz_lds_u64*)hot, wlo, wn, labels, next_frontier,
                                          outdeg, gid + part_base.off[sl], stat_mode, 0};
        if (sl == lists.cold_slice) {
            vote += gz_cc_walk(lists.cold(), c, op, edges);
        } else if (lists.is_block(sl)) {
            if (level == cb) {
                const gz_cc_pull_op_hot<true> all{op.in, op.hot, wlo, wn, labels, next_frontier,
                                                  outdeg, op.gid, stat_mode, 0};
                vote += gz_cc_walk(lists.block(sl), c, all, edges);
            } else {
                vote += gz_cc_walk(lists.block(sl), c, op, edges);
            }
        } else {
            vote += gz_cc_walk(lists.window(sl), c, op, edges);
        }
        vectors += c.count;
    }
    gz_lds_advance(hot, labels, lo, block_vertices, cb, G);
    const uint64_t total = gz_wave_sum_u64(vote);
    if (lane == 0) partials[wave] = total;
    if (collect) {
        const uint64_t tot_edges = gz_wave_sum_u64(edges);
        if (lane == 0) {
            stats[2 * wave] = vectors;
            stats[2 * wave + 1] = tot_edges;
        }
    }
}

template <int HOT>
static uint64_t gz_cc_hot_grid()
{
    int per_cu = 0;
    GZ_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gz_k_edge_pull_cc_sliced_hot<HOT>,
                                                          gz_hot_wg<HOT>::value, 0));
    if (per_cu < 1) per_cu = 1;
    uint64_t per_xcd = gz_cu_count / GZ_XCDS;
    if (per_xcd == 0) per_xcd = 1;
    return (uint64_t)GZ_XCDS * per_xcd * (uint64_t)per_cu;
}

template <int HOT>
static void gz_launch_cc_hot(const gz_xcd_lens& lens, uint64_t hot_count, uint64_t blocks)
{
    hipLaunchKernelGGL(gz_k_edge_pull_cc_sliced_hot<HOT>, dim3((unsigned)blocks), dim3(gz_hot_wg<HOT>::value), 0,
                       gz_stream, gz_slices_lists(), gz_slices.table, lens,
                       (uint64_t*)gz_graph.vertex_props, (uint64_t*)gz_graph.vertex_accumulators,
                       gz_graph.vertex_outdegrees, gz_phase_config.stat_mode, gz_slices.gid, gz_slices.part_base,
                       (uint64_t*)gz_scratch.partials, gz_scratch.stats, gz_phase_config.collect_stats,
                       gz_slices.counters, gz_slices.slice_vertices, hot_count, gz_slices.block_vertices);
    GZ_LAUNCH_CHECK();
}

void gz_pull_cc_sliced(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    if (gz_slices.nbr == nullptr || gz_slices.base_list != edge_list || gz_slices.base_count != edge_list_count) {
        fprintf(stderr,
                "grazelle: the source-sliced pull needs a sliced layout of this edge list "
                "(gz_graph_build_source_slices)\n");
        exit(72);
    }
    gz_slices_prepare_table(gz_chunk_len());
    /* without a configured cache the 32 KiB variant runs with an empty window */
    const uint32_t hot = gz_slices_hot_cache ? gz_slices_hot_cache : 4096;
    const uint64_t hot_count =
        gz_slices_hot_cache ? (hot < gz_slices.slice_vertices ? hot : gz_slices.slice_vertices) : 0;
    const uint64_t blocks = hot == 20480   ? gz_cc_hot_grid<20480>()
                            : hot == 16384 ? gz_cc_hot_grid<16384>()
                            : hot == 8192  ? gz_cc_hot_grid<8192>()
                                           : gz_cc_hot_grid<4096>();
    const uint64_t waves = blocks * (gz_hot_wg<4096>::value / GZ_WAVE);
    gz_scratch_reserve(waves, gz_graph.num_vertices);
    gz_scratch.partials_is_double = 0;
    gz_scratch.partials_count = waves;
    gz_scratch.stats_count = gz_phase_config.collect_stats ? waves : 0;
    gz_xcd_lens lens;
    for (int x = 0; x < GZ_XCDS; ++x) lens.len[x] = gz_slices.xcd_chunks[x];
    gz_device_zero(gz_slices.counters, GZ_XCDS * GZ_COUNTER_STRIDE * sizeof(uint32_t));
    gz_ktimer_begin(GZ_K_PULL_CC);
    if (hot == 20480)
        gz_launch_cc_hot<20480>(lens, hot_count, blocks);
    else if (hot == 16384)
        gz_launch_cc_hot<16384>(lens, hot_count, blocks);
    else if (hot == 8192)
        gz_launch_cc_hot<8192>(lens, hot_count, blocks);
    else
        gz_launch_cc_hot<4096>(lens, hot_count, blocks);
    gz_ktimer_end(GZ_K_PULL_CC);
}

/* ---- the phase ---------------------------------------------------------------- */

void gz_pull_pr_sliced(const gz_vec_t* edge_list, uint64_t edge_list_count)
{
    if (gz_slices.nbr == nullptr || gz_slices.base_list != edge_list || gz_slices.base_count != edge_list_count) {
        fprintf(stderr,
                "grazelle: the source-sliced pull needs a sliced layout of this edge list "
                "(gz_graph_build_source_slices)\n");
        exit(72);
    }
    const uint64_t chunk_len = gz_chunk_len();
    gz_slices_prepare_table(chunk_len);
    const uint64_t slots = gz_slices.slots;
    const uint32_t hot = gz_slices_hot_cache;
    const uint64_t waves = hot ? gz_slices.records : slots;
    gz_scratch_reserve(waves, gz_graph.num_vertices);
    gz_scratch.partials_is_double = 1;
    gz_scratch.partials_count = waves;
    gz_scratch.stats_count = gz_phase_config.collect_stats ? waves : 0;
    if (hot && gz_slices.records > 0) {
        const uint64_t hot_count = hot < gz_slices.slice_vertices ? hot : gz_slices.slice_vertices;
        gz_xcd_lens lens;
        for (int x = 0; x < GZ_XCDS; ++x) lens.len[x] = gz_slices.xcd_chunks[x];
        gz_device_zero(gz_slices.counters, GZ_XCDS * GZ_COUNTER_STRIDE * sizeof(uint32_t));
        gz_ktimer_begin(GZ_K_PULL_PR);
        if (hot == 20480)
            gz_launch_hot<20480>(lens, hot_count);
        else if (hot == 16384)
            gz_launch_hot<16384>(lens, hot_count);
        else if (hot == 8192)
            gz_launch_hot<8192>(lens, hot_count);
        else
            gz_launch_hot<4096>(lens, hot_count);
        gz_ktimer_end(GZ_K_PULL_PR);
    } else if (!hot && slots > 0) {
        const uint64_t blocks = slots / GZ_WAVES_PER_BLOCK;
        const uint32_t vpl = gz_vectors_per_lane();
        const uint32_t cold_slice = gz_slices.nslices - 1;
        gz_ktimer_begin(GZ_K_PULL_PR);
#define GZ_SLICED_ARGS                                                                                          \
    dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, gz_slices_lists(), gz_slices.table,                     \
        (const double*)gz_graph.vertex_props, gz_slices.parts, gz_slices.part_base, cold_slice, gz_slices.merge,  \
        (double*)gz_scratch.partials, gz_scratch.stats, gz_phase_config.collect_stats
        const bool cs = gz_slices_cold_stream != 0;
        if (vpl == 4 && cs)
            hipLaunchKernelGGL((gz_k_edge_pull_pr_sliced<4, true>), GZ_SLICED_ARGS);
        else if (vpl == 4)
            hipLaunchKernelGGL((gz_k_edge_pull_pr_sliced<4, false>), GZ_SLICED_ARGS);
        else if (vpl == 2 && cs)
            hipLaunchKernelGGL((gz_k_edge_pull_pr_sliced<2, true>), GZ_SLICED_ARGS);
        else if (vpl == 2)
            hipLaunchKernelGGL((gz_k_edge_pull_pr_sliced<2, false>), GZ_SLICED_ARGS);
        else if (cs)
            hipLaunchKernelGGL((gz_k_edge_pull_pr_sliced<1, true>), GZ_SLICED_ARGS);
        else
            hipLaunchKernelGGL((gz_k_edge_pull_pr_sliced<1, false>), GZ_SLICED_ARGS);
#undef GZ_SLICED_ARGS
        GZ_LAUNCH_CHECK();
        gz_ktimer_end(GZ_K_PULL_PR);
    }
    gz_ktimer_begin(GZ_K_MERGE);
    if (gz_slices.records > 0) {
        gz_slice_bounds bounds;
        memset(&bounds, 0, sizeof(bounds));
        for (uint32_t x = 0; x < gz_slices.nslices; ++x) {
            const uint64_t n = gz_slices.begin[x + 1] - gz_slices.begin[x];
            bounds.rec[x + 1] = bounds.rec[x] + (n + chunk_len - 1) / chunk_len;
        }
        const uint64_t mblocks = (gz_slices.records + GZ_BLOCK - 1) / GZ_BLOCK;
        hipLaunchKernelGGL(gz_k_merge_sliced, dim3((unsigned)mblocks), dim3(GZ_BLOCK), 0, gz_stream, gz_slices.merge,
                           bounds, gz_slices.nslices, gz_slices.parts, gz_slices.part_base);
        GZ_LAUNCH_CHECK();
    }
    gz_ktimer_end(GZ_K_MERGE);
}
