/*
 * gz_frontier.hip -- compaction of a frontier bitmap into the active-vertex
 * list consumed by the frontier-driven edge kernels (see gz_frontier.h).
 * Three passes: per-block counts, a single-block exclusive scan of the block
 * counts, and a write pass that re-derives each vertex's position.
 *
 * Layout: a block covers GZ_COMPACT_VPB = GZ_COMPACT_VPT x GZ_BLOCK
 * consecutive vertices, GZ_COMPACT_VPT per thread, so a wavefront covers a
 * few bitmap words and one contiguous range of offsets; list positions come
 * from one block scan of the per-thread counts, in vertex order.
 */
#include "gz_frontier.h"

/* Vertices v0 .. v0+GZ_COMPACT_VPT-1 (v0 a multiple of GZ_COMPACT_VPT):
 * vector counts c[k] of the vertices whose bit is set, 0 otherwise; a vertex
 * is active when c[k] > 0.  Neighbouring lanes read neighbouring bits of one
 * bitmap word.  kNarrow reads the 32-bit vector counts (gz_frontier_counts:
 * 64 bytes per thread in four 16-byte loads), else the 64-bit offsets. */
template <bool kNarrow>
__device__ __forceinline__ void gz_run_counts(const uint64_t* __restrict__ bitmap, const void* __restrict__ src,
                                              uint64_t num_vertices, uint64_t v0, uint64_t c[GZ_COMPACT_VPT])
{
#pragma unroll
    for (int k = 0; k < GZ_COMPACT_VPT; ++k) c[k] = 0;
    if (v0 >= num_vertices) return;
    uint64_t bits = (bitmap[v0 >> 6] >> (v0 & 63)) & ((1ull << GZ_COMPACT_VPT) - 1ull);
    const uint64_t left = num_vertices - v0;
    if (left < GZ_COMPACT_VPT) bits &= (1ull << left) - 1ull;
    if (bits == 0) return;
    if (kNarrow) {
        const uint4* q = (const uint4*)((const uint32_t*)src + v0);
#pragma unroll
        for (int k = 0; k < GZ_COMPACT_VPT / 4; ++k) {
            const uint4 x = q[k];
            c[4 * k] = ((bits >> (4 * k)) & 1ull) ? x.x : 0;
            c[4 * k + 1] = ((bits >> (4 * k + 1)) & 1ull) ? x.y : 0;
            c[4 * k + 2] = ((bits >> (4 * k + 2)) & 1ull) ? x.z : 0;
            c[4 * k + 3] = ((bits >> (4 * k + 3)) & 1ull) ? x.w : 0;
        }
        return;
    }
    const uint64_t* __restrict__ offsets = (const uint64_t*)src;
    uint64_t o[GZ_COMPACT_VPT + 1];
#pragma unroll
    for (int k = 0; k <= GZ_COMPACT_VPT; ++k) o[k] = offsets[(v0 + k < num_vertices) ? v0 + k : num_vertices];
#pragma unroll
    for (int k = 0; k < GZ_COMPACT_VPT; ++k) c[k] = ((bits >> k) & 1ull) ? o[k + 1] - o[k] : 0;
}

/* counts[v] = offsets[v + 1] - offsets[v] for v < num_vertices, 0 up to
 * `padded` (a multiple of GZ_COMPACT_VPT). */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_vector_counts(const uint64_t* __restrict__ offsets,
                                                               uint64_t num_vertices, uint64_t padded,
                                                               uint32_t* __restrict__ counts)
{
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < padded;
         v += (uint64_t)gridDim.x * blockDim.x)
        counts[v] = (v < num_vertices) ? (uint32_t)(offsets[v + 1] - offsets[v]) : 0u;
}

template <typename T>
__device__ __forceinline__ T gz_block_exclusive_scan(T x, T* lds, T& block_total)
{
    /* wave inclusive scan */
    const int lane = gz_lane();
    T incl = x;
#pragma unroll
    for (int o = 1; o < GZ_WAVE; o <<= 1) {
        const T t = __shfl_up(incl, o, GZ_WAVE);
        if (lane >= o) incl += t;
    }
    const int w = threadIdx.x / GZ_WAVE;
    if (lane == GZ_WAVE - 1) lds[w] = incl;
    __syncthreads();
    T wave_off = 0, total = 0;
    for (int i = 0; i < GZ_WAVES_PER_BLOCK; ++i) {
        if (i < w) wave_off += lds[i];
        total += lds[i];
    }
    __syncthreads();
    block_total = total;
    return wave_off + incl - x;
}

/* ---- the scan-free form of the passes -------------------------------------
 * The count pass also adds each block's totals to the counters of its
 * segment: GZ_SEGMENTS ranges of seg_blocks consecutive blocks, whose
 * counters sit on separate 256-byte lines so that their atomics proceed in
 * parallel (4096 workgroup atomics on one word take about 50 us on MI355X;
 * spread over 64 lines, no longer than the launch).  The write pass then finds
 * its block's position from the counters of the segments before it and the
 * block sums of its own segment before it, which replaces the scan kernel.
 * The first block of the write pass also writes the totals and the prefix
 * sentinel, and zeroes the other bank of counters, which the next compaction
 * counts into (the compaction before this one read that bank, and its write
 * pass has finished: launches on one stream run in order). */
#define GZ_SEGMENTS 64
#define GZ_SEG_STRIDE 32  /* words between counters */
#define GZ_SEG_VECTORS 16 /* a segment's vector counter follows its entry counter, 128 bytes on */

__device__ __forceinline__ void gz_seg_add(unsigned long long* seg, uint64_t s, unsigned long long entries,
                                           unsigned long long vectors)
{
    atomicAdd(&seg[s * GZ_SEG_STRIDE], entries);
    if (vectors != 0) atomicAdd(&seg[s * GZ_SEG_STRIDE + GZ_SEG_VECTORS], vectors);
}

/* Entries and vectors of the blocks before block b.  Wavefront-collective. */
__device__ __forceinline__ void gz_seg_base(const unsigned long long* __restrict__ seg,
                                            const uint64_t* __restrict__ block_sums, uint64_t b, uint64_t seg_blocks,
                                            uint64_t& entries, uint64_t& vectors)
{
    const int lane = gz_lane();
    const uint64_t s = b / seg_blocks;
    uint64_t e = 0, v = 0;
    if ((uint64_t)lane < s) {
        e = seg[(uint64_t)lane * GZ_SEG_STRIDE];
        v = seg[(uint64_t)lane * GZ_SEG_STRIDE + GZ_SEG_VECTORS];
    }
    for (uint64_t q = s * seg_blocks + (uint64_t)lane; q < b; q += GZ_WAVE) {
        e += block_sums[2 * q];
        v += block_sums[2 * q + 1];
    }
    entries = gz_wave_sum_u64(e);
    vectors = gz_wave_sum_u64(v);
}

/* First block of a write pass: totals, the sentinel prefix[entries] =
 * vectors, and the other bank cleared.  Run by wavefront 0 of block 0. */
__device__ __forceinline__ void gz_seg_finish(const unsigned long long* __restrict__ seg,
                                              unsigned long long* __restrict__ other, uint64_t* __restrict__ totals,
                                              uint64_t* __restrict__ prefix)
{
    const int lane = gz_lane();
    const uint64_t e = gz_wave_sum_u64(seg[(uint64_t)lane * GZ_SEG_STRIDE]);
    const uint64_t v = gz_wave_sum_u64(seg[(uint64_t)lane * GZ_SEG_STRIDE + GZ_SEG_VECTORS]);
    other[(uint64_t)lane * GZ_SEG_STRIDE] = 0;
    other[(uint64_t)lane * GZ_SEG_STRIDE + GZ_SEG_VECTORS] = 0;
    if (lane == 0) {
        totals[0] = e;
        totals[1] = v;
        if (prefix != nullptr) prefix[e] = v;
    }
}

/* Pass 1: per block of GZ_COMPACT_VPB vertices, its active vertices and
 * their vectors.  Workgroups stride over the `nblocks` vertex blocks. */
template <bool kNarrow>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_frontier_count(const uint64_t* __restrict__ bitmap,
                                                                 const void* __restrict__ src,
                                                                 uint64_t num_vertices,
                                                                 uint64_t* __restrict__ block_sums,
                                                                 uint64_t* __restrict__ stamp, uint64_t nblocks,
                                                                 gz_guard guard, unsigned long long* seg,
                                                                 uint64_t seg_blocks)
{
    __shared__ unsigned long long lds[GZ_WAVES_PER_BLOCK];
    if (gz_guard_off(guard)) return;
    if (stamp != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *stamp = (uint64_t)wall_clock64();
    for (uint64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const uint64_t v0 = b * GZ_COMPACT_VPB + GZ_COMPACT_VPT * (uint64_t)threadIdx.x;
        uint64_t c[GZ_COMPACT_VPT];
        gz_run_counts<kNarrow>(bitmap, src, num_vertices, v0, c);
        unsigned long long nv = 0, nvec = 0;
#pragma unroll
        for (int k = 0; k < GZ_COMPACT_VPT; ++k) {
            nv += c[k] ? 1 : 0;
            nvec += c[k];
        }
        unsigned long long tv, tvec;
        gz_block_exclusive_scan<unsigned long long>(nv, lds, tv);
        gz_block_exclusive_scan<unsigned long long>(nvec, lds, tvec);
        if (threadIdx.x == 0) {
            block_sums[2 * b] = tv;
            block_sums[2 * b + 1] = tvec;
            if (seg != nullptr && tv != 0) gz_seg_add(seg, b / seg_blocks, tv, tvec);
        }
    }
}

/* Single block: exclusive scan of the per-block pairs, totals to totals[0..1],
 * and the sentinel prefix[total vertices] = total vectors.  A pass stages up
 * to GZ_SCAN_PASS pairs in LDS with coalesced loads; each thread then scans
 * GZ_SCAN_PASS / 1024 consecutive pairs serially around one block scan of
 * the thread sums, and the pass is written back coalesced. */
#define GZ_SCAN_PASS 4096
__global__ __launch_bounds__(1024) void gz_k_frontier_scan(uint64_t* __restrict__ block_sums, uint64_t blocks,
                                                           uint64_t* __restrict__ totals, uint64_t* __restrict__ prefix,
                                                           gz_guard guard)
{
    __shared__ unsigned long long stage[2 * GZ_SCAN_PASS];
    __shared__ unsigned long long lds_v[2][1024 / GZ_WAVE];
    __shared__ unsigned long long carry[2];
    if (gz_guard_off(guard)) return;
    const int per = GZ_SCAN_PASS / 1024;
    if (threadIdx.x == 0) {
        carry[0] = 0;
        carry[1] = 0;
    }
    const int lane = gz_lane();
    const int w = threadIdx.x / GZ_WAVE;
    const int nw = blockDim.x / GZ_WAVE;
    for (uint64_t base = 0; base < blocks; base += GZ_SCAN_PASS) {
        const uint64_t n = (blocks - base < GZ_SCAN_PASS) ? blocks - base : GZ_SCAN_PASS;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 2 * GZ_SCAN_PASS / 1024; ++r) {
            const uint64_t i = (uint64_t)r * 1024 + threadIdx.x;
            stage[i] = (i < 2 * n) ? block_sums[2 * base + i] : 0ull;
        }
        __syncthreads();
        unsigned long long x[2] = {0, 0};
        for (int k = 0; k < per; ++k) {
            x[0] += stage[2 * (threadIdx.x * per + k)];
            x[1] += stage[2 * (threadIdx.x * per + k) + 1];
        }
        unsigned long long incl[2] = {x[0], x[1]};
#pragma unroll
        for (int o = 1; o < GZ_WAVE; o <<= 1) {
            const unsigned long long t0 = __shfl_up(incl[0], o, GZ_WAVE);
            const unsigned long long t1 = __shfl_up(incl[1], o, GZ_WAVE);
            if (lane >= o) {
                incl[0] += t0;
                incl[1] += t1;
            }
        }
        if (lane == GZ_WAVE - 1) {
            lds_v[0][w] = incl[0];
            lds_v[1][w] = incl[1];
        }
        __syncthreads();
        unsigned long long off[2] = {carry[0], carry[1]}, tot[2] = {0, 0};
        for (int k = 0; k < nw; ++k) {
            if (k < w) {
                off[0] += lds_v[0][k];
                off[1] += lds_v[1][k];
            }
            tot[0] += lds_v[0][k];
            tot[1] += lds_v[1][k];
        }
        off[0] += incl[0] - x[0];
        off[1] += incl[1] - x[1];
        for (int k = 0; k < per; ++k) {
            const int i = 2 * (threadIdx.x * per + k);
            const unsigned long long y0 = stage[i], y1 = stage[i + 1];
            stage[i] = off[0];
            stage[i + 1] = off[1];
            off[0] += y0;
            off[1] += y1;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 2 * GZ_SCAN_PASS / 1024; ++r) {
            const uint64_t i = (uint64_t)r * 1024 + threadIdx.x;
            if (i < 2 * n) block_sums[2 * base + i] = stage[i];
        }
        if (threadIdx.x == 0) {
            carry[0] += tot[0];
            carry[1] += tot[1];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        totals[0] = carry[0];
        totals[1] = carry[1];
        prefix[carry[0]] = carry[1];
    }
}

/* Pass 3: each block's active vertices and their vector prefix, in order. */
template <bool kNarrow>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_frontier_write(const uint64_t* __restrict__ bitmap,
                                                                 const void* __restrict__ src,
                                                                 uint64_t num_vertices,
                                                                 const uint64_t* __restrict__ block_sums,
                                                                 uint64_t* __restrict__ list,
                                                                 uint64_t* __restrict__ prefix, uint64_t nblocks,
                                                                 gz_guard guard, const unsigned long long* seg,
                                                                 unsigned long long* seg_other, uint64_t seg_blocks,
                                                                 uint64_t* __restrict__ totals)
{
    __shared__ unsigned long long lds[GZ_WAVES_PER_BLOCK];
    /* the block's entries in list order (vertex as an offset from the block's
     * first, vector prefix as an offset from the block's first vector), so
     * that list and prefix are written by consecutive lanes */
    __shared__ uint16_t stage_v[GZ_COMPACT_VPB];
    __shared__ uint32_t stage_p[GZ_COMPACT_VPB];
    static_assert(GZ_COMPACT_VPB <= 65536, "block offsets are 16-bit");
    if (gz_guard_off(guard)) return;
    if (seg != nullptr && blockIdx.x == 0 && threadIdx.x < GZ_WAVE) gz_seg_finish(seg, seg_other, totals, prefix);
    for (uint64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const uint64_t vb = b * GZ_COMPACT_VPB;
        const uint32_t t0 = GZ_COMPACT_VPT * (uint32_t)threadIdx.x;
        const uint64_t v0 = vb + t0;
        uint64_t c[GZ_COMPACT_VPT];
        gz_run_counts<kNarrow>(bitmap, src, num_vertices, v0, c);
        unsigned long long nv = 0, nvec = 0;
#pragma unroll
        for (int k = 0; k < GZ_COMPACT_VPT; ++k) {
            nv += c[k] ? 1 : 0;
            nvec += c[k];
        }
        uint64_t pos0, vpos0;
        if (seg != nullptr)
            gz_seg_base(seg, block_sums, b, seg_blocks, pos0, vpos0);
        else {
            pos0 = block_sums[2 * b];
            vpos0 = block_sums[2 * b + 1];
        }
        unsigned long long tv, tvec;
        /* each scan ends with a barrier, after every thread's copy-out of the
         * previous block, so the stage is free again */
        const uint64_t p0 = gz_block_exclusive_scan<unsigned long long>(nv, lds, tv);
        const uint64_t q0 = gz_block_exclusive_scan<unsigned long long>(nvec, lds, tvec);
        if (tvec < (1ull << 32)) { /* block-uniform */
            uint32_t p = (uint32_t)p0, q = (uint32_t)q0;
#pragma unroll
            for (int k = 0; k < GZ_COMPACT_VPT; ++k) {
                if (c[k]) {
                    stage_v[p] = (uint16_t)(t0 + (uint32_t)k);
                    stage_p[p] = q;
                    p += 1;
                    q += (uint32_t)c[k];
                }
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < (uint32_t)tv; i += GZ_BLOCK) {
                list[pos0 + i] = vb + stage_v[i];
                prefix[pos0 + i] = vpos0 + stage_p[i];
            }
        } else {
            uint64_t pos = pos0 + p0, vpos = vpos0 + q0;
#pragma unroll
            for (int k = 0; k < GZ_COMPACT_VPT; ++k) {
                if (c[k]) {
                    list[pos] = v0 + (uint64_t)k;
                    prefix[pos] = vpos;
                    pos += 1;
                    vpos += c[k];
                }
            }
        }
    }
}

/* The vertex list alone (no vector prefix): vertices whose bit is set in
 * `bitmap` and in `hasvec` (the vertices that own vectors), in order. */
__device__ __forceinline__ uint32_t gz_list_bits(const uint64_t* __restrict__ bitmap,
                                                 const uint64_t* __restrict__ hasvec, uint64_t num_vertices,
                                                 uint64_t v0)
{
    if (v0 >= num_vertices) return 0u;
    const uint64_t mask = (1ull << GZ_COMPACT_VPT) - 1ull;
    uint64_t bits = ((bitmap[v0 >> 6] & hasvec[v0 >> 6]) >> (v0 & 63)) & mask;
    const uint64_t left = num_vertices - v0;
    if (left < GZ_COMPACT_VPT) bits &= (1ull << left) - 1ull;
    return (uint32_t)bits;
}

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_frontier_list_count(const uint64_t* __restrict__ bitmap,
                                                                     const uint64_t* __restrict__ hasvec,
                                                                     uint64_t num_vertices,
                                                                     uint64_t* __restrict__ block_sums,
                                                                     uint64_t* __restrict__ stamp, uint64_t nblocks,
                                                                     unsigned long long* seg, uint64_t seg_blocks)
{
    __shared__ unsigned long long lds[GZ_WAVES_PER_BLOCK];
    if (stamp != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *stamp = (uint64_t)wall_clock64();
    for (uint64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const uint64_t v0 = b * GZ_COMPACT_VPB + GZ_COMPACT_VPT * (uint64_t)threadIdx.x;
        const unsigned long long nv = (unsigned long long)__popc(gz_list_bits(bitmap, hasvec, num_vertices, v0));
        unsigned long long tv;
        gz_block_exclusive_scan<unsigned long long>(nv, lds, tv);
        if (threadIdx.x == 0) {
            block_sums[2 * b] = tv;
            block_sums[2 * b + 1] = 0;
            if (seg != nullptr && tv != 0) gz_seg_add(seg, b / seg_blocks, tv, 0);
        }
    }
}

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_frontier_list_write(const uint64_t* __restrict__ bitmap,
                                                                     const uint64_t* __restrict__ hasvec,
                                                                     uint64_t num_vertices,
                                                                     const uint64_t* __restrict__ block_sums,
                                                                     uint64_t* __restrict__ list, uint64_t nblocks,
                                                                     const unsigned long long* seg,
                                                                     unsigned long long* seg_other, uint64_t seg_blocks,
                                                                     uint64_t* __restrict__ totals,
                                                                     uint64_t* __restrict__ prefix)
{
    __shared__ unsigned long long lds[GZ_WAVES_PER_BLOCK];
    /* the block's ids as offsets from its first vertex, in list order, so
     * that the list is written by consecutive lanes (a thread writing its own
     * ids would store 64 lanes up to 16 entries apart) */
    __shared__ uint16_t stage[GZ_COMPACT_VPB];
    static_assert(GZ_COMPACT_VPB <= 65536, "block offsets are 16-bit");
    if (seg != nullptr && blockIdx.x == 0 && threadIdx.x < GZ_WAVE) gz_seg_finish(seg, seg_other, totals, prefix);
    for (uint64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const uint64_t vb = b * GZ_COMPACT_VPB;
        const uint32_t t0 = GZ_COMPACT_VPT * (uint32_t)threadIdx.x;
        uint32_t bits = gz_list_bits(bitmap, hasvec, num_vertices, vb + t0);
        uint64_t pos0, unused;
        if (seg != nullptr)
            gz_seg_base(seg, block_sums, b, seg_blocks, pos0, unused);
        else
            pos0 = block_sums[2 * b];
        unsigned long long tv;
        /* the scan ends with a barrier, after every thread's copy-out of the
         * previous block, so the stage is free again */
        uint32_t p = (uint32_t)gz_block_exclusive_scan<unsigned long long>((unsigned long long)__popc(bits), lds, tv);
        while (bits) {
            const int k = __ffs(bits) - 1;
            bits &= bits - 1u;
            stage[p++] = (uint16_t)(t0 + (uint32_t)k);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < (uint32_t)tv; i += GZ_BLOCK) list[pos0 + i] = vb + stage[i];
    }
}

/* hasvec bit v: vertex v owns at least one vector of the list. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_has_vectors(const uint64_t* __restrict__ offsets,
                                                             uint64_t num_vertices, uint64_t words,
                                                             uint64_t* __restrict__ hasvec)
{
    const int lane = gz_lane();
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / GZ_WAVE; w < words;
         w += (uint64_t)gridDim.x * blockDim.x / GZ_WAVE) {
        const uint64_t v = w * 64 + (uint64_t)lane;
        const bool has = v < num_vertices && offsets[v + 1] > offsets[v];
        const uint64_t word = __ballot(has);
        if (lane == 0) hasvec[w] = word;
    }
}

static int gz_compact_narrow = 1;

extern "C" void gz_frontier_compact_narrow(int on) { gz_compact_narrow = on ? 1 : 0; }

/* 32-bit vector counts of the two lists' vertices, built on first use for
 * an offsets array, vertex count and list length, and kept until the graph
 * is released or another is published (gz_graph_release and every load
 * path's finish call gz_frontier_counts_release). */
static struct {
    const uint64_t* offsets;
    uint64_t num_vertices, vector_count;
    uint32_t* counts;
} gz_counts_cache[2];
static int gz_counts_next;

/* Bitmaps of the vertices that own vectors, per offsets array, kept as the
 * counts are. */
static struct {
    const uint64_t* offsets;
    uint64_t num_vertices, vector_count;
    uint64_t* bits;
} gz_hasvec_cache[2];
static int gz_hasvec_next;

extern "C" void gz_frontier_counts_release(void)
{
    for (auto& e : gz_counts_cache) {
        gz_device_free(e.counts);
        e = {nullptr, 0, 0, nullptr};
    }
    gz_counts_next = 0;
    for (auto& e : gz_hasvec_cache) {
        gz_device_free(e.bits);
        e = {nullptr, 0, 0, nullptr};
    }
    gz_hasvec_next = 0;
}

static const uint64_t* gz_frontier_hasvec(const uint64_t* offsets, uint64_t num_vertices, uint64_t vector_count)
{
    for (auto& e : gz_hasvec_cache)
        if (e.bits != nullptr && e.offsets == offsets && e.num_vertices == num_vertices &&
            e.vector_count == vector_count)
            return e.bits;
    auto& e = gz_hasvec_cache[gz_hasvec_next];
    gz_hasvec_next ^= 1;
    gz_device_free(e.bits);
    const uint64_t words = (num_vertices + 63) / 64;
    e.bits = (uint64_t*)gz_device_malloc((words + 8) * sizeof(uint64_t));
    uint64_t blocks = (words * GZ_WAVE + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gz_k_has_vectors, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, offsets, num_vertices,
                       words, e.bits);
    GZ_LAUNCH_CHECK();
    e.offsets = offsets;
    e.num_vertices = num_vertices;
    e.vector_count = vector_count;
    return e.bits;
}

/* The counts of a list if they are built, else nullptr (allocates nothing). */
static const uint32_t* gz_frontier_counts_cached(const uint64_t* offsets, uint64_t num_vertices, uint64_t vector_count)
{
    for (auto& e : gz_counts_cache)
        if (e.counts != nullptr && e.offsets == offsets && e.num_vertices == num_vertices &&
            e.vector_count == vector_count)
            return e.counts;
    return nullptr;
}

static const uint32_t* gz_frontier_counts(const uint64_t* offsets, uint64_t num_vertices, uint64_t vector_count)
{
    for (auto& e : gz_counts_cache)
        if (e.counts != nullptr && e.offsets == offsets && e.num_vertices == num_vertices &&
            e.vector_count == vector_count)
            return e.counts;
    auto& e = gz_counts_cache[gz_counts_next];
    gz_counts_next ^= 1;
    gz_device_free(e.counts);
    const uint64_t padded = (num_vertices + GZ_COMPACT_VPT - 1) / GZ_COMPACT_VPT * GZ_COMPACT_VPT;
    e.counts = (uint32_t*)gz_device_malloc(padded * sizeof(uint32_t));
    uint64_t blocks = (padded + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gz_k_vector_counts, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, offsets,
                       num_vertices, padded, e.counts);
    GZ_LAUNCH_CHECK();
    e.offsets = offsets;
    e.num_vertices = num_vertices;
    e.vector_count = vector_count;
    return e.counts;
}

/* Two banks of segment counters for the scan-free passes. */
static unsigned long long* gz_seg_banks = nullptr;
static int gz_seg_parity = 0;

/* The bank the next scan-free compaction counts into and the one its write
 * pass clears, then the parity advances.  False while the stream is being
 * captured: a graph would replay one parity every time, so captured
 * compactions keep the scan kernel. */
static bool gz_seg_take(unsigned long long** bank, unsigned long long** other)
{
    hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
    GZ_CHECK(hipStreamIsCapturing(gz_stream, &status));
    if (status != hipStreamCaptureStatusNone) return false;
    const size_t words = (size_t)GZ_SEGMENTS * GZ_SEG_STRIDE;
    if (gz_seg_banks == nullptr) {
        gz_seg_banks = (unsigned long long*)gz_device_malloc(2 * words * sizeof(uint64_t));
        gz_device_zero(gz_seg_banks, 2 * words * sizeof(uint64_t));
    }
    *bank = gz_seg_banks + (size_t)gz_seg_parity * words;
    *other = gz_seg_banks + (size_t)(gz_seg_parity ^ 1) * words;
    gz_seg_parity ^= 1;
    return true;
}

static uint64_t gz_seg_blocks(uint64_t nblocks) { return (nblocks + GZ_SEGMENTS - 1) / GZ_SEGMENTS; }

static void gz_frontier_compact_impl(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices,
                                     uint64_t vector_count, gz_guard guard, uint64_t max_grid, bool allow_build,
                                     const gz_compact_out& out)
{
    const uint64_t blocks = gz_compact_blocks(num_vertices);
    if (blocks == 0) {
        gz_device_zero(out.totals, 2 * sizeof(uint64_t));
        gz_device_zero(out.prefix, sizeof(uint64_t));
        return;
    }
    const uint64_t grid = blocks < max_grid ? blocks : max_grid;
    /* below 2^32 vectors in the list every vertex's count fits 32 bits */
    bool narrow = gz_compact_narrow && vector_count < (1ull << 32);
    const void* src = (const void*)offsets;
    if (narrow) {
        src = allow_build ? (const void*)gz_frontier_counts(offsets, num_vertices, vector_count)
                          : (const void*)gz_frontier_counts_cached(offsets, num_vertices, vector_count);
        if (src == nullptr) {
            narrow = false;
            src = (const void*)offsets;
        }
    }
    uint64_t* stamp = guard.word == nullptr ? gz_scratch.stamp_target : nullptr;
    if (guard.word == nullptr) gz_scratch.stamp_target = nullptr;
    /* guarded passes are captured into the device loop's graph */
    unsigned long long *seg = nullptr, *other = nullptr;
    if (guard.word == nullptr && !gz_seg_take(&seg, &other)) seg = other = nullptr;
    const uint64_t seg_blocks = gz_seg_blocks(blocks);
    gz_ktimer_begin(GZ_K_COMPACT);
    if (narrow)
        hipLaunchKernelGGL(gz_k_frontier_count<true>, dim3((unsigned)grid), dim3(GZ_BLOCK), 0, gz_stream, bitmap, src,
                           num_vertices, out.block_sums, stamp, blocks, guard, seg, seg_blocks);
    else
        hipLaunchKernelGGL(gz_k_frontier_count<false>, dim3((unsigned)grid), dim3(GZ_BLOCK), 0, gz_stream, bitmap,
                           src, num_vertices, out.block_sums, stamp, blocks, guard, seg, seg_blocks);
    GZ_LAUNCH_CHECK();
    if (seg == nullptr) {
        hipLaunchKernelGGL(gz_k_frontier_scan, dim3(1), dim3(1024), 0, gz_stream, out.block_sums, blocks, out.totals,
                           out.prefix, guard);
        GZ_LAUNCH_CHECK();
    }
    if (narrow)
        hipLaunchKernelGGL(gz_k_frontier_write<true>, dim3((unsigned)grid), dim3(GZ_BLOCK), 0, gz_stream, bitmap, src,
                           num_vertices, out.block_sums, out.list, out.prefix, blocks, guard, seg, other, seg_blocks,
                           out.totals);
    else
        hipLaunchKernelGGL(gz_k_frontier_write<false>, dim3((unsigned)grid), dim3(GZ_BLOCK), 0, gz_stream, bitmap,
                           src, num_vertices, out.block_sums, out.list, out.prefix, blocks, guard, seg, other,
                           seg_blocks, out.totals);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_COMPACT);
}

static gz_compact_out gz_compact_scratch_out(void)
{
    return gz_compact_out{gz_scratch.active_vertices, gz_scratch.active_prefix, gz_scratch.active_block_sums,
                          gz_scratch.active_totals};
}

void gz_frontier_compact(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices,
                         uint64_t vector_count)
{
    gz_frontier_compact_impl(bitmap, offsets, num_vertices, vector_count, gz_guard{nullptr, 0}, ~0ull, true,
                             gz_compact_scratch_out());
}

void gz_frontier_list(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices, uint64_t vector_count)
{
    const uint64_t blocks = gz_compact_blocks(num_vertices);
    if (blocks == 0) {
        gz_device_zero(gz_scratch.active_totals, 2 * sizeof(uint64_t));
        return;
    }
    const uint64_t* hasvec = gz_frontier_hasvec(offsets, num_vertices, vector_count);
    uint64_t* stamp = gz_scratch.stamp_target;
    gz_scratch.stamp_target = nullptr;
    unsigned long long *seg = nullptr, *other = nullptr;
    if (!gz_seg_take(&seg, &other)) seg = other = nullptr;
    const uint64_t seg_blocks = gz_seg_blocks(blocks);
    gz_ktimer_begin(GZ_K_COMPACT);
    hipLaunchKernelGGL(gz_k_frontier_list_count, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, bitmap, hasvec,
                       num_vertices, gz_scratch.active_block_sums, stamp, blocks, seg, seg_blocks);
    GZ_LAUNCH_CHECK();
    if (seg == nullptr) {
        hipLaunchKernelGGL(gz_k_frontier_scan, dim3(1), dim3(1024), 0, gz_stream, gz_scratch.active_block_sums, blocks,
                           gz_scratch.active_totals, gz_scratch.active_prefix, gz_guard{nullptr, 0});
        GZ_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(gz_k_frontier_list_write, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, bitmap, hasvec,
                       num_vertices, gz_scratch.active_block_sums, gz_scratch.active_vertices, blocks, seg, other,
                       seg_blocks, gz_scratch.active_totals, gz_scratch.active_prefix);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_COMPACT);
}

void gz_frontier_compact_to(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices,
                            uint64_t vector_count, const gz_compact_out& out)
{
    gz_frontier_compact_impl(bitmap, offsets, num_vertices, vector_count, gz_guard{nullptr, 0}, ~0ull, true, out);
}

void gz_frontier_compact_guarded(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices,
                                 uint64_t vector_count, gz_guard guard, uint64_t max_grid)
{
    gz_frontier_compact_impl(bitmap, offsets, num_vertices, vector_count, guard, max_grid, false,
                             gz_compact_scratch_out());
}
