/*
 * gz_frontier.hip -- compaction of a frontier bitmap into the active-vertex
 * list consumed by the frontier-driven edge kernels (see gz_frontier.h).
 * Three passes: per-block counts, a single-block exclusive scan of the block
 * counts, and a write pass that re-derives each thread's offsets.
 */
#include "gz_frontier.h"

/* Each thread owns one 64-vertex bitmap word. */
__device__ __forceinline__ void gz_word_counts(const uint64_t* __restrict__ bitmap, const uint64_t* __restrict__ offsets,
                                               uint64_t num_vertices, uint64_t w, uint64_t words, uint64_t& nv,
                                               uint64_t& nvec)
{
    nv = 0;
    nvec = 0;
    if (w >= words) return;
    uint64_t bits = bitmap[w];
    const uint64_t base = w * 64;
    if (base + 64 > num_vertices) {
        const uint64_t keep = num_vertices - base;
        bits &= (keep >= 64) ? ~0ull : ((1ull << keep) - 1ull);
    }
    while (bits) {
        const int b = __ffsll((unsigned long long)bits) - 1;
        bits &= bits - 1;
        const uint64_t v = base + (uint64_t)b;
        const uint64_t c = offsets[v + 1] - offsets[v];
        if (c) {
            nv += 1;
            nvec += c;
        }
    }
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

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_frontier_count(const uint64_t* __restrict__ bitmap,
                                                                 const uint64_t* __restrict__ offsets,
                                                                 uint64_t num_vertices, uint64_t words,
                                                                 uint64_t* __restrict__ block_sums)
{
    __shared__ unsigned long long lds[GZ_WAVES_PER_BLOCK];
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t nv, nvec;
    gz_word_counts(bitmap, offsets, num_vertices, w, words, nv, nvec);
    unsigned long long tv, tvec;
    gz_block_exclusive_scan<unsigned long long>(nv, lds, tv);
    gz_block_exclusive_scan<unsigned long long>(nvec, lds, tvec);
    if (threadIdx.x == 0) {
        block_sums[2 * blockIdx.x] = tv;
        block_sums[2 * blockIdx.x + 1] = tvec;
    }
}

/* Single block: exclusive scan of the per-block pairs, totals to totals[0..1],
 * and the sentinel prefix[total vertices] = total vectors. */
__global__ __launch_bounds__(1024) void gz_k_frontier_scan(uint64_t* __restrict__ block_sums, uint64_t blocks,
                                                           uint64_t* __restrict__ totals, uint64_t* __restrict__ prefix)
{
    __shared__ unsigned long long lds_v[1024 / GZ_WAVE];
    __shared__ unsigned long long carry[2];
    if (threadIdx.x == 0) {
        carry[0] = 0;
        carry[1] = 0;
    }
    __syncthreads();
    const int lane = gz_lane();
    const int w = threadIdx.x / GZ_WAVE;
    const int nw = blockDim.x / GZ_WAVE;
    for (uint64_t base = 0; base < blocks; base += blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        for (int field = 0; field < 2; ++field) {
            unsigned long long x = (i < blocks) ? block_sums[2 * i + field] : 0ull;
            unsigned long long incl = x;
#pragma unroll
            for (int o = 1; o < GZ_WAVE; o <<= 1) {
                const unsigned long long t = __shfl_up(incl, o, GZ_WAVE);
                if (lane >= o) incl += t;
            }
            if (lane == GZ_WAVE - 1) lds_v[w] = incl;
            __syncthreads();
            unsigned long long off = carry[field], tot = 0;
            for (int k = 0; k < nw; ++k) {
                if (k < w) off += lds_v[k];
                tot += lds_v[k];
            }
            if (i < blocks) block_sums[2 * i + field] = off + incl - x;
            __syncthreads();
            if (threadIdx.x == 0) carry[field] += tot;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        totals[0] = carry[0];
        totals[1] = carry[1];
        prefix[carry[0]] = carry[1];
    }
}

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_frontier_write(const uint64_t* __restrict__ bitmap,
                                                                 const uint64_t* __restrict__ offsets,
                                                                 uint64_t num_vertices, uint64_t words,
                                                                 const uint64_t* __restrict__ block_sums,
                                                                 uint64_t* __restrict__ list,
                                                                 uint64_t* __restrict__ prefix)
{
    __shared__ unsigned long long lds[GZ_WAVES_PER_BLOCK];
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t nv, nvec;
    gz_word_counts(bitmap, offsets, num_vertices, w, words, nv, nvec);
    unsigned long long tv, tvec;
    uint64_t pos = block_sums[2 * blockIdx.x] + gz_block_exclusive_scan<unsigned long long>(nv, lds, tv);
    uint64_t vpos = block_sums[2 * blockIdx.x + 1] + gz_block_exclusive_scan<unsigned long long>(nvec, lds, tvec);
    if (w >= words || nv == 0) return;
    uint64_t bits = bitmap[w];
    const uint64_t base = w * 64;
    if (base + 64 > num_vertices) {
        const uint64_t keep = num_vertices - base;
        bits &= (keep >= 64) ? ~0ull : ((1ull << keep) - 1ull);
    }
    while (bits) {
        const int b = __ffsll((unsigned long long)bits) - 1;
        bits &= bits - 1;
        const uint64_t v = base + (uint64_t)b;
        const uint64_t c = offsets[v + 1] - offsets[v];
        if (c) {
            list[pos] = v;
            prefix[pos] = vpos;
            pos += 1;
            vpos += c;
        }
    }
}

void gz_frontier_compact(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices)
{
    const uint64_t words = (num_vertices + 63) / 64;
    const uint64_t blocks = (words + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks == 0) {
        gz_device_zero(gz_scratch.active_totals, 2 * sizeof(uint64_t));
        gz_device_zero(gz_scratch.active_prefix, sizeof(uint64_t));
        return;
    }
    gz_ktimer_begin(GZ_K_COMPACT);
    hipLaunchKernelGGL(gz_k_frontier_count, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, bitmap, offsets,
                       num_vertices, words, gz_scratch.active_block_sums);
    GZ_LAUNCH_CHECK();
    hipLaunchKernelGGL(gz_k_frontier_scan, dim3(1), dim3(1024), 0, gz_stream, gz_scratch.active_block_sums, blocks,
                       gz_scratch.active_totals, gz_scratch.active_prefix);
    GZ_LAUNCH_CHECK();
    hipLaunchKernelGGL(gz_k_frontier_write, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, bitmap, offsets,
                       num_vertices, words, gz_scratch.active_block_sums, gz_scratch.active_vertices,
                       gz_scratch.active_prefix);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_COMPACT);
}
