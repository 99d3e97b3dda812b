/*
 * gz_init.hip -- device-side initialisation of vertex arrays and frontiers
 * (graph_helper_initialize_vertex_info, graphdata.c:279-287, and
 * graph_helper_create_and_initialize_frontiers, graphdata.c:290-319, with the
 * per-algorithm callbacks execution_pr.c:64-88, execution_cc.c:234-258,
 * execution_bfs.c:70-116).  Filling on the device keeps repeated runs (many
 * BFS roots) off PCIe.
 */
#include "gz_internal.h"
#include "gz_host_internal.h"

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_init_pr(double* __restrict__ props, double* __restrict__ acc,
                                                         const double* __restrict__ outdeg, uint64_t n)
{
    const double nv = (double)n;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n + 8; v += (uint64_t)gridDim.x * blockDim.x) {
        if (v < n) {
            const double od = outdeg[v];
            props[v] = (1.0 / nv) / (0.0 == od ? nv : od);
        } else {
            props[v] = 0.0;
        }
        acc[v] = 0.0;
    }
}

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_fill(uint64_t* __restrict__ p, uint64_t value, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_iota(uint64_t* __restrict__ p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = i;
}

__global__ void gz_k_store(uint64_t* __restrict__ p, uint64_t index, uint64_t value) { p[index] = value; }

/* The start of a BFS from `root` in one launch (execution_bfs.c:106-116 and
 * :70-102): every word unvisited but the root's (no parent), HasInfo* clear,
 * HasInfo = {root}, WantsInfo = every vertex but the root. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_bfs_reset(uint64_t* __restrict__ words, uint64_t n,
                                                           uint64_t* __restrict__ acc, uint64_t* __restrict__ has,
                                                           uint64_t* __restrict__ wants, uint64_t fwords,
                                                           uint64_t root)
{
    const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = first; i < n; i += stride) words[i] = (i == root) ? GZ_BFS_PARENT_NONE : GZ_BFS_UNVISITED;
    const uint64_t rw = root >> 6, rbit = 1ull << (root & 63);
    for (uint64_t w = first; w < fwords; w += stride) {
        acc[w] = 0;
        has[w] = (w == rw) ? rbit : 0ull;
        wants[w] = (w == rw) ? ~rbit : ~0ull;
    }
}

static unsigned gz_blocks_for(uint64_t n)
{
    uint64_t b = (n + GZ_BLOCK - 1) / GZ_BLOCK;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (unsigned)b;
}

void gz_dev_bfs_reset(uint64_t* words, uint64_t n, uint64_t* acc, uint64_t* has, uint64_t* wants, uint64_t fwords,
                      uint64_t root)
{
    hipLaunchKernelGGL(gz_k_bfs_reset, dim3(gz_blocks_for(n > fwords ? n : fwords)), dim3(GZ_BLOCK), 0, gz_stream, words,
                       n, acc, has, wants, fwords, root);
    GZ_LAUNCH_CHECK();
}

extern "C" void gz_dev_init_pr(double* props, double* acc, const double* outdeg, uint64_t n)
{
    hipLaunchKernelGGL(gz_k_init_pr, dim3(gz_blocks_for(n + 8)), dim3(GZ_BLOCK), 0, gz_stream, props, acc, outdeg, n);
    GZ_LAUNCH_CHECK();
}

extern "C" void gz_dev_fill_u64(uint64_t* p, uint64_t value, uint64_t n)
{
    if (n == 0) return;
    hipLaunchKernelGGL(gz_k_fill, dim3(gz_blocks_for(n)), dim3(GZ_BLOCK), 0, gz_stream, p, value, n);
    GZ_LAUNCH_CHECK();
}

extern "C" void gz_dev_iota_u64(uint64_t* p, uint64_t n)
{
    if (n == 0) return;
    hipLaunchKernelGGL(gz_k_iota, dim3(gz_blocks_for(n)), dim3(GZ_BLOCK), 0, gz_stream, p, n);
    GZ_LAUNCH_CHECK();
}

extern "C" void gz_dev_store_u64(uint64_t* p, uint64_t index, uint64_t value)
{
    hipLaunchKernelGGL(gz_k_store, dim3(1), dim3(1), 0, gz_stream, p, index, value);
    GZ_LAUNCH_CHECK();
}

/* In-degrees from either layout: the pull list holds each destination's
 * in-edges as its vectors' valid lanes; in the push list every valid lane
 * names one destination. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_indegrees(const gz_vec_t* __restrict__ vecs, uint64_t count,
                                                           int pull_layout, uint32_t* __restrict__ indeg)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const gz_lanes v = gz_load_vector(vecs, i);
        const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
        if (pull_layout) {
            uint32_t c = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) c += gz_lane_valid(l[k]) ? 1u : 0u;
            atomicAdd(&indeg[gz_shared_vertex(v)], c);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (gz_lane_valid(l[k])) atomicAdd(&indeg[gz_lane_id(l[k])], 1u);
        }
    }
}

extern "C" void gz_dev_indegrees(const gz_vec_t* vecs, uint64_t count, int pull_layout, uint32_t* indeg, uint64_t n)
{
    gz_device_zero(indeg, (n + 8) * sizeof(uint32_t));
    if (count == 0) return;
    hipLaunchKernelGGL(gz_k_indegrees, dim3(gz_blocks_for(count)), dim3(GZ_BLOCK), 0, gz_stream, vecs, count,
                       pull_layout, indeg);
    GZ_LAUNCH_CHECK();
}
