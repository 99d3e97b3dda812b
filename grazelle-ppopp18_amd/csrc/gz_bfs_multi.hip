/*
 * gz_bfs_multi.hip -- breadth-first searches from up to 64 roots at once (no
 * reference counterpart; the reference searches from one root,
 * execution_bfs.c:28,135).
 *
 * Each vertex holds three 64-bit words whose bit r belongs to search r:
 * seen (visited before this level), front (visited in the previous level)
 * and next (visited in this level).  One pass over the graph per level
 * advances every search, with the reference's level-synchronous hybrid
 * (execution_bfs.c:150-168) on the union of the frontiers: the frontier
 * statistic is 1 + outdeg(v) summed over the vertices some search visited in
 * the last level, and a level pulls when it exceeds E/5.
 *
 *  - Push (top-down): the vertices with a nonzero front word are compacted
 *    (gz_frontier_compact over a bitmap of them) and their vectors walked;
 *    each valid lane ORs front[s] & ~seen[d] into next[d].
 *  - Pull (bottom-up): the chunk walker of the pull sweeps (gz_walk.h) with
 *    bitwise OR in place of the sum: each destination's run ORs the front
 *    words of its sources, and the bits it still needs go into next[d] with
 *    an atomic OR.  A vector whose destination every search has seen skips
 *    its gathers.
 *  - Per level, one pass over the vertices folds next into seen, records the
 *    level as the depth of each new (search, vertex) pair (one byte per
 *    pair, the 64 searches of a vertex side by side), rebuilds the frontier
 *    bitmap and sums the frontier statistic.
 * OR is associative and commutative, so the levels, and with them the depths,
 * are the BFS distances whatever the order of the atomics: the same depths a
 * single search from each root gives.  Parents are derived per search on
 * request: the smallest in-neighbour one level up (a valid BFS tree).  Depths
 * are bytes, so a batch stops with an error past 254 levels.
 */
#include "grazelle_host.h"
#include "gz_host_internal.h"
#include "gz_slots.h"
#include "gz_walk.h"

#include <string.h>
#include <time.h>

#define GZ_MSBFS_UNREACHED 0xffu

namespace {

struct gz_msbfs_state {
    uint64_t vertices;          /* V the buffers are sized for */
    uint64_t* seen;             /* [V] */
    uint64_t* front;            /* [V] */
    uint64_t* next;             /* [V] */
    uint8_t* depth;             /* [V][64] */
    uint64_t* fbits;            /* [ceil(V/64)] front[v] != 0 */
    uint64_t* roots;            /* [64] device roots of the last batch */
    uint32_t count;             /* searches in the last batch */
    uint64_t levels;
};

gz_msbfs_state M;

/* Roots: front and seen get the searches' bits, depth 0 for each. */
__global__ void gz_k_msbfs_init(const uint64_t* __restrict__ roots, uint32_t count, uint64_t* __restrict__ front,
                                uint64_t* __restrict__ seen, uint8_t* __restrict__ depth, uint64_t* __restrict__ fbits)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (uint32_t r = 0; r < count; ++r) {
        const uint64_t v = roots[r];
        front[v] |= 1ull << r;
        seen[v] |= 1ull << r;
        depth[v * 64 + r] = 0;
        fbits[v >> 6] |= 1ull << (v & 63);
    }
}

/* Pull: OR of the sources' front words per destination run. */
struct gz_msbfs_pull_op {
    const uint64_t* __restrict__ front;
    const uint64_t* __restrict__ seen;
    uint64_t* __restrict__ next;
    uint64_t full;
    __device__ __forceinline__ uint64_t identity() const { return 0; }
    __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) const { return a | b; }
    __device__ __forceinline__ uint64_t gather(const gz_lanes& v, bool in_range, uint32_t& edges) const
    {
        const bool need = in_range && (full & ~seen[gz_shared_vertex(v)]) != 0;
        const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
        uint64_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = front[(need && gz_lane_valid(l[k])) ? gz_lane_id(l[k]) : 0];
        uint64_t m = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ok = need && gz_lane_valid(l[k]);
            m |= ok ? x[k] : 0ull;
            edges += ok ? 1u : 0u;
        }
        return m;
    }
    __device__ __forceinline__ void store(uint64_t d, uint64_t m)
    {
        const uint64_t add = m & full & ~seen[d];
        if (add != 0 && (add & ~next[d]) != 0) atomicOr((unsigned long long*)&next[d], (unsigned long long)add);
    }
};

template <int VPL>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_msbfs_pull(const gz_vec_t* __restrict__ vecs, uint64_t nvec,
                                                            uint64_t chunk_len, uint64_t nchunks,
                                                            gz_msbfs_pull_op op)
{
    const uint64_t chunk = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    if (chunk >= nchunks) return;
    const uint64_t b = chunk * chunk_len;
    const uint64_t e = (b + chunk_len < nvec) ? b + chunk_len : nvec;
    uint64_t first_d, carry_d, carry_m;
    uint32_t edges = 0;
    gz_walk_chunk<VPL, uint64_t>(vecs, b, e, op, first_d, carry_d, carry_m, edges);
    if (gz_lane() == 0) op.store(carry_d, carry_m);
}

/* Push over the compacted frontier's vectors. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_msbfs_push(const gz_vec_t* __restrict__ vecs, gz_slot_source src,
                                                            uint64_t chunk_len, const uint64_t* __restrict__ front,
                                                            const uint64_t* __restrict__ seen,
                                                            uint64_t* __restrict__ next, uint64_t full)
{
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t W = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    const uint64_t T = gz_slot_total<true>(src);
    for (uint64_t chunk = gw; chunk * chunk_len < T; chunk += W) {
        const uint64_t j0 = chunk * chunk_len;
        const uint64_t jend = (j0 + chunk_len < T) ? j0 + chunk_len : T;
        uint64_t a = gz_slot_begin<true>(src, j0);
        for (uint64_t base = j0; base < jend; base += GZ_WAVE) {
            const bool valid = base + (uint64_t)lane < jend;
            uint64_t entry;
            const uint64_t vi = gz_slot_step<true>(src, a, base, lane, valid, &entry);
            if (!valid) continue;
            const gz_lanes v = gz_load_vector(vecs, vi);
            const uint64_t f = front[gz_shared_vertex(v)] & full;
            const uint64_t l[4] = {v.l0, v.l1, v.l2, v.l3};
            uint64_t s[4], n[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t d = gz_lane_valid(l[k]) ? gz_lane_id(l[k]) : 0;
                s[k] = seen[d];
                n[k] = next[d];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!gz_lane_valid(l[k])) continue;
                const uint64_t add = f & ~s[k];
                if ((add & ~n[k]) != 0)
                    atomicOr((unsigned long long*)&next[gz_lane_id(l[k])], (unsigned long long)add);
            }
        }
    }
}

/* 0xff in byte j of the result for each set bit j of the low 8 bits of b. */
__device__ __forceinline__ uint64_t gz_byte_mask(uint64_t b)
{
    uint64_t m = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) m |= ((b >> j) & 1ull) ? (0xffull << (8 * j)) : 0ull;
    return m;
}

/* End of a level: seen |= next, depths of the new pairs, front = next,
 * next = 0, the frontier bitmap (one word per wavefront step) and the
 * frontier statistic 1 + outdeg(v) of every vertex some search reached. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_msbfs_level_end(uint64_t num_vertices, uint64_t level, uint64_t full,
                                                                 uint64_t* __restrict__ front,
                                                                 uint64_t* __restrict__ seen,
                                                                 uint64_t* __restrict__ next,
                                                                 uint8_t* __restrict__ depth,
                                                                 uint64_t* __restrict__ fbits,
                                                                 const double* __restrict__ outdeg,
                                                                 uint64_t* __restrict__ partials)
{
    const int lane = gz_lane();
    const uint64_t gw = (uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + (threadIdx.x / GZ_WAVE);
    const uint64_t W = (uint64_t)gridDim.x * GZ_WAVES_PER_BLOCK;
    const uint64_t level_bytes = level * 0x0101010101010101ull;
    uint64_t vote = 0;
    for (uint64_t base = gw * GZ_WAVE; base < num_vertices; base += W * GZ_WAVE) {
        const uint64_t v = base + (uint64_t)lane;
        const bool in = v < num_vertices;
        const uint64_t n = in ? next[v] : 0;
        const uint64_t f = in ? front[v] : 0;
        if (n != 0) {
            seen[v] |= n;
            next[v] = 0;
            uint64_t* dw = reinterpret_cast<uint64_t*>(depth + v * 64);
#pragma unroll
            for (int w = 0; w < 8; ++w) {
                const uint64_t nb = (n >> (8 * w)) & 0xffull;
                if (nb == 0) continue;
                const uint64_t m = gz_byte_mask(nb);
                dw[w] = (dw[w] & ~m) | (level_bytes & m);
            }
            vote += 1 + (uint64_t)llrint(outdeg[v]);
        }
        if (in && f != n) front[v] = n;
        const uint64_t fb = __ballot(n != 0);
        if (lane == 0) fbits[base >> 6] = fb;
    }
    const uint64_t total = gz_wave_sum_u64(vote);
    if (lane == 0) partials[gw] = total;
}

/* Depths of search r as words (UINT64_MAX unreached) and parents (device
 * ids, ~0 for the root and unreached vertices): the smallest in-neighbour one
 * level up, in the pull list's source order. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_msbfs_result(const gz_vec_t* __restrict__ pull,
                                                              const uint64_t* __restrict__ offsets,
                                                              uint64_t num_vertices, uint32_t r,
                                                              const uint8_t* __restrict__ depth,
                                                              uint64_t* __restrict__ out_depth,
                                                              uint64_t* __restrict__ out_parent, int want_parents)
{
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < num_vertices;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = depth[v * 64 + r];
        out_depth[v] = (d == GZ_MSBFS_UNREACHED) ? ~0ull : (uint64_t)d;
        if (!want_parents) continue;
        uint64_t parent = ~0ull;
        if (d != GZ_MSBFS_UNREACHED && d != 0) {
            for (uint64_t i = offsets[v]; i < offsets[v + 1] && parent == ~0ull; ++i) {
                const gz_lanes w = gz_load_vector(pull, i);
                const uint64_t l[4] = {w.l0, w.l1, w.l2, w.l3};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (parent != ~0ull || !gz_lane_valid(l[k])) continue;
                    const uint64_t u = gz_lane_id(l[k]);
                    if ((uint32_t)depth[u * 64 + r] + 1u == d) parent = u;
                }
            }
        }
        out_parent[v] = parent;
    }
}

void gz_msbfs_release_buffers()
{
    gz_device_free(M.seen);
    gz_device_free(M.front);
    gz_device_free(M.next);
    gz_device_free(M.depth);
    gz_device_free(M.fbits);
    gz_device_free(M.roots);
    memset(&M, 0, sizeof(M));
}

void gz_msbfs_reserve(uint64_t V)
{
    if (M.seen != nullptr && M.vertices == V) return;
    gz_msbfs_release_buffers();
    const uint64_t words = (V + 63) / 64 + 8;
    M.seen = (uint64_t*)gz_device_malloc((V + 8) * sizeof(uint64_t));
    M.front = (uint64_t*)gz_device_malloc((V + 8) * sizeof(uint64_t));
    M.next = (uint64_t*)gz_device_malloc((V + 8) * sizeof(uint64_t));
    M.depth = (uint8_t*)gz_device_malloc((V + 8) * 64);
    M.fbits = (uint64_t*)gz_device_malloc(words * sizeof(uint64_t));
    M.roots = (uint64_t*)gz_device_malloc(64 * sizeof(uint64_t));
    M.vertices = V;
}

}  // namespace

extern "C" void gz_bfs_batch_release(void) { gz_msbfs_release_buffers(); }

extern "C" int gz_bfs_batch_run(const uint64_t* roots, uint32_t count, gz_run_result_t* res)
{
    memset(res, 0, sizeof(*res));
    const uint64_t V = gz_graph.num_vertices, E = gz_graph.num_edges;
    if (count == 0 || count > GZ_BFS_BATCH_MAX_ROOTS) {
        fprintf(stderr, "grazelle: a BFS batch holds 1..%d roots (got %u)\n", GZ_BFS_BATCH_MAX_ROOTS, count);
        return -1;
    }
    if (gz_graph.pull_vectors == nullptr || gz_graph.push_vectors == nullptr || gz_graph.push_vertex_offsets == nullptr ||
        gz_graph.pull_vertex_offsets == nullptr) {
        fprintf(stderr, "grazelle: a BFS batch needs both layouts\n");
        return -1;
    }
    uint64_t dev_roots[GZ_BFS_BATCH_MAX_ROOTS];
    for (uint32_t r = 0; r < count; ++r) {
        if (roots[r] >= V) {
            fprintf(stderr, "grazelle: BFS root %llu is not a vertex of a %llu-vertex graph\n",
                    (unsigned long long)roots[r], (unsigned long long)V);
            return -1;
        }
        dev_roots[r] = gz_graph_device_vertex(roots[r]);
    }
    gz_msbfs_reserve(V);
    gz_scratch_reserve(GZ_PERSISTENT_BLOCKS * GZ_WAVES_PER_BLOCK, V);
    const uint64_t words = (V + 63) / 64;
    gz_device_zero(M.seen, V * sizeof(uint64_t));
    gz_device_zero(M.front, V * sizeof(uint64_t));
    gz_device_zero(M.next, V * sizeof(uint64_t));
    gz_dev_fill_u64((uint64_t*)M.depth, ~0ull, V * 8);
    gz_device_zero(M.fbits, words * sizeof(uint64_t));
    gz_copy_to_device(M.roots, dev_roots, count * sizeof(uint64_t));
    M.count = count;
    const uint64_t full = count >= 64 ? ~0ull : ((1ull << count) - 1ull);
    /* the roots' frontier statistic: 1 + outdeg for each distinct root */
    const double* od = gz_graph_outdegrees_host();
    uint64_t vote = 0;
    for (uint32_t r = 0; r < count; ++r) {
        int dup = 0;
        for (uint32_t q = 0; q < r; ++q) dup |= dev_roots[q] == dev_roots[r];
        if (!dup) vote += 1 + (uint64_t)od[dev_roots[r]];
    }
    uint64_t* reduce = gz_scratch.reduce_tmp + 2032; /* one word the vote reduction writes */
    gz_synchronize();
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    hipLaunchKernelGGL(gz_k_msbfs_init, dim3(1), dim3(64), 0, gz_stream, M.roots, count, M.front, M.seen, M.depth,
                       M.fbits);
    GZ_LAUNCH_CHECK();
    const uint64_t threshold = E / 5;
    const uint64_t chunk_len = gz_chunk_len();
    const uint64_t nchunks = gz_chunks_for(gz_graph.pull_vector_count);
    const uint64_t pblocks = (nchunks + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    uint64_t eblocks = (words + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK;
    if (eblocks > GZ_PERSISTENT_BLOCKS) eblocks = GZ_PERSISTENT_BLOCKS;
    if (eblocks == 0) eblocks = 1;
    uint64_t level = 0;
    int rc = 0;
    while (vote != 0) {
        level += 1;
        if (level >= GZ_MSBFS_UNREACHED) {
            fprintf(stderr, "grazelle: a BFS batch records depths up to %u levels\n", GZ_MSBFS_UNREACHED - 1);
            rc = -1;
            break;
        }
        if (vote > threshold) {
            res->pull_iterations += 1;
            const gz_msbfs_pull_op op{M.front, M.seen, M.next, full};
            if (nchunks > 0) {
                const uint32_t vpl = gz_vectors_per_lane();
                if (vpl == 4)
                    hipLaunchKernelGGL(gz_k_msbfs_pull<4>, dim3((unsigned)pblocks), dim3(GZ_BLOCK), 0, gz_stream,
                                       gz_graph.pull_vectors, gz_graph.pull_vector_count, chunk_len, nchunks, op);
                else if (vpl == 2)
                    hipLaunchKernelGGL(gz_k_msbfs_pull<2>, dim3((unsigned)pblocks), dim3(GZ_BLOCK), 0, gz_stream,
                                       gz_graph.pull_vectors, gz_graph.pull_vector_count, chunk_len, nchunks, op);
                else
                    hipLaunchKernelGGL(gz_k_msbfs_pull<1>, dim3((unsigned)pblocks), dim3(GZ_BLOCK), 0, gz_stream,
                                       gz_graph.pull_vectors, gz_graph.pull_vector_count, chunk_len, nchunks, op);
                GZ_LAUNCH_CHECK();
            }
        } else {
            res->push_iterations += 1;
            gz_frontier_compact(M.fbits, gz_graph.push_vertex_offsets, V, gz_graph.push_vector_count);
            gz_slot_source src;
            src.offsets = gz_graph.push_vertex_offsets;
            src.list = gz_scratch.active_vertices;
            src.prefix = gz_scratch.active_prefix;
            src.totals = gz_scratch.active_totals;
            src.sweep_count = gz_graph.push_vector_count;
            hipLaunchKernelGGL(gz_k_msbfs_push, dim3(GZ_PERSISTENT_BLOCKS), dim3(GZ_BLOCK), 0, gz_stream,
                               gz_graph.push_vectors, src, (uint64_t)256, M.front, M.seen, M.next, full);
            GZ_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(gz_k_msbfs_level_end, dim3((unsigned)eblocks), dim3(GZ_BLOCK), 0, gz_stream, V, level, full,
                           M.front, M.seen, M.next, M.depth, M.fbits, gz_graph.vertex_outdegrees,
                           (uint64_t*)gz_scratch.partials);
        GZ_LAUNCH_CHECK();
        gz_scratch.partials_is_double = 0;
        gz_scratch.partials_count = eblocks * GZ_WAVES_PER_BLOCK;
        gz_reduce_partials(reduce);
        gz_copy_to_host(&vote, reduce, sizeof(uint64_t));
    }
    gz_synchronize();
    clock_gettime(CLOCK_MONOTONIC, &t1);
    res->elapsed_ms = (double)(t1.tv_sec - t0.tv_sec) * 1e3 + (double)(t1.tv_nsec - t0.tv_nsec) * 1e-6;
    res->iterations = level;
    M.levels = level;
    return rc;
}

extern "C" void gz_bfs_batch_result(uint32_t search, int64_t* parents, uint64_t* depths)
{
    const uint64_t V = gz_graph.num_vertices;
    if (M.depth == nullptr || search >= M.count || M.vertices != V) {
        fprintf(stderr, "grazelle: no BFS batch result for search %u\n", search);
        exit(66);
    }
    uint64_t* d_depth = (uint64_t*)gz_device_malloc((V + 8) * sizeof(uint64_t));
    uint64_t* d_parent = parents ? (uint64_t*)gz_device_malloc((V + 8) * sizeof(uint64_t)) : nullptr;
    uint64_t blocks = (V + GZ_BLOCK - 1) / GZ_BLOCK;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(gz_k_msbfs_result, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, gz_graph.pull_vectors,
                       gz_graph.pull_vertex_offsets, V, search, M.depth, d_depth, d_parent, parents != nullptr);
    GZ_LAUNCH_CHECK();
    uint64_t* h = (uint64_t*)malloc((V ? V : 1) * sizeof(uint64_t));
    uint64_t* hp = parents ? (uint64_t*)malloc((V ? V : 1) * sizeof(uint64_t)) : nullptr;
    gz_copy_to_host(h, d_depth, V * sizeof(uint64_t));
    if (parents) gz_copy_to_host(hp, d_parent, V * sizeof(uint64_t));
    gz_device_free(d_depth);
    gz_device_free(d_parent);
    /* original-id order: result[original id] = value at its device vertex */
    uint64_t* orig = (uint64_t*)malloc((V ? V : 1) * sizeof(uint64_t));
    gz_graph_original_ids(orig);
    for (uint64_t v = 0; v < V; ++v) {
        if (depths) depths[orig[v]] = h[v];
        if (parents) parents[orig[v]] = (hp[v] == ~0ull) ? -1 : (int64_t)orig[hp[v]];
    }
    free(orig);
    free(h);
    free(hp);
}
