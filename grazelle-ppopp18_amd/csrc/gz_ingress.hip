/*
 * gz_ingress.hip -- device Vector-Sparse ingress: the edge lists are built in
 * HBM instead of by the host packer.
 *
 * The reference loads -pull / -push files with a producer thread and one
 * packing consumer per layout (source/graphdata.c:595-951), which at a
 * billion edges is minutes of host time.  Here the same result is computed
 * as a device pipeline over 64-bit keys (group << b | neighbour), b = bits of
 * V-1:
 *   1. keys: drawn by the R-MAT kernel (the host generator's exact stream) or
 *      converted from uploaded (src, dst) pairs, optionally both directions;
 *   2. LSD radix sort over the 2b key bits (pull: group = dst);
 *   3. drop self-loops and duplicates (a key equal to its predecessor);
 *   4. per-vertex edge offsets by binary search in the sorted keys, vector
 *      counts ceil(deg / 4) and their exclusive scan = the CSR vector offsets;
 *   5. pack: edge i of group g at rank r goes to lane r % 4 of vector
 *      offsets[g] + r / 4, which is the flush rule of graphdata.c:679-686 on a
 *      group-sorted list (a vector closes when the shared vertex changes or it
 *      holds four edges);
 *   6. the push layout: the kept keys with group and neighbour swapped,
 *      sorted again; FP64 outdegrees are its group sizes (graphdata.c:814
 *      counts sources of the pull file, which is the same number).
 * The result equals gz_edges_sort(..., dedupe) + gz_pack_vectors on the host
 * bit for bit, invalid lanes included (shared-vertex piece, neighbour 0).
 */
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "grazelle_host.h"
#include "gz_host_internal.h"
#include "gz_internal.h"

namespace {

__device__ __forceinline__ uint64_t gz_splitmix64_dev(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

/* The R-MAT draw of rmat.c for edges [0, n): key (dst << b | src), and with
 * `both` also (src << b | dst) at the odd slot.  Ids are shifted by id_base
 * (one graph of a disjoint union). */
__global__ void gz_k_rmat_keys(uint64_t seed, uint64_t edge_offset, uint32_t scale, uint64_t nv, uint32_t ta,
                               uint32_t tb, uint32_t tc, uint64_t n, uint32_t b, int both, uint64_t id_base,
                               uint64_t* __restrict__ keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        uint64_t x = gz_splitmix64_dev(seed ^ gz_splitmix64_dev(edge_offset + e));
        uint64_t s = 0, d = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            if (l && (l & 3) == 0) x = gz_splitmix64_dev(x);
            const uint32_t u = (uint32_t)((x >> (16 * (l & 3))) & 0xffff);
            const uint32_t sb = (u >= tb) ? 1u : 0u;
            const uint32_t db = ((u >= ta && u < tb) || (u >= tc)) ? 1u : 0u;
            s = (s << 1) | sb;
            d = (d << 1) | db;
        }
        if (s >= nv) s %= nv;
        if (d >= nv) d %= nv;
        s += id_base;
        d += id_base;
        if (both) {
            keys[2 * e] = (d << b) | s;
            keys[2 * e + 1] = (s << b) | d;
        } else {
            keys[e] = (d << b) | s;
        }
    }
}

__global__ void gz_k_pairs_keys(const uint64_t* __restrict__ pairs, uint64_t n, uint32_t b, int both,
                                uint64_t* __restrict__ keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const uint64_t s = pairs[2 * e], d = pairs[2 * e + 1];
        if (both) {
            keys[2 * e] = (d << b) | s;
            keys[2 * e + 1] = (s << b) | d;
        } else {
            keys[e] = (d << b) | s;
        }
    }
}

/* keep = first of its value and not a self-loop */
__global__ void gz_k_mark_kept(const uint64_t* __restrict__ keys, uint64_t n, uint32_t b, uint8_t* __restrict__ keep)
{
    const uint64_t mask = (1ull << b) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        const bool dup = (i > 0) && keys[i - 1] == k;
        keep[i] = (!dup && (k >> b) != (k & mask)) ? 1 : 0;
    }
}

__global__ void gz_k_swap_keys(const uint64_t* __restrict__ in, uint64_t n, uint32_t b, uint64_t* __restrict__ out)
{
    const uint64_t mask = (1ull << b) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = in[i];
        out[i] = ((k & mask) << b) | (k >> b);
    }
}

/* eoff[v] = first i with keys[i] >> b >= v, for v in [0, V] (eoff[V] = m). */
__global__ void gz_k_group_offsets(const uint64_t* __restrict__ keys, uint64_t m, uint32_t b, uint64_t nv,
                                   uint64_t* __restrict__ eoff)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= nv; v += stride) {
        if (v == nv) {
            eoff[v] = m;
            continue;
        }
        const uint64_t target = v << b;
        uint64_t lo = 0, hi = m;
        while (lo < hi) {
            const uint64_t mid = lo + ((hi - lo) >> 1);
            if (keys[mid] < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        eoff[v] = lo;
    }
}

__global__ void gz_k_vector_counts(const uint64_t* __restrict__ eoff, uint64_t nv, uint64_t* __restrict__ cnt)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= nv; v += stride)
        cnt[v] = (v < nv) ? (eoff[v + 1] - eoff[v] + 3) / 4 : 0;
}

__device__ __forceinline__ uint64_t gz_piece(uint64_t g, int l)
{
    const int shift = 15 * l;
    const uint64_t mask = (l == 3) ? 0x7ull : 0x7fffull;
    return ((g >> shift) & mask) << 48;
}

__global__ void gz_k_pack(const uint64_t* __restrict__ keys, uint64_t m, uint32_t b, const uint64_t* __restrict__ eoff,
                          const uint64_t* __restrict__ voff, uint64_t* __restrict__ lanes)
{
    const uint64_t mask = (1ull << b) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const uint64_t k = keys[i];
        const uint64_t g = k >> b;
        const uint64_t r = i - eoff[g];
        const uint64_t j = voff[g] + r / 4;
        const int l = (int)(r & 3);
        lanes[4 * j + l] = GZ_LANE_VALID_BIT | gz_piece(g, l) | (k & mask);
        if (l == 0) {
            const uint64_t left = eoff[g + 1] - eoff[g] - r;
            for (int q = (left < 4 ? (int)left : 4); q < 4; ++q) lanes[4 * j + q] = gz_piece(g, q);
        }
    }
}

__global__ void gz_k_outdegrees(const uint64_t* __restrict__ eoff, uint64_t nv, double* __restrict__ outdeg)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv + 8; v += stride)
        outdeg[v] = (v < nv) ? (double)(eoff[v + 1] - eoff[v]) : 0.0;
}

unsigned gz_grid_for(uint64_t n)
{
    const uint64_t blocks = (n + GZ_BLOCK - 1) / GZ_BLOCK;
    return (unsigned)(blocks < 65536 ? (blocks ? blocks : 1) : 65536);
}

/* Slice size actually used: at most V / (8 T), so the windows fit, and at
 * most 2^23, so that packed window offsets fit in 23 bits. */
uint64_t gz_slices_effective(uint64_t slice_vertices, uint32_t tiers, uint64_t nv)
{
    uint64_t cap = nv / ((uint64_t)GZ_XCDS * tiers);
    if (cap > (1ull << 23)) cap = 1ull << 23;
    return slice_vertices < cap ? slice_vertices : cap;
}

uint32_t gz_id_bits(uint64_t nv)
{
    uint32_t b = 1;
    while (b < 64 && (nv - 1) >> b) ++b;
    return b;
}

/* Sort keys[0, n) over bits [0, 2b); the sorted keys end in *sorted (one of a
 * or b_buf). */
void gz_sort_keys(uint64_t* a, uint64_t* b_buf, uint64_t n, uint32_t bits, uint64_t** sorted)
{
    rocprim::double_buffer<uint64_t> db(a, b_buf);
    size_t tmp_bytes = 0;
    GZ_CHECK(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, (size_t)n, 0, bits, gz_stream));
    void* tmp = gz_device_malloc(tmp_bytes ? tmp_bytes : 1);
    GZ_CHECK(rocprim::radix_sort_keys(tmp, tmp_bytes, db, (size_t)n, 0, bits, gz_stream));
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_device_free(tmp);
    *sorted = db.current();
}

/* Build one layout from sorted, duplicate-free keys[0, m). */
void gz_build_layout(const uint64_t* keys, uint64_t m, uint32_t b, uint64_t nv, gz_vec_t** d_vecs,
                     uint64_t** d_offsets, uint64_t* vector_count, uint64_t* eoff)
{
    const unsigned tb = GZ_BLOCK;
    hipLaunchKernelGGL(gz_k_group_offsets, dim3(gz_grid_for(nv + 1)), dim3(tb), 0, gz_stream, keys, m, b, nv, eoff);
    GZ_LAUNCH_CHECK();
    uint64_t* cnt = (uint64_t*)gz_device_malloc((nv + 1) * sizeof(uint64_t));
    uint64_t* voff = (uint64_t*)gz_device_malloc((nv + 1) * sizeof(uint64_t));
    hipLaunchKernelGGL(gz_k_vector_counts, dim3(gz_grid_for(nv + 1)), dim3(tb), 0, gz_stream, eoff, nv, cnt);
    GZ_LAUNCH_CHECK();
    size_t tmp_bytes = 0;
    GZ_CHECK(rocprim::exclusive_scan(nullptr, tmp_bytes, cnt, voff, (uint64_t)0, (size_t)(nv + 1),
                                     rocprim::plus<uint64_t>(), gz_stream));
    void* tmp = gz_device_malloc(tmp_bytes ? tmp_bytes : 1);
    GZ_CHECK(rocprim::exclusive_scan(tmp, tmp_bytes, cnt, voff, (uint64_t)0, (size_t)(nv + 1),
                                     rocprim::plus<uint64_t>(), gz_stream));
    uint64_t total = 0;
    GZ_CHECK(hipMemcpyAsync(&total, voff + nv, sizeof(uint64_t), hipMemcpyDeviceToHost, gz_stream));
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_device_free(tmp);
    gz_device_free(cnt);
    gz_vec_t* vecs = (gz_vec_t*)gz_device_malloc((total ? total : 1) * sizeof(gz_vec_t));
    if (m) {
        hipLaunchKernelGGL(gz_k_pack, dim3(gz_grid_for(m)), dim3(tb), 0, gz_stream, keys, m, b, eoff, voff,
                           (uint64_t*)vecs);
        GZ_LAUNCH_CHECK();
    }
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    *d_vecs = vecs;
    *d_offsets = voff;
    *vector_count = total;
}

__global__ void gz_k_count_sources(const uint64_t* __restrict__ keys, uint64_t n, uint32_t b, uint32_t* __restrict__ cnt)
{
    const uint64_t mask = (1ull << b) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&cnt[keys[i] & mask], 1u);
}

/* sort key: out-degree descending, then id ascending */
__global__ void gz_k_degree_order(const uint32_t* __restrict__ cnt, uint64_t nv, uint64_t* __restrict__ order)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride)
        order[v] = ((uint64_t)(0xffffffffu - cnt[v]) << 32) | v;
}

/* Rank i of the degree order becomes id i, except that each of the first
 * `tiers` runs of 8*S ranks is dealt round robin into eight S-id windows:
 * rank t*8S + q -> t*8S + (q % 8) * S + q / 8. */
__global__ void gz_k_permutation(const uint64_t* __restrict__ order, uint64_t nv, uint64_t slice_vertices,
                                 uint32_t tiers, uint64_t* __restrict__ old_of_new, uint32_t* __restrict__ new_of_old)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tier = (uint64_t)GZ_XCDS * slice_vertices;
    const uint64_t hot = tier * tiers;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        const uint64_t old = order[i] & 0xffffffffull;
        uint64_t id = i;
        if (i < hot) {
            const uint64_t t = i / tier, q = i - t * tier;
            id = t * tier + (q % GZ_XCDS) * slice_vertices + q / GZ_XCDS;
        }
        old_of_new[id] = old;
        new_of_old[old] = (uint32_t)id;
    }
}

__global__ void gz_k_remap_keys(uint64_t* __restrict__ keys, uint64_t n, uint32_t b,
                                const uint32_t* __restrict__ new_of_old)
{
    const uint64_t mask = (1ull << b) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        keys[i] = ((uint64_t)new_of_old[k >> b] << b) | (uint64_t)new_of_old[k & mask];
    }
}

/* Renumber the vertices by raw out-degree, descending (ties by id), and
 * rewrite the keys in place.  Returns the device array of original ids,
 * indexed by new id.  The hot sources of a power-law graph then share cache
 * lines, which is what the pull gathers need once the vertex array no longer
 * fits in the caches. */
uint64_t* gz_relabel_by_outdegree(uint64_t* keys, uint64_t n, uint32_t b, uint64_t nv, uint64_t slice_vertices,
                                  uint32_t tiers)
{
    uint32_t* cnt = (uint32_t*)gz_device_malloc(nv * sizeof(uint32_t));
    GZ_CHECK(hipMemsetAsync(cnt, 0, nv * sizeof(uint32_t), gz_stream));
    if (n) {
        hipLaunchKernelGGL(gz_k_count_sources, dim3(gz_grid_for(n)), dim3(GZ_BLOCK), 0, gz_stream, keys, n, b, cnt);
        GZ_LAUNCH_CHECK();
    }
    uint64_t* order = (uint64_t*)gz_device_malloc(nv * sizeof(uint64_t));
    uint64_t* order_b = (uint64_t*)gz_device_malloc(nv * sizeof(uint64_t));
    hipLaunchKernelGGL(gz_k_degree_order, dim3(gz_grid_for(nv)), dim3(GZ_BLOCK), 0, gz_stream, cnt, nv, order);
    GZ_LAUNCH_CHECK();
    uint64_t* sorted = order;
    if (nv > 1) gz_sort_keys(order, order_b, nv, 64, &sorted);
    uint64_t* old_of_new = (uint64_t*)gz_device_malloc((nv + 8) * sizeof(uint64_t));
    uint32_t* new_of_old = cnt; /* counts are no longer needed */
    hipLaunchKernelGGL(gz_k_permutation, dim3(gz_grid_for(nv)), dim3(GZ_BLOCK), 0, gz_stream, sorted, nv,
                       slice_vertices, tiers, old_of_new, new_of_old);
    GZ_LAUNCH_CHECK();
    if (n) {
        hipLaunchKernelGGL(gz_k_remap_keys, dim3(gz_grid_for(n)), dim3(GZ_BLOCK), 0, gz_stream, keys, n, b, new_of_old);
        GZ_LAUNCH_CHECK();
    }
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_device_free(order);
    gz_device_free(order_b);
    gz_device_free(cnt);
    return old_of_new;
}

/* Keys already in a[0, n): optionally renumber, sort, drop self-loops and
 * duplicates, build both layouts and hand them to the graph record. */
void gz_ingress_from_keys(uint64_t nv, uint64_t* a, uint64_t* b_buf, uint64_t n, uint32_t b, uint32_t flags)
{
    const uint32_t slice_log2 = (flags >> 8) & 0x3fu;
    uint32_t tiers = (flags >> 14) & 0x3u;
    if (tiers == 0) tiers = 1;
    const uint32_t blocks = (flags >> 16) & 0x7u;
    const uint64_t slice_vertices = slice_log2 ? gz_slices_effective(1ull << slice_log2, tiers, nv) : 0;
    uint64_t* old_of_new = nullptr;
    if (flags & GZ_INGRESS_RELABEL_BY_OUTDEGREE)
        old_of_new = gz_relabel_by_outdegree(a, n, b, nv, slice_vertices, tiers);
    uint64_t* sorted = a;
    if (n > 1) gz_sort_keys(a, b_buf, n, 2 * b, &sorted);
    uint64_t* other = (sorted == a) ? b_buf : a;
    uint64_t m = 0;
    if (n) {
        uint8_t* keep = (uint8_t*)gz_device_malloc(n);
        uint64_t* d_count = (uint64_t*)gz_device_malloc(sizeof(uint64_t));
        hipLaunchKernelGGL(gz_k_mark_kept, dim3(gz_grid_for(n)), dim3(GZ_BLOCK), 0, gz_stream, sorted, n, b, keep);
        GZ_LAUNCH_CHECK();
        size_t tmp_bytes = 0;
        GZ_CHECK(rocprim::select(nullptr, tmp_bytes, sorted, keep, other, d_count, (size_t)n, gz_stream));
        void* tmp = gz_device_malloc(tmp_bytes ? tmp_bytes : 1);
        GZ_CHECK(rocprim::select(tmp, tmp_bytes, sorted, keep, other, d_count, (size_t)n, gz_stream));
        GZ_CHECK(hipMemcpyAsync(&m, d_count, sizeof(uint64_t), hipMemcpyDeviceToHost, gz_stream));
        GZ_CHECK(hipStreamSynchronize(gz_stream));
        gz_device_free(tmp);
        gz_device_free(keep);
        gz_device_free(d_count);
    }

    gz_device_layouts_t L;
    L.num_vertices = nv;
    L.num_edges = m;
    L.original_ids = old_of_new;
    uint64_t* eoff = (uint64_t*)gz_device_malloc((nv + 1) * sizeof(uint64_t));
    gz_build_layout(other, m, b, nv, &L.pull, &L.pull_offsets, &L.pull_count, eoff);
    /* push layout: swap group and neighbour, sort again (still unique) */
    if (m) {
        hipLaunchKernelGGL(gz_k_swap_keys, dim3(gz_grid_for(m)), dim3(GZ_BLOCK), 0, gz_stream, other, m, b, sorted);
        GZ_LAUNCH_CHECK();
    }
    uint64_t* push_sorted = sorted;
    if (m > 1) gz_sort_keys(sorted, other, m, 2 * b, &push_sorted);
    gz_build_layout(push_sorted, m, b, nv, &L.push, &L.push_offsets, &L.push_count, eoff);
    L.outdeg = (double*)gz_device_malloc((nv + 8) * sizeof(double));
    hipLaunchKernelGGL(gz_k_outdegrees, dim3(gz_grid_for(nv + 8)), dim3(GZ_BLOCK), 0, gz_stream, eoff, nv, L.outdeg);
    GZ_LAUNCH_CHECK();
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_device_free(eoff);
    gz_graph_adopt_device(&L);
    if (slice_log2) gz_graph_build_source_slices(1ull << slice_log2, tiers, blocks, GZ_SLICE_BLOCK_VERTICES);
}

void gz_check_vertex_count(uint64_t nv)
{
    if (nv == 0 || nv > 0xffffffffull) {
        fprintf(stderr, "grazelle: device ingress needs 1 <= V <= 2^32 (got %llu)\n", (unsigned long long)nv);
        exit(65);
    }
}

}  // namespace

extern "C" int gz_graph_build_pairs_device(uint64_t num_vertices, uint64_t num_edges, const uint64_t* pairs,
                                           int symmetrize, uint32_t flags)
{
    gz_runtime_init(0);
    gz_graph_release();
    gz_check_vertex_count(num_vertices);
    for (uint64_t i = 0; i < 2 * num_edges; ++i)
        if (pairs[i] >= num_vertices) {
            fprintf(stderr, "grazelle: edge endpoint %llu is not below V = %llu\n", (unsigned long long)pairs[i],
                    (unsigned long long)num_vertices);
            exit(65);
        }
    const uint32_t b = gz_id_bits(num_vertices);
    const uint64_t n = symmetrize ? 2 * num_edges : num_edges;
    uint64_t* d_pairs = (uint64_t*)gz_device_malloc((num_edges ? 2 * num_edges : 2) * sizeof(uint64_t));
    if (num_edges) gz_copy_to_device(d_pairs, pairs, 2 * num_edges * sizeof(uint64_t));
    uint64_t* a = (uint64_t*)gz_device_malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t* bb = (uint64_t*)gz_device_malloc((n ? n : 1) * sizeof(uint64_t));
    if (num_edges) {
        hipLaunchKernelGGL(gz_k_pairs_keys, dim3(gz_grid_for(num_edges)), dim3(GZ_BLOCK), 0, gz_stream, d_pairs,
                           num_edges, b, symmetrize, a);
        GZ_LAUNCH_CHECK();
    }
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_device_free(d_pairs);
    gz_ingress_from_keys(num_vertices, a, bb, n, b, flags);
    gz_device_free(a);
    gz_device_free(bb);
    return 0;
}

extern "C" int gz_graph_build_rmat_device(const gz_rmat_params_t* p, uint64_t num_vertices, int symmetrize,
                                          uint64_t extra_edges, const uint64_t* extra_pairs, uint32_t flags)
{
    gz_runtime_init(0);
    gz_graph_release();
    const uint64_t rmat_nv = p->num_vertices ? p->num_vertices : (1ull << p->scale);
    uint64_t nv = num_vertices > rmat_nv ? num_vertices : rmat_nv;
    for (uint64_t i = 0; i < 2 * extra_edges; ++i)
        if (extra_pairs[i] >= nv) {
            fprintf(stderr, "grazelle: extra edge endpoint %llu is not below V = %llu\n",
                    (unsigned long long)extra_pairs[i], (unsigned long long)nv);
            exit(65);
        }
    gz_check_vertex_count(nv);
    const uint32_t b = gz_id_bits(nv);
    const uint64_t per = symmetrize ? 2 : 1;
    const uint64_t n = per * (p->num_edges + extra_edges);
    uint64_t* a = (uint64_t*)gz_device_malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t* bb = (uint64_t*)gz_device_malloc((n ? n : 1) * sizeof(uint64_t));
    if (p->num_edges) {
        const uint32_t ta = (uint32_t)(p->a * 65536.0);
        const uint32_t tb = (uint32_t)((p->a + p->b) * 65536.0);
        const uint32_t tc = (uint32_t)((p->a + p->b + p->c) * 65536.0);
        hipLaunchKernelGGL(gz_k_rmat_keys, dim3(gz_grid_for(p->num_edges)), dim3(GZ_BLOCK), 0, gz_stream, p->seed,
                           p->edge_offset, p->scale, rmat_nv, ta, tb, tc, p->num_edges, b, symmetrize, (uint64_t)0, a);
        GZ_LAUNCH_CHECK();
    }
    if (extra_edges) {
        uint64_t* d_pairs = (uint64_t*)gz_device_malloc(2 * extra_edges * sizeof(uint64_t));
        gz_copy_to_device(d_pairs, extra_pairs, 2 * extra_edges * sizeof(uint64_t));
        hipLaunchKernelGGL(gz_k_pairs_keys, dim3(gz_grid_for(extra_edges)), dim3(GZ_BLOCK), 0, gz_stream, d_pairs,
                           extra_edges, b, symmetrize, a + per * p->num_edges);
        GZ_LAUNCH_CHECK();
        GZ_CHECK(hipStreamSynchronize(gz_stream));
        gz_device_free(d_pairs);
    }
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_ingress_from_keys(nv, a, bb, n, b, flags);
    gz_device_free(a);
    gz_device_free(bb);
    return 0;
}

extern "C" int gz_graph_build_rmat_union_device(const gz_rmat_params_t* params, uint32_t count, int symmetrize,
                                                uint64_t* vertex_bases)
{
    gz_runtime_init(0);
    gz_graph_release();
    uint64_t nv = 0, n = 0;
    const uint64_t per = symmetrize ? 2 : 1;
    for (uint32_t g = 0; g < count; ++g) {
        vertex_bases[g] = nv;
        nv += params[g].num_vertices ? params[g].num_vertices : (1ull << params[g].scale);
        n += per * params[g].num_edges;
    }
    vertex_bases[count] = nv;
    gz_check_vertex_count(nv);
    const uint32_t b = gz_id_bits(nv);
    uint64_t* a = (uint64_t*)gz_device_malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t* bb = (uint64_t*)gz_device_malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t at = 0;
    for (uint32_t g = 0; g < count; ++g) {
        const gz_rmat_params_t* p = &params[g];
        if (p->num_edges == 0) continue;
        const uint32_t ta = (uint32_t)(p->a * 65536.0);
        const uint32_t tb = (uint32_t)((p->a + p->b) * 65536.0);
        const uint32_t tc = (uint32_t)((p->a + p->b + p->c) * 65536.0);
        const uint64_t gnv = vertex_bases[g + 1] - vertex_bases[g];
        hipLaunchKernelGGL(gz_k_rmat_keys, dim3(gz_grid_for(p->num_edges)), dim3(GZ_BLOCK), 0, gz_stream, p->seed,
                           p->edge_offset, p->scale, gnv, ta, tb, tc, p->num_edges, b, symmetrize, vertex_bases[g],
                           a + at);
        GZ_LAUNCH_CHECK();
        at += per * p->num_edges;
    }
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_ingress_from_keys(nv, a, bb, n, b, 0);
    gz_device_free(a);
    gz_device_free(bb);
    return 0;
}

/* ---- source-sliced pull layout ---------------------------------------------
 * From the loaded pull vectors: every valid lane becomes the key
 * (slice << 2b | dst << b | src) (invalid lanes all ones), a stable radix pass
 * over the 7 slice bits groups the keys by slice while keeping the (dst, src)
 * order, and each slice's keys are packed by the layout builder above. */
namespace {

__global__ void gz_k_slice_keys(const gz_vec_t* __restrict__ vecs, uint64_t nvec, uint32_t b, uint64_t slice_vertices,
                                uint32_t tiers, uint32_t blocks, uint64_t block_vertices, uint64_t* __restrict__ keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 4 * nvec; i += stride) {
        const gz_vec_t* v = vecs + i / 4;
        const uint64_t l = v->lane[i % 4];
        if (!gz_lane_valid(l)) {
            keys[i] = ~0ull;
            continue;
        }
        gz_lanes w;
        w.l0 = v->lane[0];
        w.l1 = v->lane[1];
        w.l2 = v->lane[2];
        w.l3 = v->lane[3];
        const uint64_t d = gz_shared_vertex(w);
        const uint64_t s = gz_lane_id(l);
        const uint64_t x = gz_slice_of_source(s, slice_vertices, tiers, blocks, block_vertices);
        keys[i] = (x << (2 * b)) | (d << b) | s;
    }
}

/* first[x] = first key whose slice field is >= x, x in [0, nslices] */
__global__ void gz_k_slice_bounds(const uint64_t* __restrict__ keys, uint64_t n, uint32_t b, uint32_t nslices,
                                  uint64_t* __restrict__ first)
{
    const uint64_t x = threadIdx.x;
    if (x > nslices) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if ((keys[mid] >> (2 * b)) < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    first[x] = lo;
}

__global__ void gz_k_strip_slice(uint64_t* __restrict__ keys, uint64_t n, uint32_t b)
{
    const uint64_t mask = (1ull << (2 * b)) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) keys[i] &= mask;
}

/* mask bit for slice `bit`, and has[v] = 1 when v has vectors in the slice
 * (has[nv] = 0, so its exclusive scan ends in the slice's vertex count) */
template <typename M>
__global__ void gz_k_slice_mask(const uint64_t* __restrict__ voff, uint64_t nv, uint32_t bit, M* __restrict__ mask,
                                uint64_t* __restrict__ has)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= nv; v += stride) {
        const bool in = v < nv && voff[v + 1] > voff[v];
        if (in) mask[v] |= (M)1 << bit;
        has[v] = in ? 1 : 0;
    }
}

/* gid[rank[v]] = v for the vertices with vectors in the slice */
__global__ void gz_k_slice_gid(const uint64_t* __restrict__ has, const uint64_t* __restrict__ rank, uint64_t nv,
                               uint32_t* __restrict__ gid)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride)
        if (has[v]) gid[rank[v]] = (uint32_t)v;
}

/* index in the concatenated slice accumulators (slice x's run starts at
 * `part_off`) of the first vertex of each 64-vertex block, at
 * out[block * nslices + x] */
/* 32-bit copy of the FP64 out-degrees (integral, below 2^32: V <= 2^28 here). */
__global__ void gz_k_outdeg32(const double* __restrict__ outdeg, uint64_t n, uint32_t* __restrict__ out)
{
    /* grid-stride: gz_grid_for caps the grid at 65,536 workgroups */
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)outdeg[i];
}

__global__ void gz_k_rank_base(const uint64_t* __restrict__ rank, uint64_t blocks, uint32_t x, uint32_t nslices,
                               uint64_t part_off, uint32_t* __restrict__ out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < blocks; k += stride)
        out[k * nslices + x] = (uint32_t)(part_off + rank[k * GZ_WAVE]);
}

/* the shared vertex of every vector becomes the destination's local rank */
__global__ void gz_k_local_dst(gz_vec_t* __restrict__ vecs, uint64_t n, const uint64_t* __restrict__ rank)
{
    const uint64_t clear = ~(0x7fffull << 48);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        gz_lanes w;
        w.l0 = vecs[j].lane[0];
        w.l1 = vecs[j].lane[1];
        w.l2 = vecs[j].lane[2];
        w.l3 = vecs[j].lane[3];
        const uint64_t r = rank[gz_shared_vertex(w)];
        vecs[j].lane[0] = (w.l0 & clear) | gz_piece(r, 0);
        vecs[j].lane[1] = (w.l1 & clear) | gz_piece(r, 1);
        vecs[j].lane[2] = (w.l2 & clear) | gz_piece(r, 2);
        vecs[j].lane[3] = (w.l3 & clear) | gz_piece(r, 3);
    }
}

/* Vector-Sparse vectors of one slice -> their 32-bit shared vertices, and
 * (nbr != nullptr) 16-byte packed vectors: 32-bit neighbour ids (empty lanes
 * GZ_PACKED_NONE), bit 31 of the first set on the first vector of each
 * destination. */
__global__ void gz_k_pack_u32(const gz_vec_t* __restrict__ vecs, uint64_t n, gz_u32x4* __restrict__ nbr,
                              uint32_t* __restrict__ shared)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        gz_lanes w;
        w.l0 = vecs[j].lane[0];
        w.l1 = vecs[j].lane[1];
        w.l2 = vecs[j].lane[2];
        w.l3 = vecs[j].lane[3];
        const uint64_t d = gz_shared_vertex(w);
        bool first = true;
        if (j > 0) {
            gz_lanes p;
            p.l0 = vecs[j - 1].lane[0];
            p.l1 = vecs[j - 1].lane[1];
            p.l2 = vecs[j - 1].lane[2];
            p.l3 = vecs[j - 1].lane[3];
            first = gz_shared_vertex(p) != d;
        }
        gz_u32x4 a;
        a.x = (uint32_t)gz_lane_id(w.l0) | (first ? 0x80000000u : 0u); /* the first lane holds an edge */
        a.y = gz_lane_valid(w.l1) ? (uint32_t)gz_lane_id(w.l1) : GZ_PACKED_NONE;
        a.z = gz_lane_valid(w.l2) ? (uint32_t)gz_lane_id(w.l2) : GZ_PACKED_NONE;
        a.w = gz_lane_valid(w.l3) ? (uint32_t)gz_lane_id(w.l3) : GZ_PACKED_NONE;
        if (nbr) nbr[j] = a;
        shared[j] = (uint32_t)d;
    }
}

/* Vector-Sparse vectors of one block slice -> 8-byte packed vectors: four
 * 16-bit offsets from the block's first vertex (GZ_PACKED16_NONE for an empty
 * lane), bit 15 of the first set on the first vector of each destination. */
__global__ void gz_k_pack_u16(const gz_vec_t* __restrict__ vecs, uint64_t n, uint64_t window,
                              uint64_t* __restrict__ words)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        gz_lanes w;
        w.l0 = vecs[j].lane[0];
        w.l1 = vecs[j].lane[1];
        w.l2 = vecs[j].lane[2];
        w.l3 = vecs[j].lane[3];
        bool first = true;
        if (j > 0) {
            gz_lanes p;
            p.l0 = vecs[j - 1].lane[0];
            p.l1 = vecs[j - 1].lane[1];
            p.l2 = vecs[j - 1].lane[2];
            p.l3 = vecs[j - 1].lane[3];
            first = gz_shared_vertex(p) != gz_shared_vertex(w);
        }
        const uint64_t f0 = (gz_lane_id(w.l0) - window) | (first ? GZ_PACKED16_FIRST : 0u);
        const uint64_t f1 = gz_lane_valid(w.l1) ? gz_lane_id(w.l1) - window : GZ_PACKED16_NONE;
        const uint64_t f2 = gz_lane_valid(w.l2) ? gz_lane_id(w.l2) - window : GZ_PACKED16_NONE;
        const uint64_t f3 = gz_lane_valid(w.l3) ? gz_lane_id(w.l3) - window : GZ_PACKED16_NONE;
        words[j] = f0 | (f1 << 16) | (f2 << 32) | (f3 << 48);
    }
}

/* Vector-Sparse vectors of one window slice -> 12-byte packed vectors: four
 * 24-bit offsets from the window's first vertex (empty lanes all ones), bit
 * 23 of the first set on the first vector of each destination. */
__global__ void gz_k_pack_u24(const gz_vec_t* __restrict__ vecs, uint64_t n, uint64_t window,
                              uint32_t* __restrict__ words)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        gz_lanes w;
        w.l0 = vecs[j].lane[0];
        w.l1 = vecs[j].lane[1];
        w.l2 = vecs[j].lane[2];
        w.l3 = vecs[j].lane[3];
        bool first = true;
        if (j > 0) {
            gz_lanes p;
            p.l0 = vecs[j - 1].lane[0];
            p.l1 = vecs[j - 1].lane[1];
            p.l2 = vecs[j - 1].lane[2];
            p.l3 = vecs[j - 1].lane[3];
            first = gz_shared_vertex(p) != gz_shared_vertex(w);
        }
        const uint32_t f0 = (uint32_t)(gz_lane_id(w.l0) - window) | (first ? 0x800000u : 0u);
        const uint32_t f1 = gz_lane_valid(w.l1) ? (uint32_t)(gz_lane_id(w.l1) - window) : 0xffffffu;
        const uint32_t f2 = gz_lane_valid(w.l2) ? (uint32_t)(gz_lane_id(w.l2) - window) : 0xffffffu;
        const uint32_t f3 = gz_lane_valid(w.l3) ? (uint32_t)(gz_lane_id(w.l3) - window) : 0xffffffu;
        words[3 * j] = f0 | (f1 << 24);
        words[3 * j + 1] = (f1 >> 8) | (f2 << 16);
        words[3 * j + 2] = (f2 >> 16) | (f3 << 8);
    }
}

}  // namespace

extern "C" int gz_graph_build_source_slices(uint64_t slice_vertices, uint32_t tiers, uint32_t blocks,
                                            uint64_t block_vertices)
{
    gz_runtime_init(0);
    if (gz_graph.pull_vectors == nullptr) {
        fprintf(stderr, "grazelle: the source-sliced layout needs a pull layout\n");
        exit(65);
    }
    const uint64_t nv = gz_graph.num_vertices, nvec = gz_graph.pull_vector_count;
    const uint32_t b = gz_id_bits(nv);
    if (2 * b + GZ_SLICE_KEY_BITS > 64) {
        fprintf(stderr, "grazelle: the source-sliced layout needs V <= 2^28 (V = %llu)\n", (unsigned long long)nv);
        exit(65);
    }
    if (tiers < 1 || tiers > GZ_MAX_TIERS || blocks + tiers > GZ_MAX_LEVELS) {
        fprintf(stderr, "grazelle: source-sliced tiers must be 1..%d and tiers + blocks at most %d (got %u, %u)\n",
                GZ_MAX_TIERS, GZ_MAX_LEVELS, tiers, blocks);
        exit(65);
    }
    gz_source_slices_release();
    const uint64_t S = gz_slices_effective(slice_vertices, tiers, nv);
    /* as many of the requested blocks as fit in a window after its block 0 */
    const uint64_t H = blocks ? block_vertices : 0;
    if (H > GZ_PACKED16_NONE) {
        fprintf(stderr, "grazelle: LDS blocks hold at most %u sources (got %llu)\n", GZ_PACKED16_NONE,
                (unsigned long long)H);
        exit(65);
    }
    if (blocks && (H == 0 || S / H < 1))
        blocks = 0;
    else if (blocks && S / H - 1 < blocks)
        blocks = (uint32_t)(S / H - 1);
    const uint32_t ns = GZ_XCDS * (tiers + blocks) + 1;
    const bool wide_mask = ns > 32;
    const uint64_t n = 4 * nvec;
    uint64_t* a = (uint64_t*)gz_device_malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t* bb = (uint64_t*)gz_device_malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t* first = (uint64_t*)gz_device_malloc((GZ_MAX_SLICES + 1) * sizeof(uint64_t));
    uint64_t lo[GZ_MAX_SLICES + 1];
    uint64_t* sorted = a;
    if (n) {
        hipLaunchKernelGGL(gz_k_slice_keys, dim3(gz_grid_for(n)), dim3(GZ_BLOCK), 0, gz_stream, gz_graph.pull_vectors,
                           nvec, b, S, tiers, blocks, H, a);
        GZ_LAUNCH_CHECK();
        rocprim::double_buffer<uint64_t> db(a, bb);
        size_t tmp_bytes = 0;
        GZ_CHECK(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, (size_t)n, 2 * b, 2 * b + GZ_SLICE_KEY_BITS,
                                          gz_stream));
        void* tmp = gz_device_malloc(tmp_bytes ? tmp_bytes : 1);
        GZ_CHECK(rocprim::radix_sort_keys(tmp, tmp_bytes, db, (size_t)n, 2 * b, 2 * b + GZ_SLICE_KEY_BITS,
                                          gz_stream));
        GZ_CHECK(hipStreamSynchronize(gz_stream));
        gz_device_free(tmp);
        sorted = db.current();
        hipLaunchKernelGGL(gz_k_slice_bounds, dim3(1), dim3(64), 0, gz_stream, sorted, n, b, ns, first);
        GZ_LAUNCH_CHECK();
        gz_copy_to_host(lo, first, (ns + 1) * sizeof(uint64_t));
        hipLaunchKernelGGL(gz_k_strip_slice, dim3(gz_grid_for(lo[ns])), dim3(GZ_BLOCK), 0, gz_stream, sorted, lo[ns],
                           b);
        GZ_LAUNCH_CHECK();
    } else {
        for (uint32_t x = 0; x <= ns; ++x) lo[x] = 0;
    }
    gz_slices.mask = gz_device_malloc((nv + 8) * (wide_mask ? sizeof(uint64_t) : sizeof(uint32_t)));
    gz_slices.rank_blocks = (nv + GZ_WAVE - 1) / GZ_WAVE;
    gz_slices.rank_base = (uint32_t*)gz_device_malloc(ns * gz_slices.rank_blocks * sizeof(uint32_t));
    uint64_t* eoff = (uint64_t*)gz_device_malloc((nv + 1) * sizeof(uint64_t));
    uint64_t* has = (uint64_t*)gz_device_malloc((nv + 1) * sizeof(uint64_t));
    uint64_t* rank = (uint64_t*)gz_device_malloc((nv + 1) * sizeof(uint64_t));
    size_t scan_bytes = 0;
    GZ_CHECK(rocprim::exclusive_scan(nullptr, scan_bytes, has, rank, (uint64_t)0, (size_t)(nv + 1),
                                     rocprim::plus<uint64_t>(), gz_stream));
    void* scan_tmp = gz_device_malloc(scan_bytes ? scan_bytes : 1);
    gz_vec_t* part_vecs[GZ_MAX_SLICES];
    uint32_t* part_gid[GZ_MAX_SLICES];
    uint64_t counts[GZ_MAX_SLICES], ranked[GZ_MAX_SLICES];
    uint64_t part_off = 0;
    for (uint32_t x = 0; x < ns; ++x) {
        uint64_t* voff = nullptr;
        gz_build_layout(sorted + lo[x], lo[x + 1] - lo[x], b, nv, &part_vecs[x], &voff, &counts[x], eoff);
        if (wide_mask)
            hipLaunchKernelGGL(gz_k_slice_mask<uint64_t>, dim3(gz_grid_for(nv + 1)), dim3(GZ_BLOCK), 0, gz_stream,
                               voff, nv, x, (uint64_t*)gz_slices.mask, has);
        else
            hipLaunchKernelGGL(gz_k_slice_mask<uint32_t>, dim3(gz_grid_for(nv + 1)), dim3(GZ_BLOCK), 0, gz_stream,
                               voff, nv, x, (uint32_t*)gz_slices.mask, has);
        GZ_LAUNCH_CHECK();
        GZ_CHECK(rocprim::exclusive_scan(scan_tmp, scan_bytes, has, rank, (uint64_t)0, (size_t)(nv + 1),
                                         rocprim::plus<uint64_t>(), gz_stream));
        hipLaunchKernelGGL(gz_k_rank_base, dim3(gz_grid_for(gz_slices.rank_blocks)), dim3(GZ_BLOCK), 0, gz_stream,
                           rank, gz_slices.rank_blocks, x, ns, part_off, gz_slices.rank_base);
        GZ_LAUNCH_CHECK();
        if (counts[x]) {
            hipLaunchKernelGGL(gz_k_local_dst, dim3(gz_grid_for(counts[x])), dim3(GZ_BLOCK), 0, gz_stream,
                               part_vecs[x], counts[x], rank);
            GZ_LAUNCH_CHECK();
        }
        gz_copy_to_host(&ranked[x], rank + nv, sizeof(uint64_t));
        part_off += ranked[x];
        if (part_off + 8 > 0xffffffffull) {
            fprintf(stderr, "grazelle: the sliced layout needs fewer than 2^32 slice accumulators\n");
            exit(65);
        }
        part_gid[x] = (uint32_t*)gz_device_malloc((ranked[x] ? ranked[x] : 1) * sizeof(uint32_t));
        hipLaunchKernelGGL(gz_k_slice_gid, dim3(gz_grid_for(nv)), dim3(GZ_BLOCK), 0, gz_stream, has, rank, nv,
                           part_gid[x]);
        GZ_LAUNCH_CHECK();
        GZ_CHECK(hipStreamSynchronize(gz_stream));
        gz_device_free(voff);
    }
    gz_device_free(scan_tmp);
    gz_device_free(rank);
    gz_device_free(has);
    gz_device_free(eoff);
    gz_device_free(first);
    gz_device_free(a);
    gz_device_free(bb);
    gz_slices.begin[0] = 0;
    for (uint32_t x = 0; x < ns; ++x) gz_slices.begin[x + 1] = gz_slices.begin[x] + counts[x];
    const uint64_t total = gz_slices.begin[ns];
    const uint64_t block_first = gz_slices.begin[GZ_XCDS];
    const uint64_t block_span = gz_slices.begin[GZ_XCDS * (blocks + 1)] - block_first;
    const uint64_t window_total = gz_slices.begin[ns - 1] - block_span;
    gz_slices.nbr24 = (uint32_t*)gz_device_malloc((window_total ? 3 * window_total : 1) * sizeof(uint32_t));
    gz_slices.nbr16 = (uint64_t*)gz_device_malloc((block_span ? block_span : 1) * sizeof(uint64_t));
    gz_slices.block_span = block_span;
    gz_slices.nbr = (gz_u32x4*)gz_device_malloc((counts[ns - 1] ? counts[ns - 1] : 1) * sizeof(gz_u32x4));
    gz_slices.shared = (uint32_t*)gz_device_malloc((total ? total : 1) * sizeof(uint32_t));
    for (uint32_t x = 0; x < ns; ++x) {
        if (counts[x]) {
            /* the shared vertices of every slice, then the packed neighbours */
            hipLaunchKernelGGL(gz_k_pack_u32, dim3(gz_grid_for(counts[x])), dim3(GZ_BLOCK), 0, gz_stream, part_vecs[x],
                               counts[x], x + 1 == ns ? gz_slices.nbr : (gz_u32x4*)nullptr,
                               gz_slices.shared + gz_slices.begin[x]);
            GZ_LAUNCH_CHECK();
            const uint32_t level = x / GZ_XCDS;
            if (x + 1 < ns && level >= 1 && level <= blocks) {
                hipLaunchKernelGGL(gz_k_pack_u16, dim3(gz_grid_for(counts[x])), dim3(GZ_BLOCK), 0, gz_stream,
                                   part_vecs[x], counts[x], gz_slice_window(x, S, blocks) + (uint64_t)level * H,
                                   gz_slices.nbr16 + (gz_slices.begin[x] - block_first));
                GZ_LAUNCH_CHECK();
            } else if (x + 1 < ns) {
                const uint64_t at = gz_slices.begin[x] - (level > blocks ? block_span : 0);
                hipLaunchKernelGGL(gz_k_pack_u24, dim3(gz_grid_for(counts[x])), dim3(GZ_BLOCK), 0, gz_stream,
                                   part_vecs[x], counts[x], gz_slice_window(x, S, blocks), gz_slices.nbr24 + 3 * at);
                GZ_LAUNCH_CHECK();
            }
        }
        GZ_CHECK(hipStreamSynchronize(gz_stream));
        gz_device_free(part_vecs[x]);
    }
    memset(&gz_slices.part_base, 0, sizeof(gz_slices.part_base));
    for (uint32_t x = 0; x < ns; ++x) gz_slices.part_base.off[x + 1] = gz_slices.part_base.off[x] + ranked[x];
    gz_slices.parts = (double*)gz_device_malloc((gz_slices.part_base.off[ns] + 8) * sizeof(double));
    gz_slices.gid = (uint32_t*)gz_device_malloc((gz_slices.part_base.off[ns] + 8) * sizeof(uint32_t));
    for (uint32_t x = 0; x < ns; ++x) {
        if (ranked[x]) gz_device_copy(gz_slices.gid + gz_slices.part_base.off[x], part_gid[x], ranked[x] * sizeof(uint32_t));
        GZ_CHECK(hipStreamSynchronize(gz_stream));
        gz_device_free(part_gid[x]);
    }
    gz_slices.outdeg32 = (uint32_t*)gz_device_malloc((nv + 8) * sizeof(uint32_t));
    hipLaunchKernelGGL(gz_k_outdeg32, dim3(gz_grid_for(nv)), dim3(GZ_BLOCK), 0, gz_stream, gz_graph.vertex_outdegrees,
                       nv, gz_slices.outdeg32);
    GZ_LAUNCH_CHECK();
    gz_slices.base_list = gz_graph.pull_vectors;
    gz_slices.base_count = nvec;
    gz_slices.slice_vertices = S;
    gz_slices.tiers = tiers;
    gz_slices.blocks = blocks;
    gz_slices.block_vertices = H;
    gz_slices.nslices = ns;
    return 0;
}
