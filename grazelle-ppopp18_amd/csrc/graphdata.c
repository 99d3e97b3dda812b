/*
 * graphdata.c -- host ingress of the binary -pull / -push edge lists into the
 * Vector-Sparse layout, upload to device memory, and the -o output.
 *
 * Follows source/graphdata.c:
 *   file header / body                  :533-591
 *   stash-and-flush packing              :618-701 (pull), :705-788 (push)
 *   vector composition                   :143-173
 *   FP64 outdegrees from the pull file   :792-821
 *   vertex index                         :219-267
 *   vertex info / frontier allocation    :269-319
 *   merge buffers                        :1053-1072
 *   rank output                          :1119-1138
 * Differences: one host thread builds both lists (the reference runs a
 * producer and two consumers, :856-951); invalid lanes carry neighbour id 0
 * instead of stale stash ids (no consumer reads them); an empty edge list
 * yields zero vectors instead of one all-invalid vector.  NUMA placement has
 * no meaning on one GPU: all arrays live in HBM.
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grazelle_host.h"
#include "grazelle_phases.h"
#include "gz_host_internal.h"

/* ---- edge-list files ------------------------------------------------------ */

int gz_edge_list_write(const char* path, uint64_t num_vertices, uint64_t num_edges, const uint64_t* pairs)
{
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    const uint64_t hdr[2] = {num_vertices, num_edges};
    int ok = fwrite(hdr, sizeof(uint64_t), 2, f) == 2;
    if (ok && num_edges) ok = fwrite(pairs, sizeof(uint64_t), 2 * num_edges, f) == 2 * num_edges;
    if (fclose(f) != 0) ok = 0;
    return ok ? 0 : -4;
}

int gz_edge_list_read(const char* path, uint64_t* num_vertices, uint64_t* num_edges, uint64_t** pairs)
{
    *pairs = NULL;
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    uint64_t hdr[2];
    if (fread(hdr, sizeof(uint64_t), 2, f) != 2) {
        fclose(f);
        return -2;
    }
    *num_vertices = hdr[0];
    *num_edges = hdr[1];
    uint64_t* p = (uint64_t*)malloc((hdr[1] ? 2 * hdr[1] : 2) * sizeof(uint64_t));
    if (!p) {
        fclose(f);
        return -5;
    }
    if (hdr[1] && fread(p, sizeof(uint64_t), 2 * hdr[1], f) != 2 * hdr[1]) {
        free(p);
        fclose(f);
        return -3;
    }
    fclose(f);
    *pairs = p;
    return 0;
}

/* ---- sorting ---------------------------------------------------------------
 * LSD radix sort of 64-bit composite keys (primary << 32 | secondary) when
 * both ids fit in 32 bits, else a comparison sort on the pairs. */

static void gz_radix_sort_u64(uint64_t* keys, uint64_t n, unsigned key_bits)
{
    if (n < 2) return;
    uint64_t* tmp = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* count = (uint64_t*)malloc(65536 * sizeof(uint64_t));
    uint64_t* src = keys;
    uint64_t* dst = tmp;
    for (unsigned shift = 0; shift < key_bits; shift += 16) {
        memset(count, 0, 65536 * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; ++i) count[(src[i] >> shift) & 0xffff] += 1;
        uint64_t run = 0;
        for (unsigned d = 0; d < 65536; ++d) {
            const uint64_t c = count[d];
            count[d] = run;
            run += c;
        }
        for (uint64_t i = 0; i < n; ++i) dst[count[(src[i] >> shift) & 0xffff]++] = src[i];
        uint64_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != keys) memcpy(keys, src, n * sizeof(uint64_t));
    free(count);
    free(tmp);
}

static int gz_cmp_pair_src(const void* a, const void* b)
{
    const uint64_t* x = (const uint64_t*)a;
    const uint64_t* y = (const uint64_t*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

static int gz_cmp_pair_dst(const void* a, const void* b)
{
    const uint64_t* x = (const uint64_t*)a;
    const uint64_t* y = (const uint64_t*)b;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    return 0;
}

uint64_t gz_edges_sort(uint64_t* pairs, uint64_t num_edges, int by_dst, int dedupe)
{
    uint64_t max_id = 0;
    for (uint64_t i = 0; i < 2 * num_edges; ++i)
        if (pairs[i] > max_id) max_id = pairs[i];
    if (max_id <= 0xffffffffull) {
        unsigned bits = 0;
        while (bits < 32 && (max_id >> bits)) ++bits;
        uint64_t* keys = (uint64_t*)malloc((num_edges ? num_edges : 1) * sizeof(uint64_t));
        for (uint64_t i = 0; i < num_edges; ++i) {
            const uint64_t s = pairs[2 * i], d = pairs[2 * i + 1];
            keys[i] = by_dst ? ((d << 32) | s) : ((s << 32) | d);
        }
        gz_radix_sort_u64(keys, num_edges, 32 + ((bits + 15) / 16) * 16);
        for (uint64_t i = 0; i < num_edges; ++i) {
            const uint64_t hi = keys[i] >> 32, lo = keys[i] & 0xffffffffull;
            pairs[2 * i] = by_dst ? lo : hi;
            pairs[2 * i + 1] = by_dst ? hi : lo;
        }
        free(keys);
    } else {
        qsort(pairs, num_edges, 2 * sizeof(uint64_t), by_dst ? gz_cmp_pair_dst : gz_cmp_pair_src);
    }
    if (!dedupe) return num_edges;
    uint64_t w = 0;
    for (uint64_t i = 0; i < num_edges; ++i) {
        const uint64_t s = pairs[2 * i], d = pairs[2 * i + 1];
        if (s == d) continue;
        if (w > 0 && pairs[2 * (w - 1)] == s && pairs[2 * (w - 1) + 1] == d) continue;
        pairs[2 * w] = s;
        pairs[2 * w + 1] = d;
        ++w;
    }
    return w;
}

/* ---- Vector-Sparse packing ------------------------------------------------ */

uint64_t gz_pack_vectors(const uint64_t* pairs, uint64_t num_edges, int group_by_dst, gz_vec_t* out)
{
    uint64_t stash[4] = {0, 0, 0, 0};
    uint64_t stash_shared = 0;
    unsigned stash_count = 0;
    uint64_t nv = 0;
    for (uint64_t i = 0; i < num_edges; ++i) {
        const uint64_t s = pairs[2 * i], d = pairs[2 * i + 1];
        const uint64_t shared = group_by_dst ? d : s;
        const uint64_t other = group_by_dst ? s : d;
        /* flush rule of graphdata.c:679-686 / :766-773 */
        if ((stash_count != 0 && stash_shared != shared) || stash_count == 4) {
            gz_vec_encode(&out[nv++], stash_shared, stash, stash_count);
            stash_count = 0;
        }
        stash_shared = shared;
        stash[stash_count++] = other;
    }
    if (stash_count) gz_vec_encode(&out[nv++], stash_shared, stash, stash_count);
    return nv;
}

uint64_t gz_count_vectors(const uint64_t* pairs, uint64_t num_edges, int group_by_dst)
{
    uint64_t nv = 0, run = 0, prev = 0;
    for (uint64_t i = 0; i < num_edges; ++i) {
        const uint64_t shared = group_by_dst ? pairs[2 * i + 1] : pairs[2 * i];
        if (run != 0 && (shared != prev || run == 4)) {
            nv += 1;
            run = 0;
        }
        prev = shared;
        run += 1;
    }
    return nv + (run ? 1 : 0);
}

void gz_build_vertex_index(const gz_vec_t* vectors, uint64_t vector_count, uint64_t num_vertices, uint64_t* index,
                           uint64_t* index_start, uint64_t* index_end)
{
    const uint64_t count = num_vertices + 64;
    if (vector_count == 0) {
        for (uint64_t i = 0; i < count; ++i) index[i] = GZ_VINDEX_PAST_END;
        *index_start = 0;
        *index_end = 0;
        return;
    }
    uint64_t cur = gz_vec_shared_vertex(&vectors[0]);
    for (uint64_t i = 0; i < cur; ++i) index[i] = GZ_VINDEX_NO_EDGES;
    index[cur] = 0;
    uint64_t last = cur;
    *index_start = cur;
    for (uint64_t i = 1; i < vector_count; ++i) {
        cur = gz_vec_shared_vertex(&vectors[i]);
        if (cur != last) {
            while (last + 1 < cur) index[++last] = GZ_VINDEX_NO_EDGES;
            index[cur] = i;
            last = cur;
        }
    }
    *index_end = cur;
    for (uint64_t i = cur + 1; i < count; ++i) index[i] = GZ_VINDEX_PAST_END;
}

void gz_build_vertex_offsets(const gz_vec_t* vectors, uint64_t vector_count, uint64_t num_vertices, uint64_t* offsets)
{
    /* offsets[v] = index of the first vector whose shared vertex is >= v */
    uint64_t v = 0;
    for (uint64_t i = 0; i < vector_count; ++i) {
        const uint64_t s = gz_vec_shared_vertex(&vectors[i]);
        while (v <= s && v <= num_vertices) offsets[v++] = i;
    }
    while (v <= num_vertices) offsets[v++] = vector_count;
}

void gz_count_outdegrees(const uint64_t* pairs, uint64_t num_edges, uint64_t num_vertices, double* outdeg)
{
    memset(outdeg, 0, (num_vertices + 8) * sizeof(double));
    for (uint64_t i = 0; i < num_edges; ++i) outdeg[pairs[2 * i]] += 1.0;
}

/* ---- device graph --------------------------------------------------------- */

typedef struct gz_host_graph_t {
    uint64_t num_vertices, num_edges;
    uint64_t pull_vectors, push_vectors;
    double* outdeg;              /* host mirror */
    /* device allocations */
    gz_vec_t* d_pull;
    gz_vec_t* d_push;
    uint64_t* d_pull_offsets;
    uint64_t* d_push_offsets;
    double* d_outdeg;
    void* d_props;
    void* d_acc;
    uint64_t* d_has;
    uint64_t* d_wants;
    /* renumbered graphs only (else NULL) */
    uint64_t* d_original_ids;    /* original id by device vertex */
    uint64_t* original_ids;      /* host mirror of the above */
    uint64_t* device_ids;        /* device vertex by original id */
    uint32_t* d_indeg;           /* in-degrees, built on first use */
    uint32_t* h_indeg;           /* host copy of d_indeg, made on first use */
} gz_host_graph_t;

static gz_host_graph_t G;

uint64_t gz_graph_num_vertices(void) { return G.num_vertices; }

/* Give the frontier and accumulator arrays back their load-time roles.  The
 * CC/BFS drivers swap gz_graph.frontier_has_info and
 * gz_graph.vertex_accumulators every iteration; the accumulator array holds
 * V+8 words and the bitmaps only ceil(V/64)+8, so every run must start from
 * the original assignment. */
void gz_graph_reset_roles(void)
{
    gz_graph.vertex_accumulators = G.d_acc;
    gz_graph.frontier_has_info = G.d_has;
    gz_graph.frontier_wants_info = G.d_wants;
}
uint64_t gz_graph_num_edges(void) { return G.num_edges; }
uint64_t gz_graph_pull_vector_count(void) { return G.pull_vectors; }
uint64_t gz_graph_push_vector_count(void) { return G.push_vectors; }
const double* gz_graph_outdegrees_host(void) { return G.outdeg; }

void gz_graph_release(void)
{
    gz_source_slices_release();
    gz_frontier_counts_release();
    gz_pr_release();
    gz_bfs_batch_release();
    gz_bfs_pending_release();
    gz_device_free(G.d_pull);
    gz_device_free(G.d_push);
    gz_device_free(G.d_pull_offsets);
    gz_device_free(G.d_push_offsets);
    gz_device_free(G.d_outdeg);
    gz_device_free(G.d_props);
    gz_device_free(G.d_acc);
    gz_device_free(G.d_has);
    gz_device_free(G.d_wants);
    gz_device_free(G.d_original_ids);
    gz_device_free(G.d_indeg);
    free(G.h_indeg);
    free(G.original_ids);
    free(G.device_ids);
    free(G.outdeg);
    memset(&G, 0, sizeof(G));
    memset(&gz_graph, 0, sizeof(gz_graph));
}

static void gz_graph_finish(void);

static void gz_upload_layout(const uint64_t* pairs, uint64_t num_edges, uint64_t num_vertices, int by_dst,
                             gz_vec_t** d_vecs, uint64_t** d_offsets, uint64_t* count)
{
    gz_vec_t* vecs = (gz_vec_t*)aligned_alloc(32, (num_edges ? num_edges : 1) * sizeof(gz_vec_t));
    const uint64_t nv = gz_pack_vectors(pairs, num_edges, by_dst, vecs);
    uint64_t* offs = (uint64_t*)malloc((num_vertices + 1) * sizeof(uint64_t));
    gz_build_vertex_offsets(vecs, nv, num_vertices, offs);
    *d_vecs = (gz_vec_t*)gz_device_malloc((nv ? nv : 1) * sizeof(gz_vec_t));
    gz_copy_to_device(*d_vecs, vecs, nv * sizeof(gz_vec_t));
    *d_offsets = (uint64_t*)gz_device_malloc((num_vertices + 1) * sizeof(uint64_t));
    gz_copy_to_device(*d_offsets, offs, (num_vertices + 1) * sizeof(uint64_t));
    *count = nv;
    free(offs);
    free(vecs);
}

int gz_graph_load_pairs(uint64_t num_vertices, uint64_t pull_edges, const uint64_t* pull_pairs, uint64_t push_edges,
                        const uint64_t* push_pairs)
{
    gz_runtime_init(0);
    gz_graph_release();
    if (num_vertices == 0 || num_vertices > GZ_MAX_VERTEX_ID) {
        fprintf(stderr, "grazelle: vertex count %llu is outside [1, 2^48)\n", (unsigned long long)num_vertices);
        exit(65);
    }
    G.num_vertices = num_vertices;
    G.num_edges = pull_pairs ? pull_edges : push_edges;
    G.outdeg = (double*)malloc((num_vertices + 8) * sizeof(double));
    if (pull_pairs)
        gz_count_outdegrees(pull_pairs, pull_edges, num_vertices, G.outdeg);
    else
        gz_count_outdegrees(push_pairs, push_edges, num_vertices, G.outdeg);
    if (pull_pairs) gz_upload_layout(pull_pairs, pull_edges, num_vertices, 1, &G.d_pull, &G.d_pull_offsets, &G.pull_vectors);
    if (push_pairs) gz_upload_layout(push_pairs, push_edges, num_vertices, 0, &G.d_push, &G.d_push_offsets, &G.push_vectors);
    const size_t vbytes = (num_vertices + 8) * sizeof(uint64_t);
    G.d_outdeg = (double*)gz_device_malloc(vbytes);
    gz_copy_to_device(G.d_outdeg, G.outdeg, vbytes);
    gz_graph_finish();
    return 0;
}

void gz_graph_adopt_device(const gz_device_layouts_t* L)
{
    G.num_vertices = L->num_vertices;
    G.num_edges = L->num_edges;
    G.d_pull = L->pull;
    G.d_pull_offsets = L->pull_offsets;
    G.pull_vectors = L->pull_count;
    G.d_push = L->push;
    G.d_push_offsets = L->push_offsets;
    G.push_vectors = L->push_count;
    G.d_outdeg = L->outdeg;
    G.outdeg = (double*)malloc((L->num_vertices + 8) * sizeof(double));
    gz_copy_to_host(G.outdeg, G.d_outdeg, (L->num_vertices + 8) * sizeof(double));
    if (L->original_ids != NULL) {
        const uint64_t n = L->num_vertices;
        G.d_original_ids = L->original_ids;
        G.original_ids = (uint64_t*)malloc(n * sizeof(uint64_t));
        G.device_ids = (uint64_t*)malloc(n * sizeof(uint64_t));
        gz_copy_to_host(G.original_ids, G.d_original_ids, n * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; ++i) G.device_ids[G.original_ids[i]] = i;
    }
    gz_graph_finish();
}

const uint64_t* gz_graph_original_ids_device(void) { return G.d_original_ids; }

const uint32_t* gz_graph_indegrees_device(void)
{
    if (G.d_indeg == NULL && G.num_vertices) {
        G.d_indeg = (uint32_t*)gz_device_malloc((G.num_vertices + 8) * sizeof(uint32_t));
        if (G.d_pull != NULL)
            gz_dev_indegrees(G.d_pull, G.pull_vectors, 1, G.d_indeg, G.num_vertices);
        else
            gz_dev_indegrees(G.d_push, G.push_vectors, 0, G.d_indeg, G.num_vertices);
    }
    return G.d_indeg;
}

uint32_t gz_graph_indegree_host(uint64_t v)
{
    if (G.h_indeg == NULL && G.num_vertices) {
        const uint32_t* d = gz_graph_indegrees_device();
        G.h_indeg = (uint32_t*)malloc(G.num_vertices * sizeof(uint32_t));
        gz_copy_to_host(G.h_indeg, d, G.num_vertices * sizeof(uint32_t));
    }
    return (v < G.num_vertices) ? G.h_indeg[v] : 0u;
}

void gz_graph_indegrees(uint32_t* out)
{
    if (G.num_vertices == 0) return;
    gz_copy_to_host(out, gz_graph_indegrees_device(), G.num_vertices * sizeof(uint32_t));
}

uint64_t gz_graph_device_vertex(uint64_t original)
{
    if (G.device_ids == NULL || original >= G.num_vertices) return original;
    return G.device_ids[original];
}

void gz_graph_original_ids(uint64_t* out)
{
    for (uint64_t v = 0; v < G.num_vertices; ++v) out[v] = G.original_ids != NULL ? G.original_ids[v] : v;
}

/* out[original id] = in[device vertex] for V words (identity copy when the
 * graph is not renumbered). */
static void gz_to_original_order(uint64_t* out, const uint64_t* in)
{
    if (G.device_ids == NULL) {
        if (out != in) memcpy(out, in, G.num_vertices * sizeof(uint64_t));
        return;
    }
    for (uint64_t v = 0; v < G.num_vertices; ++v) out[v] = in[G.device_ids[v]];
}

/* Allocate the vertex arrays and frontiers and publish the graph record. */
static void gz_graph_finish(void)
{
    const uint64_t num_vertices = G.num_vertices;
    /* the compaction's per-vertex vector counts are keyed by the offsets'
     * address; a new graph may reuse an address, so they are rebuilt */
    gz_frontier_counts_release();
    const size_t vbytes = (num_vertices + 8) * sizeof(uint64_t);
    const size_t fwords = (num_vertices + 63) / 64 + 8;
    G.d_props = gz_device_malloc(vbytes);
    G.d_acc = gz_device_malloc(vbytes);
    G.d_has = (uint64_t*)gz_device_malloc(fwords * sizeof(uint64_t));
    G.d_wants = (uint64_t*)gz_device_malloc(fwords * sizeof(uint64_t));

    gz_graph.num_vertices = num_vertices;
    gz_graph.num_edges = G.num_edges;
    gz_graph.pull_vectors = G.d_pull;
    gz_graph.pull_vector_count = G.pull_vectors;
    gz_graph.pull_vertex_offsets = G.d_pull_offsets;
    gz_graph.push_vectors = G.d_push;
    gz_graph.push_vector_count = G.push_vectors;
    gz_graph.push_vertex_offsets = G.d_push_offsets;
    gz_graph.vertex_outdegrees = G.d_outdeg;
    gz_graph.vertex_props = G.d_props;
    gz_graph.vertex_accumulators = G.d_acc;
    gz_graph.frontier_has_info = G.d_has;
    gz_graph.frontier_wants_info = G.d_wants;
}

void graph_data_read_from_file(const char* filename_gather, const char* filename_scatter, uint32_t num_numa_nodes,
                               const uint32_t* numa_nodes)
{
    (void)num_numa_nodes;
    (void)numa_nodes;
    uint64_t v1 = 0, e1 = 0, v2 = 0, e2 = 0;
    uint64_t *p1 = NULL, *p2 = NULL;
    if (gz_edge_list_read(filename_gather, &v1, &e1, &p1) != 0) {
        fprintf(stderr, "Error: unable to read file \"%s\"\n", filename_gather);
        exit(255);
    }
    if (filename_scatter != NULL && gz_edge_list_read(filename_scatter, &v2, &e2, &p2) != 0) {
        fprintf(stderr, "Error: unable to read file \"%s\"\n", filename_scatter);
        exit(255);
    }
    gz_graph_load_pairs(v1, e1, p1, e2, p2);
    printf("In-edges:  created %llu vectors, efficiency = %.1lf%%\n", (unsigned long long)G.pull_vectors,
           G.pull_vectors ? (double)G.num_edges / (double)G.pull_vectors / 4.0 * 100.0 : 0.0);
    if (p2)
        printf("Out-edges: created %llu vectors, efficiency = %.1lf%%\n", (unsigned long long)G.push_vectors,
               G.push_vectors ? (double)e2 / (double)G.push_vectors / 4.0 * 100.0 : 0.0);
    free(p1);
    free(p2);
}

void graph_data_allocate_merge_buffers(uint64_t num_threads, uint32_t num_numa_nodes, const uint32_t* numa_nodes)
{
    (void)num_threads;
    (void)num_numa_nodes;
    (void)numa_nodes;
    gz_merge_buffer_allocate(gz_pull_units_for(G.pull_vectors));
}

void gz_graph_download_vectors(int pull, gz_vec_t* out)
{
    if (pull)
        gz_copy_to_host(out, G.d_pull, G.pull_vectors * sizeof(gz_vec_t));
    else
        gz_copy_to_host(out, G.d_push, G.push_vectors * sizeof(gz_vec_t));
}

void gz_graph_download_offsets(int pull, uint64_t* out)
{
    gz_copy_to_host(out, pull ? G.d_pull_offsets : G.d_push_offsets, (G.num_vertices + 1) * sizeof(uint64_t));
}

/* ---- results -------------------------------------------------------------- */

/* Device vertex words in original-id order (a fresh malloc'd array). */
static uint64_t* gz_download_vertex_words(void)
{
    const uint64_t n = G.num_vertices;
    uint64_t* w = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    gz_copy_to_host(w, gz_graph.vertex_props, n * sizeof(uint64_t));
    if (G.device_ids != NULL) {
        uint64_t* o = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
        gz_to_original_order(o, w);
        free(w);
        w = o;
    }
    return w;
}

void gz_result_pr_props(double* out)
{
    uint64_t* w = gz_download_vertex_words();
    memcpy(out, w, G.num_vertices * sizeof(double));
    free(w);
}

void gz_result_pr_ranks(double* out)
{
    gz_result_pr_props(out);
    const double nv = (double)G.num_vertices;
    for (uint64_t i = 0; i < G.num_vertices; ++i) {
        const double od = G.outdeg[gz_graph_device_vertex(i)];
        out[i] = out[i] * (0.0 == od ? nv : od);
    }
}

void gz_result_cc_labels(uint64_t* out)
{
    uint64_t* w = gz_download_vertex_words();
    memcpy(out, w, G.num_vertices * sizeof(uint64_t));
    free(w);
}

void gz_result_bfs(int64_t* parents, uint64_t* depths)
{
    uint64_t* w = gz_download_vertex_words();
    for (uint64_t i = 0; i < G.num_vertices; ++i) {
        int64_t p = -1;
        if (w[i] != GZ_BFS_UNVISITED && gz_bfs_parent(w[i]) >= 0) {
            p = gz_bfs_parent(w[i]);
            if (G.original_ids != NULL) p = (int64_t)G.original_ids[p];
        }
        if (parents) parents[i] = p;
        if (depths) depths[i] = (w[i] == GZ_BFS_UNVISITED) ? UINT64_MAX : gz_bfs_depth(w[i]);
    }
    free(w);
}

int graph_data_write_ranks_to_file(const char* filename, uint32_t algorithm)
{
    FILE* f = fopen(filename, "w");
    if (!f) return -1;
    const uint64_t n = G.num_vertices;
    if (algorithm == GZ_ALG_PR) {
        double* r = (double*)malloc(n * sizeof(double));
        gz_result_pr_ranks(r);
        for (uint64_t i = 0; i < n; ++i) fprintf(f, "%llu %.5le\n", (unsigned long long)i, r[i]);
        free(r);
    } else if (algorithm == GZ_ALG_CC) {
        uint64_t* l = (uint64_t*)malloc(n * sizeof(uint64_t));
        gz_result_cc_labels(l);
        for (uint64_t i = 0; i < n; ++i) fprintf(f, "%llu %.0lf\n", (unsigned long long)i, (double)l[i]);
        free(l);
    } else {
        int64_t* p = (int64_t*)malloc(n * sizeof(int64_t));
        gz_result_bfs(p, NULL);
        for (uint64_t i = 0; i < n; ++i) fprintf(f, "%llu %.0lf\n", (unsigned long long)i, (double)p[i]);
        free(p);
    }
    fclose(f);
    return 0;
}
