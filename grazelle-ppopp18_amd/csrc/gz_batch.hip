/*
 * gz_batch.hip -- many independent graphs in one launch per phase.
 *
 * No counterpart in the reference, which runs one graph per process.  The
 * loaded graph is the disjoint union of `count` graphs: graph g owns the
 * vertex ids [base[g], base[g+1]) and its edges never leave that range, so
 * the union's Vector-Sparse lists are the graphs' lists one after another.
 * Connected components of the union are the graphs' components (labels are
 * union ids; subtract base[g]) and run through the ordinary drivers.
 * PageRank is per graph: V and the sink mass S enter the vertex phase, so the
 * batched phases keep them per graph:
 *   - the pull sweep walks a chunk table whose chunks never straddle two
 *     graphs; each chunk's length comes from its graph's size bucket, and one
 *     persistent grid of wavefronts strides over all chunks;
 *   - the merge buffer has one record per chunk (edge_pull_op_merge_with_merge_buffer);
 *   - S[g] is the sum of graph g's chunk partials (one wavefront per graph);
 *   - the vertex phase applies ((acc + (1-S[g])/V[g]) * d + (1-d)/V[g]) / outdeg'
 *     with outdeg' = V[g] for sinks (phase_vertex_pr.asm:75-134 per graph).
 * A run of k iterations is captured once into a HIP graph and replayed.
 */
#include "gz_pr_ops.h"
#include "gz_walk.h"
#include "gz_host_internal.h"
#include "grazelle_host.h"

#include <string.h>

namespace {

struct gz_batch_t {
    uint32_t count = 0;
    uint64_t* base = nullptr;        /* host [count+1] */
    uint64_t* d_base = nullptr;      /* [count+1] */
    uint16_t* d_gid = nullptr;       /* [V] graph of each vertex */
    uint64_t* vec_base = nullptr;    /* host [count+1] first pull vector of each graph */
    uint32_t* bucket = nullptr;      /* host [count] */
    /* chunk table: (begin, end) vector pairs; chunks of a graph are
     * consecutive and never cross into another graph */
    uint64_t nchunks = 0;
    uint64_t* d_chunks = nullptr;    /* [2 * nchunks] */
    uint64_t* d_cbase = nullptr;     /* [count+1] chunk range of each graph */
    mergeaccum_t* d_merge = nullptr; /* [nchunks] */
    double* d_partials = nullptr;    /* [nchunks] */
    double* d_c = nullptr;           /* [count] (1 - S[g]) / V[g] */
    /* configuration */
    gz_batch_config_t cfg = {1, 256, {512, 512, 512, 512}};
    double damping = 0.875;
    /* captured iterations */
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_iters = 0;
};

gz_batch_t B;

__global__ void gz_k_batch_gid(const uint64_t* __restrict__ base, uint32_t count, uint64_t nv,
                               uint16_t* __restrict__ gid)
{
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = count; /* last g with base[g] <= v */
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (base[mid] <= v)
                lo = mid;
            else
                hi = mid;
        }
        gid[v] = (uint16_t)lo;
    }
}

__global__ void gz_k_batch_gather_offsets(const uint64_t* __restrict__ offsets, const uint64_t* __restrict__ base,
                                          uint32_t n, uint64_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = offsets[base[i]];
}

template <int VPL>
__global__ void gz_k_batch_pull_pr(const gz_vec_t* __restrict__ vecs, const uint64_t* __restrict__ chunks,
                                   uint64_t nchunks, const double* __restrict__ props, double* __restrict__ acc,
                                   mergeaccum_t* __restrict__ merge, double* __restrict__ partials)
{
    const int lane = gz_lane();
    const uint64_t wpb = blockDim.x / GZ_WAVE;
    const uint64_t waves = (uint64_t)gridDim.x * wpb;
    for (uint64_t chunk = (uint64_t)blockIdx.x * wpb + (threadIdx.x / GZ_WAVE); chunk < nchunks; chunk += waves) {
        const uint64_t b = chunks[2 * chunk], e = chunks[2 * chunk + 1];
        gz_pr_pull_op op{props, acc, true, 0.0};
        uint64_t first_d, carry_d;
        double carry_v;
        uint32_t edges = 0;
        gz_walk_chunk<VPL, double>(vecs, b, e, op, first_d, carry_d, carry_v, edges);
        const double S = gz_wave_sum(op.sum);
        if (lane == 0) {
            mergeaccum_t r;
            r.initial_vertex_id = first_d;
            r.final_vertex_id = carry_d;
            r.final_partial_value = carry_v;
            r.__padding__ = 0;
            merge[chunk] = r;
            partials[chunk] = S;
        }
    }
}

/* one wavefront per graph: c[g] = (1 - S[g]) / V[g] */
__global__ void gz_k_batch_sums(const double* __restrict__ partials, const uint64_t* __restrict__ cbase,
                                const uint64_t* __restrict__ base, uint32_t count, double* __restrict__ c)
{
    const uint64_t g = (uint64_t)blockIdx.x * (blockDim.x / GZ_WAVE) + threadIdx.x / GZ_WAVE;
    if (g >= count) return;
    double s = 0.0;
    for (uint64_t i = cbase[g] + (uint64_t)gz_lane(); i < cbase[g + 1]; i += GZ_WAVE) s += partials[i];
    s = gz_wave_sum(s);
    if (gz_lane() == 0) c[g] = (1.0 - s) / (double)(base[g + 1] - base[g]);
}

__global__ void gz_k_batch_vertex_pr(double* __restrict__ props, const double* __restrict__ acc,
                                     const double* __restrict__ outdeg, const uint16_t* __restrict__ gid,
                                     const uint64_t* __restrict__ base, const double* __restrict__ c, uint64_t nv,
                                     double damping)
{
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t g = gid[v];
        const double n = (double)(base[g + 1] - base[g]);
        const double k = (1.0 - damping) / n;
        double od = outdeg[v];
        if (od == 0.0) od = n;
        double x = acc[v] + c[g];
        x = x * damping;
        x = x + k;
        props[v] = x / od;
    }
}

__global__ void gz_k_batch_init_pr(double* __restrict__ props, double* __restrict__ acc,
                                   const double* __restrict__ outdeg, const uint16_t* __restrict__ gid,
                                   const uint64_t* __restrict__ base, uint64_t nv)
{
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv + 8;
         v += (uint64_t)gridDim.x * blockDim.x) {
        if (v < nv) {
            const uint32_t g = gid[v];
            const double n = (double)(base[g + 1] - base[g]);
            const double od = outdeg[v];
            props[v] = (1.0 / n) / (0.0 == od ? n : od);
        } else {
            props[v] = 0.0;
        }
        acc[v] = 0.0;
    }
}

unsigned gz_grid(uint64_t n, uint32_t block)
{
    uint64_t b = (n + block - 1) / block;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (unsigned)b;
}

void gz_batch_free_graph(void)
{
    if (B.exec) GZ_CHECK(hipGraphExecDestroy(B.exec));
    if (B.graph) GZ_CHECK(hipGraphDestroy(B.graph));
    B.exec = nullptr;
    B.graph = nullptr;
    B.graph_iters = 0;
}

void gz_batch_free_chunks(void)
{
    gz_batch_free_graph();
    gz_device_free(B.d_chunks);
    gz_device_free(B.d_cbase);
    gz_device_free(B.d_merge);
    gz_device_free(B.d_partials);
    B.d_chunks = nullptr;
    B.d_cbase = nullptr;
    B.d_merge = nullptr;
    B.d_partials = nullptr;
    B.nchunks = 0;
}

uint32_t gz_bucket_of(uint64_t nv)
{
    uint32_t bits = 0;
    while (bits < 63 && (nv - 1) >> bits) ++bits;
    const int b = ((int)bits - 10) / 2;
    return (uint32_t)(b < 0 ? 0 : (b >= GZ_BATCH_BUCKETS ? GZ_BATCH_BUCKETS - 1 : b));
}

/* Chunk table over the graphs whose bucket bit is set in `mask`. */
void gz_batch_build_chunks(uint32_t mask)
{
    gz_batch_free_chunks();
    uint64_t* cbase = (uint64_t*)malloc((B.count + 1) * sizeof(uint64_t));
    uint64_t n = 0;
    for (uint32_t g = 0; g < B.count; ++g) {
        cbase[g] = n;
        if (!((mask >> B.bucket[g]) & 1u)) continue;
        const uint64_t nvec = B.vec_base[g + 1] - B.vec_base[g];
        const uint64_t len = B.cfg.chunk[B.bucket[g]];
        n += (nvec + len - 1) / len;
    }
    cbase[B.count] = n;
    uint64_t* pairs = (uint64_t*)malloc((2 * n + 2) * sizeof(uint64_t));
    uint64_t c = 0;
    for (uint32_t g = 0; g < B.count; ++g) {
        if (!((mask >> B.bucket[g]) & 1u)) continue;
        const uint64_t len = B.cfg.chunk[B.bucket[g]];
        const uint64_t end = B.vec_base[g + 1];
        for (uint64_t s = B.vec_base[g]; s < end; s += len) {
            pairs[2 * c] = s;
            pairs[2 * c + 1] = (s + len < end) ? s + len : end;
            ++c;
        }
    }
    B.nchunks = n;
    B.d_chunks = (uint64_t*)gz_device_malloc((2 * n + 2) * sizeof(uint64_t));
    if (n) gz_copy_to_device(B.d_chunks, pairs, 2 * n * sizeof(uint64_t));
    B.d_cbase = (uint64_t*)gz_device_malloc((B.count + 1) * sizeof(uint64_t));
    gz_copy_to_device(B.d_cbase, cbase, (B.count + 1) * sizeof(uint64_t));
    B.d_merge = (mergeaccum_t*)gz_device_malloc((n ? n : 1) * sizeof(mergeaccum_t));
    B.d_partials = (double*)gz_device_malloc((n ? n : 1) * sizeof(double));
    free(pairs);
    free(cbase);
}

/* Grid of the persistent pull sweep: at most 8 waves per SIMD over 256 CUs. */
unsigned gz_batch_pull_blocks(uint64_t nchunks, uint32_t block)
{
    const uint64_t wpb = block / GZ_WAVE;
    uint64_t blocks = (nchunks + wpb - 1) / wpb;
    const uint64_t cap = (256ull * 32ull) / wpb;
    if (blocks > cap) blocks = cap;
    return (unsigned)(blocks ? blocks : 1);
}

/* One PageRank iteration of the batch, enqueued on the stream. */
void gz_batch_enqueue_iteration(void)
{
    const uint64_t V = gz_graph.num_vertices;
    double* props = (double*)gz_graph.vertex_props;
    double* acc = (double*)gz_graph.vertex_accumulators;
    if (B.nchunks) {
        const unsigned blocks = gz_batch_pull_blocks(B.nchunks, B.cfg.workgroup_size);
#define GZ_BATCH_PULL_ARGS                                                                                      \
    dim3(blocks), dim3(B.cfg.workgroup_size), 0, gz_stream, gz_graph.pull_vectors, (const uint64_t*)B.d_chunks, B.nchunks, \
        (const double*)props, acc, B.d_merge, B.d_partials
        if (B.cfg.vectors_per_lane == 4)
            hipLaunchKernelGGL(gz_k_batch_pull_pr<4>, GZ_BATCH_PULL_ARGS);
        else if (B.cfg.vectors_per_lane == 2)
            hipLaunchKernelGGL(gz_k_batch_pull_pr<2>, GZ_BATCH_PULL_ARGS);
        else
            hipLaunchKernelGGL(gz_k_batch_pull_pr<1>, GZ_BATCH_PULL_ARGS);
#undef GZ_BATCH_PULL_ARGS
        GZ_LAUNCH_CHECK();
        edge_pull_op_merge_with_merge_buffer(B.d_merge, B.nchunks, acc);
    }
    hipLaunchKernelGGL(gz_k_batch_sums, dim3((B.count + GZ_WAVES_PER_BLOCK - 1) / GZ_WAVES_PER_BLOCK), dim3(GZ_BLOCK),
                       0, gz_stream, (const double*)B.d_partials, (const uint64_t*)B.d_cbase,
                       (const uint64_t*)B.d_base, B.count, B.d_c);
    GZ_LAUNCH_CHECK();
    hipLaunchKernelGGL(gz_k_batch_vertex_pr, dim3(gz_grid(V, GZ_BLOCK)), dim3(GZ_BLOCK), 0, gz_stream, props,
                       (const double*)acc, gz_graph.vertex_outdegrees, (const uint16_t*)B.d_gid,
                       (const uint64_t*)B.d_base, (const double*)B.d_c, V, B.damping);
    GZ_LAUNCH_CHECK();
}

bool gz_batch_config_valid(const gz_batch_config_t* c)
{
    if (c->vectors_per_lane != 1 && c->vectors_per_lane != 2 && c->vectors_per_lane != 4) return false;
    if (c->workgroup_size < 64 || c->workgroup_size > 1024 || c->workgroup_size % 64) return false;
    for (int i = 0; i < GZ_BATCH_BUCKETS; ++i)
        if (c->chunk[i] == 0 || c->chunk[i] % (64u * c->vectors_per_lane)) return false;
    return true;
}

}  // namespace

/* ---- C ABI ------------------------------------------------------------------ */

extern "C" void gz_batch_release(void)
{
    gz_batch_free_chunks();
    gz_device_free(B.d_base);
    gz_device_free(B.d_gid);
    gz_device_free(B.d_c);
    free(B.base);
    free(B.vec_base);
    free(B.bucket);
    const gz_batch_config_t cfg = B.cfg;
    B = gz_batch_t();
    B.cfg = cfg;
}

extern "C" int gz_batch_attach(const uint64_t* vertex_bases, uint32_t count)
{
    gz_batch_release();
    const uint64_t V = gz_graph.num_vertices;
    if (count == 0 || count > 65535 || vertex_bases[0] != 0 || vertex_bases[count] != V ||
        gz_graph.pull_vectors == nullptr) {
        fprintf(stderr, "grazelle: batch of %u graphs does not describe the loaded graph\n", count);
        return -1;
    }
    for (uint32_t g = 0; g < count; ++g)
        if (vertex_bases[g + 1] <= vertex_bases[g]) {
            fprintf(stderr, "grazelle: batch graph %u has no vertices\n", g);
            return -1;
        }
    B.count = count;
    B.base = (uint64_t*)malloc((count + 1) * sizeof(uint64_t));
    memcpy(B.base, vertex_bases, (count + 1) * sizeof(uint64_t));
    B.d_base = (uint64_t*)gz_device_malloc((count + 1) * sizeof(uint64_t));
    gz_copy_to_device(B.d_base, B.base, (count + 1) * sizeof(uint64_t));
    B.d_gid = (uint16_t*)gz_device_malloc((V ? V : 1) * sizeof(uint16_t));
    hipLaunchKernelGGL(gz_k_batch_gid, dim3(gz_grid(V, GZ_BLOCK)), dim3(GZ_BLOCK), 0, gz_stream,
                       (const uint64_t*)B.d_base, count, V, B.d_gid);
    GZ_LAUNCH_CHECK();
    uint64_t* d_vb = (uint64_t*)gz_device_malloc((count + 1) * sizeof(uint64_t));
    hipLaunchKernelGGL(gz_k_batch_gather_offsets, dim3((count + 1 + 255) / 256), dim3(256), 0, gz_stream,
                       (const uint64_t*)gz_graph.pull_vertex_offsets, (const uint64_t*)B.d_base, count + 1, d_vb);
    GZ_LAUNCH_CHECK();
    B.vec_base = (uint64_t*)malloc((count + 1) * sizeof(uint64_t));
    gz_copy_to_host(B.vec_base, d_vb, (count + 1) * sizeof(uint64_t));
    gz_device_free(d_vb);
    B.bucket = (uint32_t*)malloc(count * sizeof(uint32_t));
    for (uint32_t g = 0; g < count; ++g) B.bucket[g] = gz_bucket_of(B.base[g + 1] - B.base[g]);
    B.d_c = (double*)gz_device_malloc(count * sizeof(double));
    gz_batch_build_chunks(~0u);
    return 0;
}

extern "C" uint32_t gz_batch_bucket_of_graph(uint32_t g) { return g < B.count ? B.bucket[g] : 0; }

extern "C" int gz_batch_set_config(const gz_batch_config_t* cfg)
{
    if (!gz_batch_config_valid(cfg)) {
        fprintf(stderr, "grazelle: invalid batch configuration (vpl %u, workgroup %u)\n", cfg->vectors_per_lane,
                cfg->workgroup_size);
        return -1;
    }
    B.cfg = *cfg;
    if (B.count) gz_batch_build_chunks(~0u);
    return 0;
}

extern "C" void gz_batch_get_config(gz_batch_config_t* cfg) { *cfg = B.cfg; }

extern "C" void gz_batch_pr_prepare(double damping)
{
    const uint64_t V = gz_graph.num_vertices;
    B.damping = damping;
    gz_graph_reset_roles();
    hipLaunchKernelGGL(gz_k_batch_init_pr, dim3(gz_grid(V + 8, GZ_BLOCK)), dim3(GZ_BLOCK), 0, gz_stream,
                       (double*)gz_graph.vertex_props, (double*)gz_graph.vertex_accumulators,
                       gz_graph.vertex_outdegrees, (const uint16_t*)B.d_gid, (const uint64_t*)B.d_base, V);
    GZ_LAUNCH_CHECK();
    gz_synchronize();
}

extern "C" double gz_batch_pr_run(uint32_t iterations, int use_graph)
{
    if (B.count == 0) {
        fprintf(stderr, "grazelle: no batch attached\n");
        exit(73);
    }
    gz_timer_t* t = gz_timer_create();
    const int timing = gz_kernel_timing_enabled();
    if (use_graph) {
        if (B.exec == nullptr || B.graph_iters != iterations) {
            gz_batch_free_graph();
            gz_kernel_timing_enable(0); /* no event records inside the capture */
            GZ_CHECK(hipStreamBeginCapture(gz_stream, hipStreamCaptureModeThreadLocal));
            for (uint32_t i = 0; i < iterations; ++i) gz_batch_enqueue_iteration();
            GZ_CHECK(hipStreamEndCapture(gz_stream, &B.graph));
            GZ_CHECK(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
            B.graph_iters = iterations;
            gz_kernel_timing_enable(timing);
        }
        gz_timer_start(t);
        GZ_CHECK(hipGraphLaunch(B.exec, gz_stream));
        gz_timer_stop(t);
    } else {
        gz_timer_start(t);
        for (uint32_t i = 0; i < iterations; ++i) gz_batch_enqueue_iteration();
        gz_timer_stop(t);
    }
    const double ms = gz_timer_elapsed_ms(t);
    gz_timer_destroy(t);
    return ms;
}

/* Pull sweep time of the graphs in `bucket` under `cfg` (autotuner probe);
 * the vertex arrays are left as they were. */
extern "C" double gz_batch_time_pull(const gz_batch_config_t* cfg, uint32_t bucket, uint32_t reps)
{
    if (!gz_batch_config_valid(cfg) || B.count == 0) return -1.0;
    const gz_batch_config_t saved = B.cfg;
    B.cfg = *cfg;
    gz_batch_build_chunks(1u << bucket);
    double best = -1.0;
    if (B.nchunks) {
        gz_timer_t* t = gz_timer_create();
        double* props = (double*)gz_graph.vertex_props;
        double* scratch_acc = (double*)gz_device_malloc((gz_graph.num_vertices + 8) * sizeof(double));
        for (uint32_t r = 0; r < reps + 1; ++r) {
            const unsigned blocks = gz_batch_pull_blocks(B.nchunks, B.cfg.workgroup_size);
            gz_timer_start(t);
#define GZ_BATCH_TUNE_ARGS                                                                                      \
    dim3(blocks), dim3(B.cfg.workgroup_size), 0, gz_stream, gz_graph.pull_vectors, (const uint64_t*)B.d_chunks, B.nchunks, \
        (const double*)props, scratch_acc, B.d_merge, B.d_partials
            if (B.cfg.vectors_per_lane == 4)
                hipLaunchKernelGGL(gz_k_batch_pull_pr<4>, GZ_BATCH_TUNE_ARGS);
            else if (B.cfg.vectors_per_lane == 2)
                hipLaunchKernelGGL(gz_k_batch_pull_pr<2>, GZ_BATCH_TUNE_ARGS);
            else
                hipLaunchKernelGGL(gz_k_batch_pull_pr<1>, GZ_BATCH_TUNE_ARGS);
#undef GZ_BATCH_TUNE_ARGS
            GZ_LAUNCH_CHECK();
            gz_timer_stop(t);
            const double ms = gz_timer_elapsed_ms(t);
            if (r > 0 && (best < 0.0 || ms < best)) best = ms; /* the first run warms up */
        }
        gz_device_free(scratch_acc);
        gz_timer_destroy(t);
    } else {
        best = 0.0;
    }
    B.cfg = saved;
    gz_batch_build_chunks(~0u);
    return best;
}

extern "C" void gz_batch_pr_ranks(double* out)
{
    const uint64_t V = gz_graph.num_vertices;
    gz_copy_to_host(out, gz_graph.vertex_props, V * sizeof(double));
    const double* od = gz_graph_outdegrees_host();
    for (uint32_t g = 0; g < B.count; ++g) {
        const double n = (double)(B.base[g + 1] - B.base[g]);
        for (uint64_t v = B.base[g]; v < B.base[g + 1]; ++v) out[v] = out[v] * (0.0 == od[v] ? n : od[v]);
    }
}
