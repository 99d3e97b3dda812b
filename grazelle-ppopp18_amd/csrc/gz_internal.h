/*
 * gz_internal.h -- declarations shared by the HIP translation units of
 * libgrazelle.so: error checking, launch geometry, device helpers for the
 * Vector-Sparse format, wavefront primitives and the internal scratch.
 */
#ifndef GZ_INTERNAL_H
#define GZ_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "grazelle_phases.h"

typedef unsigned int gz_u32x4 __attribute__((ext_vector_type(4)));

/* A packed vector as the walker sees it: the four 32-bit neighbour ids of a
 * Vector-Sparse vector (GZ_PACKED_NONE for an empty lane; valid lanes form a
 * prefix) and d, its destination or a flag while it is being loaded
 * (gz_packed_list, gz_walk.h). */
#define GZ_PACKED_NONE 0xffffffffu
/* 16-bit block offsets: empty lane, and the first-of-destination flag */
#define GZ_PACKED16_NONE 0x7fffu
#define GZ_PACKED16_FIRST 0x8000u
struct gz_packed_vec {
    uint32_t s0, s1, s2, s3, d;
};

#define GZ_WAVE 64
#define GZ_BLOCK 256
#define GZ_WAVES_PER_BLOCK (GZ_BLOCK / GZ_WAVE)

/* Persistent grid used by the frontier-driven kernels: enough waves to fill
 * 256 CUs several times over, each looping over its share of the work. */
#define GZ_PERSISTENT_BLOCKS 2048

/* Largest frontier statistic (out-degrees plus count) for which a CC or BFS push
 * scans the HasInfo bitmap directly instead of compacting the frontier. */
#define GZ_SPARSE_PUSH_VOTE (1u << 16)
/* Up to this statistic each wavefront of that push reads 64 words and walks
 * every source itself (the statistic bounds its vectors to 16 batches); above
 * it each reads 16 words and sends a round of sources with more vectors than
 * one batch to the overflow queue. */
#define GZ_SPARSE_PUSH_WIDE_VOTE 4096
/* Overflow queue: pieces of at most this many vectors (one batch of four
 * steps of 64), and room for this many pieces. */
#define GZ_PUSH_PIECE_VECTORS 256
#define GZ_PUSH_QUEUE_PIECES (1u << 17)

#define GZ_CHECK(call)                                                                   \
    do {                                                                                 \
        hipError_t gz_err_ = (call);                                                     \
        if (gz_err_ != hipSuccess) {                                                     \
            fprintf(stderr, "grazelle: HIP error '%s' from %s at %s:%d\n",               \
                    hipGetErrorString(gz_err_), #call, __FILE__, __LINE__);              \
            exit(70);                                                                    \
        }                                                                                \
    } while (0)

#define GZ_LAUNCH_CHECK() GZ_CHECK(hipGetLastError())

/* ---- internal scratch, owned by gz_runtime.hip ---------------------------- */

typedef struct gz_scratch_t {
    void* partials;              /* per-wave partial sums (double or uint64) */
    uint64_t partials_capacity;  /* entries */
    uint64_t partials_count;     /* entries written by the last edge phase */
    uint32_t partials_is_double; /* 1 for PR (S), 0 for CC/BFS (vote) */
    uint64_t* stats;             /* per-wave [vectors, edges] pairs when collecting */
    uint64_t stats_count;
    uint64_t* reduce_tmp;        /* second-level reduction slots + done counter */
    /* active-vertex compaction (frontier skip) */
    uint64_t* active_vertices;   /* [V] */
    uint64_t* active_prefix;     /* [V+1] exclusive prefix of vector counts */
    uint64_t* active_block_sums; /* per-block [vertices, vectors] */
    uint64_t* active_totals;     /* [0] = active vertices, [1] = active vectors */
    uint64_t active_capacity;
    uint64_t* bfs_first_vector;  /* [V] first qualifying pull vector of each dst, ~0 when none */
    uint64_t* features;          /* [3][partials_capacity] selector features per wave (BFS edge phases) */
    uint32_t features_on;         /* BFS edge phases fold the selector's features */
    const uint32_t* features_indeg; /* in-degrees for the in-edge feature, or nullptr (not folded) */
    uint64_t* stamp_target;      /* when set, the next compaction writes the wall clock there (once) */
    uint64_t scratch_vertices;   /* V the vertex-sized buffers were sized for */
    /* overflow queue of the sparse CC push: pieces of at most
     * GZ_PUSH_PIECE_VECTORS vectors as [source, first vector, count] */
    uint64_t* push_pieces;       /* [3 * GZ_PUSH_QUEUE_PIECES] */
    uint64_t* push_queue;        /* [2] piece counters used alternately; the unused one is 0 */
    uint32_t push_queue_parity;  /* counter the next sparse push appends to */
} gz_scratch_t;

extern gz_scratch_t gz_scratch;
/* Pieces the overflow queue accepts, 1..GZ_PUSH_QUEUE_PIECES (set by
 * gz_push_queue_capacity); a piece beyond it is pushed by the lane that made it. */
extern uint32_t gz_push_queue_cap;
extern hipStream_t gz_stream;
extern uint32_t gz_cu_count; /* compute units of the device */

/* ---- source-sliced pull layout, owned by gz_slices.hip -------------------
 * The pull vectors split by source, in levels of eight slices (slice
 * 8 level + x belongs to XCD x).  With T tiers of windows of S vertices and G
 * LDS blocks of H vertices:
 *   level 0:          tier-0 window x, sources in [x S, (x + 1) S) except
 *                     those of its blocks;
 *   level b, 1..G:    block b of window x, sources in [x S + b H, x S + (b + 1) H);
 *   level G + t:      tier-t window x (t = 1 .. T-1), sources in
 *                     [(8t + x) S, (8t + x + 1) S);
 *   slice 8 (G + T):  every other edge (the cold slice).
 * Each slice is an ordinary pull list (dst-grouped), and the slices are stored
 * one after another.  Workgroups are dispatched to the XCDs round robin, so a
 * chunk table that gives XCD x only the chunks of its slices, level by level,
 * plus a share of the cold slice keeps each XCD's gathers inside one window
 * at a time, which its own L2 can hold; the persistent pull loads each block
 * into LDS in turn, so the gathers of a block slice never leave the CU. */
#define GZ_XCDS 8
#define GZ_MAX_TIERS 3
#define GZ_MAX_LEVELS 7 /* tiers + blocks: the slice mask of a vertex fits 64 bits */
#define GZ_MAX_SLICES (GZ_XCDS * GZ_MAX_LEVELS + 1)
#define GZ_SLICE_KEY_BITS 7 /* slice field of the layout builder's sort keys */

/* First vertex of the window that slice `sl` (not the cold slice) packs its
 * 24-bit offsets against: tier-0 window x for levels 0..G, else its tier's. */
static inline __host__ __device__ uint64_t gz_slice_window(uint32_t sl, uint64_t S, uint32_t blocks)
{
    const uint32_t level = sl / GZ_XCDS, x = sl % GZ_XCDS;
    const uint32_t tier = level <= blocks ? 0u : level - blocks;
    return ((uint64_t)tier * GZ_XCDS + x) * S;
}

/* The slice that holds edges from source s (ids after the interleaved
 * renumbering; any ids for a layout built without it). */
static inline __host__ __device__ uint32_t gz_slice_of_source(uint64_t s, uint64_t S, uint32_t tiers, uint32_t blocks,
                                                             uint64_t H)
{
    if (S == 0 || s >= (uint64_t)GZ_XCDS * tiers * S) return GZ_XCDS * (blocks + tiers);
    const uint64_t w = s / S, local = s - w * S;
    const uint32_t tier = (uint32_t)(w / GZ_XCDS), x = (uint32_t)(w % GZ_XCDS);
    uint32_t level;
    if (tier > 0)
        level = blocks + tier;
    else if (blocks > 0 && local >= H && local < (uint64_t)(blocks + 1) * H)
        level = (uint32_t)(local / H);
    else
        level = 0;
    return GZ_XCDS * level + x;
}

/* Offsets of the slices' accumulator runs in one array. */
typedef struct gz_slice_offsets {
    uint64_t off[GZ_MAX_SLICES + 1];
} gz_slice_offsets;

typedef struct gz_source_slices_t {
    const gz_vec_t* base_list;        /* the pull list the layout was built from */
    uint64_t base_count;
    uint64_t slice_vertices;          /* S */
    uint32_t tiers;                   /* T */
    uint32_t blocks;                  /* G */
    uint64_t block_vertices;          /* H */
    uint32_t nslices;                 /* 8 (G + T) + 1 */
    /* packed vectors (gz_walk.h): the window slices' (levels 0 and G+1 ..) as
     * 24-bit window offsets, 12 bytes each, the block slices' (levels 1..G) as
     * 16-bit block offsets, 8 bytes each, the last slice's as 32-bit ids */
    uint32_t* nbr24;                  /* 3 words per window vector: vector i of level 0 at i, of a later
                                       * tier at i - block_span */
    uint64_t* nbr16;                  /* block vector i at i - begin[8] */
    uint64_t block_span;              /* vectors of the block slices, begin[8 (G+1)] - begin[8] */
    gz_u32x4* nbr;                    /* last slice: vector begin[nslices-1] + j at nbr[j] */
    uint32_t* shared;                 /* every vector's shared vertex (local rank) */
    uint64_t begin[GZ_MAX_SLICES + 1]; /* first vector of each slice */
    void* mask;                       /* [V] bit x: the vertex has vectors in slice x; uint32 words
                                       * for at most 32 slices, else uint64 */
    /* Inside slice x, destinations are numbered densely (local rank: the
     * count of lower vertices that have vectors in slice x), and the shared
     * vertex of slice x's vectors is that local rank. */
    double* parts;                    /* slice x's accumulators at parts + part_base.off[x], by local rank */
    gz_slice_offsets part_base;
    uint32_t* gid;                    /* global vertex of each (slice, local rank), same offsets as parts */
    uint32_t* rank_base;              /* [rank_blocks][nslices] index in parts of the accumulator of each
                                       * 64-vertex block's first vertex in slice x (part_base.off[x] + its
                                       * local rank) */
    uint64_t rank_blocks;             /* ceil(V / 64) */
    uint32_t* outdeg32;               /* [V] out-degree as a 32-bit count (0 for sinks), read by the
                                       * vertex phase in place of the FP64 array */
    gz_slice_chunk_t* table;          /* chunk slots, built for table_chunk_len */
    uint64_t slots;
    uint64_t table_chunk_len;
    uint32_t table_cold_spread;
    mergeaccum_t* merge;              /* one record per chunk, in slice order */
    uint64_t records;
    uint64_t xcd_chunks[GZ_XCDS];     /* chunks in each XCD's slot sequence */
    uint32_t* counters;               /* per-XCD work counters of the persistent pull */
} gz_source_slices_t;

extern gz_source_slices_t gz_slices;
/* PageRank pull over gz_slices (pull and per-slice merge; the vertex phase
 * sums the slices). */
void gz_pull_pr_sliced(const gz_vec_t* edge_list, uint64_t edge_list_count);
/* Connected Components pull over gz_slices. */
void gz_pull_cc_sliced(const gz_vec_t* edge_list, uint64_t edge_list_count);

/* Bracket a kernel launch for gz_kernel_timing_*; no-ops when disabled. */
void gz_ktimer_begin(int id);
void gz_ktimer_end(int id);

/* Make sure the scratch can hold `waves` partials and vertex-sized buffers for V. */
void gz_scratch_reserve(uint64_t waves, uint64_t num_vertices);
/* Number of wave chunks covering n vectors at the configured chunk length. */
uint64_t gz_chunks_for(uint64_t n);
/* Chunk length in vectors (a multiple of 64 x vectors per lane). */
uint64_t gz_chunk_len(void);
uint32_t gz_vectors_per_lane(void);
/* Reduce gz_scratch.partials[0..count) into dst[0] (bit pattern). */
void gz_reduce_partials(uint64_t* dst);
/* Deterministic two-stage FP64 sum of in[count] into out[0] (as bits): up to
 * 1024 blocks write slots[], the last one sums them and clears *done, which
 * must be 0 when the launch starts. */
void gz_reduce_f64_into(const double* in, uint64_t count, double* slots, unsigned int* done, uint64_t* out);
/* Workgroups of the first stage of that reduction for `count` inputs; each
 * sums ceil(count / blocks) consecutive inputs (at most 1024 workgroups). */
uint64_t gz_reduce_blocks(uint64_t count);
/* PageRank kernels over explicit arrays (gz_pr.hip). */
void gz_launch_pr_pull_explicit(const gz_vec_t* vecs, uint64_t nvec, uint64_t chunk_len, uint64_t nchunks,
                                const double* props, double* acc, mergeaccum_t* merge, double* partials);
void gz_launch_pr_vertex_explicit(double* props, double* acc, const double* outdeg, uint64_t num_vertices,
                                  const uint64_t* sum_word, double damping);
void gz_launch_merge_sum(const mergeaccum_t* merge, uint64_t count, double* acc);
/* Build the compacted active-vertex list from `bitmap` against `offsets`. */
void gz_frontier_compact(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices,
                         uint64_t vector_count);
/* Where a compaction writes: the vertices with vectors whose bit is set, in
 * vertex order ([V]), the exclusive prefix of their vector counts ([V+1]),
 * per-block sums (2 gz_compact_blocks(V) + 8 words) and the totals ([0]
 * vertices, [1] vectors). */
struct gz_compact_out {
    uint64_t* list;
    uint64_t* prefix;
    uint64_t* block_sums;
    uint64_t* totals;
};
/* The vertex list alone (scratch active_vertices, totals[0]; no vector
 * prefix, totals[1] = 0): vertices with their bit set that own vectors. */
void gz_frontier_list(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices,
                      uint64_t vector_count);
/* As gz_frontier_compact, into other buffers than the scratch's. */
void gz_frontier_compact_to(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices,
                            uint64_t vector_count, const gz_compact_out& out);

/* The pinned host copy of the reduce buffer (gz_runtime.hip).  A kernel
 * handed gz_reduce_mirror_target(rb) writes the eight words of rb there too
 * (gz_mirror_words, after its own writes to rb), so that
 * phase_op_combine_global_var_from_buf(rb) needs no copy; a launcher that
 * writes a reduce buffer otherwise calls gz_reduce_mirror_stale. */
extern "C" uint64_t* gz_reduce_mirror_target(const uint64_t* reduce_buffer);
extern "C" void gz_reduce_mirror_stale(void);

/* ---- device helpers ------------------------------------------------------- */

typedef unsigned long long gz_u64x2 __attribute__((ext_vector_type(2)));

struct gz_lanes {
    uint64_t l0, l1, l2, l3;
};

__device__ __forceinline__ gz_lanes gz_load_vector(const gz_vec_t* __restrict__ list, uint64_t i)
{
    const gz_u64x2* p = reinterpret_cast<const gz_u64x2*>(list + i);
    const gz_u64x2 a = __builtin_nontemporal_load(p);
    const gz_u64x2 b = __builtin_nontemporal_load(p + 1);
    gz_lanes v;
    v.l0 = a.x;
    v.l1 = a.y;
    v.l2 = b.x;
    v.l3 = b.y;
    return v;
}

__device__ __forceinline__ uint64_t gz_shared_vertex(const gz_lanes& v)
{
    return ((v.l0 >> 48) & 0x7fffull) | (((v.l1 >> 48) & 0x7fffull) << 15) |
           (((v.l2 >> 48) & 0x7fffull) << 30) | (((v.l3 >> 48) & 0x7ull) << 45);
}

__device__ __forceinline__ bool gz_lane_valid(uint64_t l) { return (l >> 63) != 0; }
__device__ __forceinline__ uint64_t gz_lane_id(uint64_t l) { return l & GZ_LANE_NEIGHBOUR_MASK; }

__device__ __forceinline__ int gz_lane() { return (int)(threadIdx.x & (GZ_WAVE - 1)); }

/* Index of the segment head at or below `lane`, given the ballot of heads. */
__device__ __forceinline__ int gz_segment_start(uint64_t heads, int lane)
{
    const uint64_t upto = heads & ((2ull << lane) - 1ull);
    return 63 - __clzll((long long)upto);
}

/* Inclusive segmented scans across the wavefront; segments are runs of equal
 * keys, described by their start lane. */
__device__ __forceinline__ double gz_segscan_sum(double x, int lane, int seg_start)
{
#pragma unroll
    for (int o = 1; o < GZ_WAVE; o <<= 1) {
        const double t = __shfl_up(x, o, GZ_WAVE);
        if (lane - o >= seg_start) x += t;
    }
    return x;
}

__device__ __forceinline__ uint64_t gz_segscan_min(uint64_t x, int lane, int seg_start)
{
#pragma unroll
    for (int o = 1; o < GZ_WAVE; o <<= 1) {
        const uint64_t t = __shfl_up(x, o, GZ_WAVE);
        if (lane - o >= seg_start && t < x) x = t;
    }
    return x;
}

__device__ __forceinline__ double gz_wave_sum(double x)
{
#pragma unroll
    for (int o = GZ_WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, GZ_WAVE);
    return x;
}

/* Sum over a workgroup: a butterfly inside each wavefront, then thread 0
 * adds the wavefront sums in wavefront order (the result is returned to
 * thread 0 only).  Every thread of the workgroup must call it. */
template <typename T>
__device__ __forceinline__ T gz_block_sum(T x, T* lds)
{
    for (int o = GZ_WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, GZ_WAVE);
    const int w = threadIdx.x / GZ_WAVE;
    if ((threadIdx.x & (GZ_WAVE - 1)) == 0) lds[w] = x;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x / GZ_WAVE); ++i) r += lds[i];
    }
    __syncthreads();
    return r;
}

/* One record of the parallel merge of phases.c:41-70.  Records sharing a
 * final vertex form a run; the first record of each run (its head) folds the
 * run's partials in record order and, when the record after the run starts
 * on the same vertex (that unit stored its first vertex directly), folds the
 * accumulator too.  A record whose initial id has bit 63 set is unused. */
__device__ __forceinline__ void gz_merge_sum_record(const mergeaccum_t* __restrict__ mb, uint64_t count, uint64_t u,
                                                    double* __restrict__ acc)
{
    const mergeaccum_t r = mb[u];
    if (r.initial_vertex_id & GZ_LANE_VALID_BIT) return;
    if (u > 0) {
        const mergeaccum_t p = mb[u - 1];
        if (!(p.initial_vertex_id & GZ_LANE_VALID_BIT) && p.final_vertex_id == r.final_vertex_id) return;
    }
    double proposed = r.final_partial_value;
    uint64_t j = u + 1;
    while (j < count && mb[j].final_vertex_id == r.final_vertex_id) {
        proposed = proposed + mb[j].final_partial_value;
        ++j;
    }
    if (j < count && mb[j].initial_vertex_id == r.final_vertex_id) proposed = proposed + acc[mb[j].initial_vertex_id];
    acc[r.final_vertex_id] = proposed;
}

/* Position of the k-th (0-based) set bit of x; k < popcount(x). */
__device__ __forceinline__ uint32_t gz_nth_set_bit(uint64_t x, uint32_t k)
{
    uint32_t pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t low = x & ((1ull << w) - 1);
        const uint32_t c = (uint32_t)__popcll(low);
        if (k >= c) {
            k -= c;
            x >>= w;
            pos += (uint32_t)w;
        } else {
            x = low;
        }
    }
    return pos;
}

__device__ __forceinline__ uint64_t gz_wave_sum_u64(uint64_t x)
{
#pragma unroll
    for (int o = GZ_WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, GZ_WAVE);
    return x;
}

/* Selector features folded into a BFS edge phase: every vertex the phase
 * activates (exactly once in BFS) adds (1, outdeg, indeg) to the lane's
 * counters, and each wavefront writes its sums next to its vote partial, one
 * component after another: out[wave], out[stride + wave] and
 * out[2 * stride + wave] (vertices, out-edges, in-edges), so that a reader
 * of one component reads it as densely as the votes.  out == nullptr turns
 * it off. */
struct gz_feature_sink {
    uint64_t* out;
    const uint32_t* indeg;
    uint64_t stride;
};

struct gz_feature_acc {
    uint64_t count, out_edges, in_edges;
};

/* Copy the eight reduce-buffer words to the pinned host copy (one thread,
 * after it wrote its own words of `out`); nothing when mirror is null. */
__device__ __forceinline__ void gz_mirror_words(uint64_t* __restrict__ mirror, const uint64_t* out)
{
    if (mirror == nullptr) return;
#pragma unroll
    for (int i = 0; i < 8; ++i) mirror[i] = out[i];
}

/* One activated vertex: its frontier statistic (gz_frontier_stat) and, when
 * the sink is on, (1, out-degree, in-degree) added to the lane's feature
 * counters, with the out-degree read once. */
__device__ __forceinline__ uint64_t gz_activate_stat(gz_feature_acc& f, const gz_feature_sink& s,
                                                     const double* __restrict__ outdeg, uint64_t v, uint32_t mode)
{
    const double od = (s.out != nullptr || mode != GZ_STAT_COUNT_ONLY) ? outdeg[v] : 0.0;
    uint64_t st = 0;
    if (mode != GZ_STAT_OUTDEG_ONLY) st += 1;
    if (mode != GZ_STAT_COUNT_ONLY) st += (uint64_t)llrint(od);
    if (s.out) {
        f.count += 1;
        f.out_edges += (uint64_t)od;
        f.in_edges += s.indeg ? (uint64_t)s.indeg[v] : 0ull;
    }
    return st;
}

/* Lane `src`'s value of x in every lane. */
__device__ __forceinline__ uint64_t gz_shfl_u64(uint64_t x, int src)
{
    return (uint64_t)__shfl((unsigned long long)x, src, GZ_WAVE);
}

__device__ __forceinline__ void gz_feature_flush(const gz_feature_acc& f, const gz_feature_sink& s, uint64_t wave)
{
    if (s.out) {
        const uint64_t c = gz_wave_sum_u64(f.count), o = gz_wave_sum_u64(f.out_edges),
                       i = gz_wave_sum_u64(f.in_edges);
        if ((threadIdx.x & (GZ_WAVE - 1)) == 0) {
            s.out[wave] = c;
            s.out[s.stride + wave] = o;
            s.out[2 * s.stride + wave] = i;
        }
    }
}

/* A kernel launched inside a captured iteration loop reads the loop's
 * engine word first and runs only when the engine is one of `mask` (bit e for
 * engine e); word == nullptr: always runs. */
struct gz_guard {
    const uint64_t* word;
    uint64_t mask;
};

__device__ __forceinline__ bool gz_guard_off(const gz_guard& g)
{
    return g.word != nullptr && ((g.mask >> (*g.word & 63)) & 1ull) == 0;
}

/* gz_frontier_compact whose three passes run under `guard`, on grids of at
 * most `max_grid` workgroups (the counts of the narrow path must already be
 * built: nothing is allocated, so the launches can be captured). */
void gz_frontier_compact_guarded(const uint64_t* bitmap, const uint64_t* offsets, uint64_t num_vertices,
                                 uint64_t vector_count, gz_guard guard, uint64_t max_grid);

/* Bitmaps are arrays of uint64 words, bit v%64 of word v/64 (LSB first).  The
 * device updates them through 32-bit atomics on the same bytes. */
__device__ __forceinline__ bool gz_bit_test(const uint64_t* __restrict__ bm, uint64_t v)
{
    return (bm[v >> 6] >> (v & 63)) & 1ull;
}

__device__ __forceinline__ void gz_bit_set(uint64_t* bm, uint64_t v)
{
    unsigned int* w = reinterpret_cast<unsigned int*>(bm) + (v >> 5);
    atomicOr(w, 1u << (v & 31));
}

__device__ __forceinline__ void gz_bit_clear(uint64_t* bm, uint64_t v)
{
    unsigned int* w = reinterpret_cast<unsigned int*>(bm) + (v >> 5);
    atomicAnd(w, ~(1u << (v & 31)));
}

/* phase_helper_add_frontier_stat (phasehelpers.inc:388-399): +1 unless the
 * count is disabled, +(uint64)outdeg(v) unless outdegrees are disabled. */
__device__ __forceinline__ uint64_t gz_frontier_stat(const double* __restrict__ outdeg, uint64_t v, uint32_t mode)
{
    uint64_t s = 0;
    if (mode != GZ_STAT_OUTDEG_ONLY) s += 1;
    if (mode != GZ_STAT_COUNT_ONLY) s += (uint64_t)llrint(outdeg[v]);
    return s;
}

#endif /* GZ_INTERNAL_H */
