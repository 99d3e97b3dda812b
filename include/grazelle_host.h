/*
 * grazelle_host.h -- host-side C API of libgrazelle.so: edge-list files,
 * Vector-Sparse ingress, the algorithm drivers and the command line.  These
 * are the reference's host files (main.c, cmdline.c, graphdata.c,
 * execution_{pr,cc,bfs}.c, scheduler.c) kept as host C code that reaches the
 * GPU only through grazelle_phases.h.
 */
#ifndef GRAZELLE_HOST_H
#define GRAZELLE_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "grazelle_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* -------- binary edge-list files (README.md:50-62, graphdata.c:533-591) --
 * little-endian uint64: V, E, then E pairs (src, dst).  A "-pull" file is
 * grouped by dst ascending, a "-push" file by src ascending.  `pairs` holds
 * 2*E words laid out as in the file. */
int gz_edge_list_write(const char* path, uint64_t num_vertices, uint64_t num_edges, const uint64_t* pairs);
/* Returns 0 and a malloc'd pair array, or -1 (missing file), -2 (truncated
 * header), -3 (truncated body). */
int gz_edge_list_read(const char* path, uint64_t* num_vertices, uint64_t* num_edges, uint64_t** pairs);

/* Sort pairs by (dst, src) when by_dst, else by (src, dst); drop self-loops
 * and duplicate edges when dedupe.  Returns the new edge count. */
uint64_t gz_edges_sort(uint64_t* pairs, uint64_t num_edges, int by_dst, int dedupe);

/* -------- Vector-Sparse ingress (graphdata.c:143-173, 618-788) ----------- */

/* Pack a grouped edge list into 4-edge vectors.  group_by_dst selects the
 * pull layout (shared vertex = dst).  `out` needs room for num_edges vectors.
 * Returns the vector count. */
uint64_t gz_pack_vectors(const uint64_t* pairs, uint64_t num_edges, int group_by_dst, gz_vec_t* out);
/* Vectors the packer would produce (sum over groups of ceil(deg/4)). */
uint64_t gz_count_vectors(const uint64_t* pairs, uint64_t num_edges, int group_by_dst);
/* Reference-format vertex index, index[V+64] (graphdata.c:219-267). */
void gz_build_vertex_index(const gz_vec_t* vectors, uint64_t vector_count, uint64_t num_vertices,
                           uint64_t* index, uint64_t* index_start, uint64_t* index_end);
/* CSR-style offsets over vectors, offsets[V+1]: vectors of v are
 * [offsets[v], offsets[v+1]). */
void gz_build_vertex_offsets(const gz_vec_t* vectors, uint64_t vector_count, uint64_t num_vertices, uint64_t* offsets);
/* FP64 outdegrees counted from the source column (graphdata.c:792-821). */
void gz_count_outdegrees(const uint64_t* pairs, uint64_t num_edges, uint64_t num_vertices, double* outdeg);

/* -------- seeded R-MAT generator ----------------------------------------- */

typedef struct gz_rmat_params_t {
    uint32_t scale;            /* vertex ids drawn from [0, 2^scale) */
    uint64_t num_vertices;     /* 0: 2^scale; else ids >= num_vertices are folded by modulo */
    uint64_t num_edges;        /* raw edges drawn before dedupe */
    double a, b, c;            /* quadrant probabilities, d = 1 - a - b - c */
    uint64_t seed;
    uint64_t edge_offset;      /* first edge counter (lets callers draw disjoint streams) */
} gz_rmat_params_t;

/* Draw raw R-MAT edges into pairs[2*num_edges] (src,dst interleaved). */
void gz_rmat_draw(const gz_rmat_params_t* p, uint64_t* pairs);
/* splitmix64 step used by the generator (exposed for tests). */
uint64_t gz_splitmix64(uint64_t x);

/* -------- graph loading into device memory (graphdata.c:957-1049) -------- */

/* Load from host pair arrays already grouped as the files would be.  Either
 * list may be NULL (then that layout is absent).  Returns 0 or dies. */
int gz_graph_load_pairs(uint64_t num_vertices, uint64_t pull_edges, const uint64_t* pull_pairs,
                        uint64_t push_edges, const uint64_t* push_pairs);
/* Reference entry point: reads "<base>-pull" and "<base>-push".  Exits 255
 * with a message when a file cannot be read, like graphdata.c:963-967. */
void graph_data_read_from_file(const char* filename_gather, const char* filename_scatter,
                               uint32_t num_numa_nodes, const uint32_t* numa_nodes);
void graph_data_allocate_merge_buffers(uint64_t num_threads, uint32_t num_numa_nodes, const uint32_t* numa_nodes);

/* Device ingress (replaces the packing consumers of graphdata.c:595-951 with
 * a device pipeline; no reference counterpart as an entry point).  Both
 * layouts are built in HBM from raw (src, dst) pairs: grouped by dst (pull)
 * and src (push) ascending, neighbours ascending inside a group, self-loops
 * and duplicates dropped, symmetrised first when `symmetrize`.  The result is
 * bit-identical to gz_edges_sort(dedupe) + gz_pack_vectors on the host.
 * Needs 1 <= V <= 2^32; dies with exit 65 otherwise.
 *
 * Flag GZ_INGRESS_RELABEL_BY_OUTDEGREE renumbers the vertices by raw
 * out-degree (descending, ties by id) before packing, so that the hot sources
 * of a power-law graph share cache lines.  The device arrays then live in the
 * new numbering; every result getter, the -o writer and the BFS root keep
 * using original ids, CC labels start from the original ids (so they are the
 * smallest original id of each component, as without renumbering), and BFS
 * parents are mapped back (a valid BFS tree; ties between equal-depth parents
 * may resolve differently). */
#define GZ_INGRESS_RELABEL_BY_OUTDEGREE 1u
/* GZ_INGRESS_SOURCE_SLICES(k) also builds the source-sliced pull layout
 * (gz_graph_build_source_slices) with S = 2^k and GZ_INGRESS_SLICE_TIERS(t)
 * tiers (1 when absent).  Together with the renumbering, each run of 8*S
 * ranks of the degree order, for the first t runs, is dealt round robin into
 * its tier's eight windows (rank i*8S + q -> id i*8S + (q % 8) * S + q / 8),
 * so that the slices of a tier carry about the same share of the edges. */
#define GZ_INGRESS_SOURCE_SLICES(k) (((uint32_t)(k) & 0x3fu) << 8)
#define GZ_INGRESS_SLICE_TIERS(t) (((uint32_t)(t) & 0x3u) << 14)
/* ... and g LDS blocks of GZ_SLICE_BLOCK_VERTICES sources per tier-0 window
 * (0 when absent; see gz_graph_build_source_slices). */
#define GZ_INGRESS_SLICE_BLOCKS(g) (((uint32_t)(g) & 0x7u) << 16)
#define GZ_SLICE_BLOCK_VERTICES 20480u
int gz_graph_build_pairs_device(uint64_t num_vertices, uint64_t num_edges, const uint64_t* pairs, int symmetrize,
                                uint32_t flags);
/* As above with the raw edges drawn on the device by the generator of
 * gz_rmat_draw (same stream), plus `extra_edges` host pairs appended (planted
 * components).  V = max(num_vertices, R-MAT vertex count). */
int gz_graph_build_rmat_device(const gz_rmat_params_t* p, uint64_t num_vertices, int symmetrize,
                               uint64_t extra_edges, const uint64_t* extra_pairs, uint32_t flags);
/* Device id of an original vertex id (identity unless renumbered). */
uint64_t gz_graph_device_vertex(uint64_t original);
/* out[v] = original id of device vertex v, V entries (identity unless
 * renumbered).  The device lists (gz_graph_download_vectors) are in device
 * numbering; this maps them back. */
void gz_graph_original_ids(uint64_t* out);
/* In-degree of every device vertex, V entries (device numbering), counted
 * on the device from the loaded lists. */
void gz_graph_indegrees(uint32_t* out);
/* Disjoint union of `count` R-MAT graphs (params[g] as for
 * gz_graph_build_rmat_device); graph g gets the ids [bases[g], bases[g+1]),
 * bases has count+1 entries and is filled here. */
int gz_graph_build_rmat_union_device(const gz_rmat_params_t* params, uint32_t count, int symmetrize,
                                     uint64_t* vertex_bases);

/* -------- source-sliced pull layout (no reference counterpart) ------------
 * Splits the loaded graph's pull vectors by source into 8 (G + T) + 1 slices,
 * in levels of eight (slice 8 level + x belongs to XCD x), T = `tiers` (1..3)
 * windows of S vertices per XCD, G = `blocks` LDS blocks of H =
 * `block_vertices` sources carved from each tier-0 window:
 *   level 0          window x of tier 0, [x*S, (x+1)*S) without its blocks
 *   level b (1..G)   block b of window x, [x*S + b*H, x*S + (b+1)*H)
 *   level G + t      window x of tier t >= 1, [(8t+x)*S, (8t+x+1)*S)
 *   slice 8 (G + T)  all other sources.
 * Each slice is an ordinary dst-grouped pull list whose destinations are
 * renumbered densely inside the slice (in vertex order).  PageRank and CC run
 * over it with pull_variant GZ_PULL_SOURCE_SLICED, which gives XCD x the
 * chunks of its slices level by level (plus a share of the last slice); with
 * an LDS cache of at least H sources (gz_slices_set_hot_cache) each workgroup
 * holds block b in LDS while it takes block b's chunks.  S is capped at
 * V / (8T) and at 2^23; G + T <= 7; G > 0 needs (G + 1) * H <= S.  Needs
 * V <= 2^28.  Replaces any earlier sliced layout. */
int gz_graph_build_source_slices(uint64_t slice_vertices, uint32_t tiers, uint32_t blocks, uint64_t block_vertices);
/* The current layout: vectors per slice (8 (G + T) + 1 entries), S, T, G and
 * H; each pointer may be NULL; returns 0 when there is none. */
int gz_graph_source_slices_info(uint64_t* vectors, uint64_t* slice_vertices, uint32_t* tiers, uint32_t* blocks,
                                uint64_t* block_vertices);
/* Copy the concatenated slice vectors and the per-vertex slice mask (bit x:
 * the vertex has vectors in slice x) to the host; either may be NULL.
 * Returns 0 when there is no sliced layout. */
int gz_graph_source_slices_download(gz_vec_t* vectors, uint64_t* mask);
/* How the chunks of slices with vectors[8 levels + 1] vectors are laid out at
 * chunk length chunk_len: XCD x (workgroup w with w % 8 == x) takes the chunks
 * of slices x, 8 + x, ... in that order, and the last slice's chunks fill the
 * XCDs up to an even share, spread between the others (cold_spread) or after
 * them.  Four slots per workgroup.  Writes up to `capacity` slots to `out`
 * (NULL: count only) and returns the slot count (0 for levels outside 1..7).
 * Host-only computation. */
uint64_t gz_slices_chunk_table(const uint64_t* vectors, uint32_t levels, uint64_t chunk_len, uint32_t cold_spread,
                               gz_slice_chunk_t* out, uint64_t capacity);
/* Layout of the last slice's chunks for the next source-sliced pull (default 1). */
void gz_slices_set_cold_spread(uint32_t cold_spread);
/* Nonzero: the last slice's gathers are nontemporal loads (default 0). */
void gz_slices_set_cold_stream(uint32_t cold_stream);
/* Keep the first `vertices` properties of each tier-0 window in LDS (0 = off,
 * else 4096, 8192, 16384 or 20480), and each LDS block in turn when the
 * blocks fit; the pull then runs as a persistent kernel.  Returns -1 for
 * another value. */
int gz_slices_set_hot_cache(uint32_t vertices);

/* PageRank runs launch iterations 2.. as one HIP graph of their launches
 * when no per-iteration record, statistic or kernel timing is asked for (on
 * by default; no reference counterpart).  A configuration (graph, options,
 * iteration count, device allocations) is captured the second time in a
 * row it runs and replayed while it repeats; results are the same. */
void gz_pr_graph_replay(int on);
/* PageRank runs that record nothing per iteration and use the
 * scheduler-aware pull, on a pull list of at most 4096 chunks, run the merge,
 * the sum of S and the vertex phase as one launch after each pull (two
 * launches per iteration instead of four; results bit-identical).  On by
 * default; 0 switches it off. */
void gz_pr_fused_tail(int on);
/* PageRank runs that record nothing per iteration and use the
 * scheduler-aware pull can run all their iterations in one cooperative
 * launch, with grid barriers between the pull and the sum-and-vertex step
 * (results bit-identical to the per-phase launches).  Off by default: a
 * barrier across the eight XCDs costs more than the launch boundary it
 * replaces.  mode 0: never (the default); 1: when 32 B per pull vector + 24 B
 * per vertex fit 128 MiB; 2: any graph.  workgroups_per_cu (1..8, capped by
 * occupancy) sizes the grid.  Returns -1 for other values. */
int gz_pr_persistent(int mode, uint32_t workgroups_per_cu);
/* Runs that took the one-launch path since the library was loaded.  When a
 * cooperative launch is refused (another process may hold the device's
 * cooperative resources), a note goes to stderr and the per-phase launches
 * run from then on. */
uint64_t gz_pr_persistent_runs(void);
/* Polls a grid barrier makes before it gives up (default 2^22, a few
 * seconds).  The kernel then leaves at once, a note goes to stderr, and the
 * run restarts from the initial ranks with the per-phase launches.  For
 * tests of that path. */
void gz_pr_persistent_spin_limit(uint32_t polls);
/* One-launch runs whose barrier gave up since the library was loaded. */
uint64_t gz_pr_persistent_gave_up(void);
/* CC runs that record nothing per iteration, with the fixed threshold or a
 * forced engine over the scheduler-aware pull, keep the vote and the engine
 * choice on the device and run their iterations as a captured HIP graph,
 * four iterations per launch (off by default, since each captured iteration
 * also dispatches the kernels of the engines it does not run; no reference
 * counterpart, execution_cc.c:124-142 decides on the host).  A configuration
 * is captured the second time in a row it runs; labels are the same either
 * way. */
void gz_cc_graph_replay(int on);
/* CC runs that record nothing per iteration, with the fixed threshold or a
 * forced push, run their push iterations over a frontier whose statistic is
 * at most `max_vote` (clamped to 16384; 0 switches it off) back to back in
 * one workgroup, the round trip to the host for each vote saved (default
 * 4096; no reference counterpart, execution_cc.c:124-142 returns to the
 * host after each iteration).  Every edge gets the same update, so labels
 * are the same; iterations keep their count, one per round. */
void gz_cc_push_tail(uint64_t max_vote);
/* Push iterations run that way since the library was loaded. */
uint64_t gz_cc_push_tail_rounds(void);
/* Graph captures and graph launches made so far (no reference counterpart). */
void gz_graph_replay_counts(uint64_t* captures, uint64_t* launches);

/* Frontier compaction reads 32-bit per-vertex vector counts when the list has
 * fewer than 2^32 vectors (default on); 0 makes it read the 64-bit offsets, as
 * it does for larger lists, so that both paths can be compared (no reference
 * counterpart). */
void gz_frontier_compact_narrow(int on);

/* Pieces of at most 256 vectors that the overflow queue of the small-frontier
 * push accepts (default and maximum 131072, which a frontier within that
 * push's bound cannot fill).  A piece beyond the cap is pushed by the lane
 * that cut it, with the same results; a smaller cap lets tests reach that
 * path.  Returns -1 for 0 or a value above the maximum (no reference
 * counterpart). */
int gz_push_queue_capacity(uint32_t pieces);

/* -------- batches of independent graphs (no reference counterpart) -------
 * The loaded graph is a disjoint union; gz_batch_attach names the graphs.
 * CC and BFS need nothing more (the ordinary drivers work on the union).
 * PageRank keeps V and the sink mass per graph: gz_batch_pr_prepare
 * initialises it, gz_batch_pr_run runs the iterations (captured into a HIP
 * graph when use_graph) and returns their device time in ms. */
#define GZ_BATCH_BUCKETS 4
typedef struct gz_batch_config_t {
    uint32_t vectors_per_lane;          /* 1, 2 or 4 */
    uint32_t workgroup_size;            /* 64..1024, a multiple of 64 */
    uint32_t chunk[GZ_BATCH_BUCKETS];   /* pull chunk length (vectors) per size bucket,
                                         * a multiple of 64 x vectors_per_lane; bucket
                                         * b holds graphs of 2^(10+2b) .. 2^(12+2b) vertices */
} gz_batch_config_t;
int gz_batch_attach(const uint64_t* vertex_bases, uint32_t count);
void gz_batch_release(void);
uint32_t gz_batch_bucket_of_graph(uint32_t graph);
int gz_batch_set_config(const gz_batch_config_t* cfg);
void gz_batch_get_config(gz_batch_config_t* cfg);
void gz_batch_pr_prepare(double damping);
double gz_batch_pr_run(uint32_t iterations, int use_graph);
/* Best-of-reps device time (ms) of the batched pull sweep over one bucket's
 * graphs under cfg (the autotuner's probe); -1 for an invalid cfg. */
double gz_batch_time_pull(const gz_batch_config_t* cfg, uint32_t bucket, uint32_t reps);
/* PageRank ranks of the whole union (rank = props * outdeg', outdeg' = V of
 * the vertex's graph for sinks), V entries. */
void gz_batch_pr_ranks(double* out);
void gz_graph_release(void);

/* Host mirrors filled at load time (for output, statistics and tests). */
uint64_t gz_graph_num_vertices(void);
uint64_t gz_graph_num_edges(void);
uint64_t gz_graph_pull_vector_count(void);
uint64_t gz_graph_push_vector_count(void);
const double* gz_graph_outdegrees_host(void);
/* Copy the device Vector-Sparse lists back (tests). */
void gz_graph_download_vectors(int pull, gz_vec_t* out);
void gz_graph_download_offsets(int pull, uint64_t* out);

/* -------- algorithm drivers (execution_{pr,cc,bfs}.c) -------------------- */

enum gz_algorithm { GZ_ALG_PR = 0, GZ_ALG_CC = 1, GZ_ALG_BFS = 2 };
enum gz_engine_mode {
    GZ_ENGINE_HYBRID = 0,       /* fixed threshold (execution_cc.c:124-142) */
    GZ_ENGINE_PULL = 1,         /* EXPERIMENT_EDGE_FORCE_PULL */
    GZ_ENGINE_PUSH = 2,         /* EXPERIMENT_EDGE_FORCE_PUSH */
    GZ_ENGINE_LEARNED = 3,      /* MLP selector (gz_selector_*), fixed threshold when no weights */
    GZ_ENGINE_PROBE = 4         /* every iteration: time both engines from the same state (best of
                                 * 3 runs each), then continue with the fixed-threshold choice;
                                 * records carry both times and the selector features */
};

typedef struct gz_run_options_t {
    uint32_t algorithm;         /* enum gz_algorithm */
    uint32_t engine;            /* enum gz_engine_mode; PR honours PULL/PUSH only */
    uint32_t iterations;        /* PR iteration count (-N) */
    uint32_t stat_mode;         /* enum gz_stat_mode */
    uint32_t pull_variant;      /* enum gz_pull_variant */
    uint32_t edge_only;         /* EXPERIMENT_EDGE_ONLY (PR) */
    uint32_t vertex_only;       /* EXPERIMENT_VERTEX_ONLY (PR) */
    uint32_t profile;           /* record per-iteration engine / time / frontier (ITERATION_PROFILE) */
    uint32_t collect_stats;     /* record vectors / edges per iteration (ITERATION_STATS) */
    uint32_t max_iterations;    /* CC/BFS safety cap, 0 = none */
    double damping;             /* PR damping factor */
    uint64_t root;              /* BFS search root */
} gz_run_options_t;

void gz_run_options_default(gz_run_options_t* o);

typedef struct gz_iteration_record_t {
    uint64_t iteration;         /* 1-based */
    uint32_t engine;            /* 1 = pull, 2 = push */
    uint32_t pad;
    uint64_t frontier_stat;     /* converge_vote that selected the engine */
    uint64_t vectors;           /* vectors examined (collect_stats) */
    uint64_t edges;             /* valid lanes examined (collect_stats) */
    double edge_ms;             /* edge phase time (profile) */
    float features[8];          /* learned / probe engines: selector features (before normalisation) */
    float predicted[2];         /* learned engine: predicted log ms of pull, push */
    double pull_ms, push_ms;    /* probe engine: edge-phase time of each engine */
    double variant_ms[3];       /* BFS probe with gz_probe_wide(1): the pull over every vector (no
                                 * compaction of WantsInfo), the push after a compaction (never the
                                 * direct bitmap read), the push over every vector; 0 otherwise */
} gz_iteration_record_t;

typedef struct gz_run_result_t {
    uint64_t iterations;        /* total_iterations_executed */
    uint64_t pull_iterations;   /* total_iterations_used_gather */
    uint64_t push_iterations;   /* total_iterations_used_scatter */
    double elapsed_ms;          /* wall time of the iteration loop, device-synchronised */
} gz_run_result_t;

/* The BFS probe engine also times three more edge-phase variants per level
 * (gz_iteration_record_t.variant_ms), each from the same state, for bounds
 * on what a selector over a wider set of choices could gain.  Off by
 * default. */
void gz_probe_wide(int on);

/* Initialise props, accumulators and frontiers for the algorithm
 * (graph_helper_initialize_vertex_info / _create_and_initialize_frontiers). */
void gz_execution_prepare(const gz_run_options_t* o);
/* Run the iteration loop of execution_impl_{pr,cc,bfs}. */
void gz_execution_run(const gz_run_options_t* o, gz_run_result_t* r);
/* Per-iteration records of the last run when profile/collect_stats was set. */
uint64_t gz_execution_records(gz_iteration_record_t* out, uint64_t capacity);

/* Results.  PR: rank(v) = props(v) * outdeg'(v); CC: label; BFS: parent (-1
 * for root / unreached) and depth (UINT64_MAX unreached). */
void gz_result_pr_ranks(double* out);
void gz_result_pr_props(double* out);
void gz_result_cc_labels(uint64_t* out);
void gz_result_bfs(int64_t* parents, uint64_t* depths);
/* The -o text output (graphdata.c:1119-1138). */
int graph_data_write_ranks_to_file(const char* filename, uint32_t algorithm);

/* -------- breadth-first searches from many roots at once ------------------
 * No reference counterpart (the reference searches from one root,
 * execution_bfs.c:28,135).  Up to 64 searches share one pass over the graph
 * per level: every vertex keeps a 64-bit word per state (seen, frontier,
 * next) with bit i for search i, the level pushes or pulls by the reference's
 * E/5 rule on the union frontier's statistic (1 + outdeg per vertex), and the
 * depths are the BFS distances of each root, as a single search from it
 * gives them.  Needs both layouts; depths are kept as bytes, so a batch
 * fails past 254 levels.  gz_bfs_batch_run returns 0, or -1 with a message
 * (bad count or root, missing layout, too many levels); r->iterations is the
 * level count, r->elapsed_ms the wall time of the level loop. */
#define GZ_BFS_BATCH_MAX_ROOTS 64
int gz_bfs_batch_run(const uint64_t* roots, uint32_t count, gz_run_result_t* r);
/* Depths (UINT64_MAX unreached) and, when `parents` is not NULL, parents
 * (-1 for the root and unreached vertices) of search i of the last batch, in
 * original ids; the parent is the smallest in-neighbour one level up (a
 * valid BFS tree, derived from the depths on request). */
void gz_bfs_batch_result(uint32_t search, int64_t* parents, uint64_t* depths);
/* Frees the batch's device buffers (also done by gz_graph_release). */
void gz_bfs_batch_release(void);

/* Command line (main.c + cmdline.c); returns the process exit code. */
int gz_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif

#endif /* GRAZELLE_HOST_H */
