/*
 * gz_host_internal.h -- device helpers the host drivers call that are not
 * part of the reference's phase interface.
 */
#ifndef GZ_HOST_INTERNAL_H
#define GZ_HOST_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "grazelle_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device-built layouts handed to the graph record (device ingress). */
typedef struct gz_device_layouts_t {
    uint64_t num_vertices, num_edges;
    gz_vec_t* pull;
    uint64_t* pull_offsets;      /* [V+1] */
    uint64_t pull_count;
    gz_vec_t* push;
    uint64_t* push_offsets;      /* [V+1] */
    uint64_t push_count;
    double* outdeg;              /* [V+8] */
    uint64_t* original_ids;      /* [V+8] original id of each device vertex, or NULL (not renumbered) */
} gz_device_layouts_t;

/* Take ownership of device layouts and allocate the vertex arrays. */
void gz_graph_adopt_device(const gz_device_layouts_t* layouts);
/* Device array of original ids by device vertex, NULL when not renumbered. */
const uint64_t* gz_graph_original_ids_device(void);

void gz_dev_init_pr(double* props, double* acc, const double* outdeg, uint64_t n);
void gz_dev_fill_u64(uint64_t* p, uint64_t value, uint64_t n);
void gz_dev_iota_u64(uint64_t* p, uint64_t n);
void gz_dev_store_u64(uint64_t* p, uint64_t index, uint64_t value);
/* A BFS start from device vertex `root` in one launch: words[0..n) unvisited
 * but the root's (no parent), acc[0..fwords) = 0, has = {root}, wants = all
 * but root. */
void gz_dev_bfs_reset(uint64_t* words, uint64_t n, uint64_t* acc, uint64_t* has, uint64_t* wants, uint64_t fwords,
                      uint64_t root);
void gz_graph_reset_roles(void);
void gz_dev_indegrees(const gz_vec_t* vecs, uint64_t count, int pull_layout, uint32_t* indeg, uint64_t n);
/* In-degrees of the loaded graph (u32[V+8]), built on first use. */
const uint32_t* gz_graph_indegrees_device(void);
/* In-degree of device vertex v from a host copy of the array above (made on
 * first use). */
uint32_t gz_graph_indegree_host(uint64_t v);
/* The loaded selector's forward pass on the host, in the device kernel's
 * order of operations: features x[8] -> predictions y[2]; returns 1 (pull)
 * or 2 (push), 0 without weights. */
uint32_t gz_selector_host_decide(const float* x, float* y);
/* Launch one graph of `times` rounds of the launches enqueue(ctx) makes on
 * the phase stream and return 1.  The graph is kept while key[0..n), `times`
 * and the device allocations stay the same.  It is captured the second time
 * in a row the same request comes; the first time nothing is launched and 0
 * tells the caller to run the rounds itself (enqueue must not allocate or
 * wait). */
int gz_graph_replay_cached(const uint64_t* key, size_t n, void (*enqueue)(void*), void* ctx, uint32_t times);
/* The same in cache slot `slot` (0: PageRank iterations, 1: CC iterations). */
int gz_graph_replay_cached_slot(uint32_t slot, const uint64_t* key, size_t n, void (*enqueue)(void*), void* ctx,
                                uint32_t times);
/* Frontier statistic of the sources the next CC push will read (the vote that
 * built its frontier); a small one lets the push skip the compaction. */
void gz_cc_push_vote_hint(uint64_t vote);
/* The one-workgroup push tail (gz_cc_push_tail): whether a frontier of
 * statistic `vote` qualifies, and its launch, which compacts HasInfo, runs up
 * to max_rounds push iterations while their statistic stays at most
 * vote_cap, and writes the last vote to reduce_buffer[0] and the rounds run
 * to reduce_buffer[7]. */
int gz_cc_push_tail_applies(uint64_t vote);
void gz_cc_launch_push_tail(uint64_t vote_cap, uint64_t max_rounds, uint64_t* reduce_buffer);
/* Control words of a device-driven iteration loop (CC, and the log of
 * (engine, vote) pairs from GZ_LOOP_LOG on). */
#define GZ_LOOP_ENGINE 0          /* engine of the iteration about to run: 0 stop, 1 pull, 2 push over the
                                   * compacted frontier, 3 / 4 the small-frontier push reading 64 / 16
                                   * bitmap words per wavefront (gz_loop_engine) */
#define GZ_LOOP_VOTE 1            /* the vote that chose it */
#define GZ_LOOP_ITERATIONS 2
#define GZ_LOOP_PULLS 3
#define GZ_LOOP_PUSHES 4
#define GZ_LOOP_THRESHOLD 5
#define GZ_LOOP_MODE 6            /* enum gz_engine_mode: hybrid, pull or push */
#define GZ_LOOP_MAX_ITERATIONS 7  /* 0: none */
#define GZ_LOOP_LOG_CAP 8
#define GZ_LOOP_PULL_NEXT 9       /* next pull chunk to hand out (cleared after each iteration) */
#define GZ_LOOP_SPARSE_OK 10      /* the vote bounds the frontier's vectors: small-frontier pushes allowed */
#define GZ_LOOP_LOG 16
#define GZ_LOOP_WORDS(cap) (GZ_LOOP_LOG + 2 * (cap))

/* The CC iteration loop driven by control words on the device (gz_cc.hip):
 * the vote partials each engine writes, and one iteration's launches over
 * the current frontier roles. */
uint64_t gz_cc_loop_waves(uint64_t* pull_waves, uint64_t* push_waves);
/* The engine code of an iteration chosen by `vote` (pull: 1, push: 2..4),
 * used by the host before the loop and by the loop's control kernel. */
#ifdef __HIPCC__
#define GZ_HOST_DEVICE __host__ __device__
#else
#define GZ_HOST_DEVICE
#endif
static inline GZ_HOST_DEVICE uint64_t gz_loop_engine(uint64_t vote, int pull, uint64_t sparse_ok)
{
    if (pull) return 1;
    if (!sparse_ok || vote > (1u << 16)) return 2;
    return vote <= 4096 ? 3 : 4;
}
void gz_cc_loop_enqueue(uint64_t* ctrl);
/* gz_scratch_reserve for host C code: per-wave buffers for `waves`
 * wavefronts and vertex-sized buffers for V. */
void gz_scratch_reserve_public(uint64_t waves, uint64_t num_vertices);
/* The same for the next BFS push (the vote of the level that built its
 * frontier). */
void gz_bfs_push_vote_hint(uint64_t vote);
/* Drop the frontier compaction's per-vertex vector counts (graph release). */
void gz_frontier_counts_release(void);
/* Frees the BFS bottom-up step's pending-destination buffers. */
void gz_bfs_pending_release(void);
/* The PageRank push's fixed-point accumulators: freed with the graph, and
 * cleared when a PageRank run starts. */
void gz_pr_release(void);
/* Every PageRank iteration of the loaded graph in one cooperative launch
 * (the scheduler-aware pull, merge, sum and vertex phase); returns 1 when it
 * ran, 0 when it is switched off, does not apply or cannot launch
 * (gz_pr_persistent), -1 when a grid barrier gave up (the vertex arrays
 * then hold a partial run). */
int gz_pr_run_persistent(uint64_t iterations, uint64_t* reduce_buffer);
/* The merge, the sum of S and the PageRank vertex phase in one launch after
 * a scheduler-aware pull of the loaded graph's pull list: applies when it
 * has at most 4096 chunks (gz_pr_fused_tail_applies); the per-vertex run
 * heads are built once per run from the first pull's records
 * (gz_pr_run_heads_build); the result is bit-identical to the merge, reduce
 * and vertex launches. */
int gz_pr_fused_tail_applies(void);
void gz_pr_run_heads_build(void);
void gz_pr_vertex_merged(uint64_t* reduce_buffer);
void gz_pr_clear_fixed(void);
/* The source-sliced pull's settings: hot cache, cold spread, cold stream. */
void gz_slices_settings(uint32_t* out3);

/* The 8 reduce-buffer words as read by the last
 * phase_op_combine_global_var_from_buf (pinned host memory). */
const uint64_t* gz_reduce_words(void);
/* The pinned copy no longer matches a reduce buffer another writer changed
 * (gz_runtime.hip). */
void gz_reduce_mirror_stale(void);
/* phase_op_combine_global_var_from_buf for the library's own drivers: reads
 * the pinned copy without a copy when the vote kernels wrote it for this
 * buffer. */
uint64_t gz_combine_vote(uint64_t* reduce_buffer);

/* Learned selector on the device (selector_kernels.hip). */
enum { GZ_SEL_FROM_BITMAP = 0, GZ_SEL_FROM_EDGE_PHASE = 1, GZ_SEL_FROM_ROOT = 2 };
/* root < V stages a one-vertex initial frontier (BFS); pass V for none. */
void gz_sel_dev_begin(uint64_t num_vertices, uint64_t num_edges, uint64_t root, const uint32_t* indeg,
                      int root_removed);
void gz_sel_dev_stamp(int by_compaction);
/* BFS with the learned engine deciding on the host: the last edge phase's
 * vote and frontier sums into the reduce words (0: vote; 2, 3, 4: activated
 * vertices, their out-edges, their in-edges; 5: the phase's device-clock ms
 * as a float). */
void gz_sel_dev_sums(uint64_t* reduce_buffer, const uint32_t* indeg);
void gz_sel_dev_step(const uint64_t* frontier, const uint32_t* indeg, uint64_t* reduce_buffer, int have_model,
                     int bfs, uint64_t iteration, uint32_t prev_engine, float prev_ms, int source, int with_vote);
/* BFS edge phases fold the selector's features (vertices, out-edges and,
 * with `indeg`, in-edges of the vertices they activate) when `on`. */
void gz_features_fold(int on, const uint32_t* indeg);
/* Free the source-sliced layout (graph release). */
void gz_source_slices_release(void);
void gz_clear_bitmap_range(uint64_t* bitmap, uint64_t vertex_start, uint64_t vertex_count);

#ifdef __cplusplus
}
#endif

#endif /* GZ_HOST_INTERNAL_H */
