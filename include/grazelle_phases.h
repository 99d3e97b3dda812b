/*
 * grazelle_phases.h -- the drop-in boundary: the per-phase C ABI that the
 * host drivers call once per phase per iteration, implemented by hand-written
 * HIP kernels for gfx950 (libgrazelle.so).
 *
 * Reference interface replaced (include/phases.h:49-86):
 *   phase_op_reset_global_accum            phases.h:53  (asm, phases.asm:24-30)
 *   phase_op_write_global_accum_to_buf     phases.h:57  (asm, phases.asm:34-44)
 *   phase_op_combine_global_var_from_buf   phases.h:61  (C,   phases.c:27-37)
 *   edge_pull_op_merge_with_merge_buffer   phases.h:70  (C,   phases.c:41-70)
 *   perform_edge_pull_phase_{pr,cc,bfs}    phases.h:77  (asm, phase_edge_pull_*.asm)
 *   perform_edge_push_phase_{pr,cc,bfs}    phases.h:81  (asm, phase_edge_push_*.asm)
 *   perform_vertex_phase_{pr,cc,bfs}       phases.h:86  (asm, phase_vertex_*.asm)
 * The algorithm suffix is chosen by the caller, as phases.h:28-46 does with
 * macros.  The runtime boundary that threads.c / numanodes.c provided
 * (threads.h:37-82, numanodes.h:27-61) is replaced by the gz_runtime_* hooks
 * below: one host thread drives the GPU, stream order replaces barriers.
 *
 * Conventions that differ from the reference, and why:
 *   - Every pointer handed to a phase is a DEVICE pointer (edge lists, merge
 *     buffer, accumulators, reduce buffer).  The reduce buffer comes from
 *     gz_reduce_buffer_create(); only phase_op_combine_global_var_from_buf()
 *     copies its first element to the host (one synchronisation).
 *   - The implicit inputs the assembly read from graphdata globals
 *     (graphdata.inc:21-50) are read from the exported `gz_graph` record,
 *     filled by the host graph loader.  The drivers swap
 *     gz_graph.frontier_has_info and gz_graph.vertex_accumulators exactly as
 *     execution_cc.c:245-250 swaps graph_frontier_has_info and
 *     graph_vertex_accumulators.
 *   - Compile-time constants and EXPERIMENT_* switches of the reference that
 *     touch this path are runtime fields of `gz_phase_config` (damping,
 *     threshold statistic, scheduler-aware vs traditional pull).
 *   - Element types: PR props/accumulators are FP64 as in the reference.  CC
 *     labels are uint64 (the reference keeps the same integral values in
 *     FP64).  BFS props are uint64 words (depth << 48 | parent), parent
 *     0xffffffffffff meaning "none"; gz_bfs_parent() / gz_bfs_depth() decode.
 *   - Errors: phases return void; any HIP error prints a message naming the
 *     failing call and exits the process, as the reference dies on bad input
 *     (graphdata.c:963-967).
 */
#ifndef GRAZELLE_PHASES_H
#define GRAZELLE_PHASES_H

#include <stddef.h>
#include <stdint.h>

#include "grazelle_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* -------- device graph record (reference: graphdata.h:30-114 globals) ----- */

typedef struct gz_phase_graph_t {
    uint64_t num_vertices;                 /* graph_num_vertices */
    uint64_t num_edges;                    /* graph_num_edges */
    const gz_vec_t* pull_vectors;          /* graph_edges_gather_list_block_bufs_numa[0][0] */
    uint64_t pull_vector_count;            /* graph_edges_gather_list_block_counts_numa[0][0] */
    const uint64_t* pull_vertex_offsets;   /* [V+1] first pull vector of each dst; derived from graph_vertex_gather_index_numa */
    const gz_vec_t* push_vectors;          /* graph_edges_scatter_list_block_bufs_numa[0][0] */
    uint64_t push_vector_count;            /* graph_edges_scatter_list_block_counts_numa[0][0] */
    const uint64_t* push_vertex_offsets;   /* [V+1] first push vector of each src; derived from graph_vertex_scatter_index_numa */
    const double* vertex_outdegrees;       /* graph_vertex_outdegrees [V+8], FP64 */
    void* vertex_props;                    /* graph_vertex_props */
    void* vertex_accumulators;             /* graph_vertex_accumulators (PR: FP64; CC/BFS: HasInfo* bitmap) */
    uint64_t* frontier_has_info;           /* graph_frontier_has_info   [ceil(V/64)] */
    uint64_t* frontier_wants_info;         /* graph_frontier_wants_info [ceil(V/64)] */
} gz_phase_graph_t;

extern gz_phase_graph_t gz_graph;

/* -------- runtime configuration (reference: compile-time switches) -------- */

enum gz_stat_mode {                        /* phase_helper_add_frontier_stat, phasehelpers.inc:388-399 */
    GZ_STAT_COUNT_AND_OUTDEG = 0,          /* default: +1 +outdeg(v) */
    GZ_STAT_OUTDEG_ONLY = 1,               /* EXPERIMENT_THRESHOLD_WITHOUT_COUNT */
    GZ_STAT_COUNT_ONLY = 2                 /* EXPERIMENT_THRESHOLD_WITHOUT_OUTDEGREES */
};

enum gz_pull_variant {
    GZ_PULL_SCHED_AWARE = 0,               /* merge buffers, plain stores (default) */
    GZ_PULL_TRADITIONAL = 1,               /* EXPERIMENT_EDGE_PULL_WITHOUT_SCHED_AWARE: atomics, no merge */
    GZ_PULL_SOURCE_SLICED = 2,             /* over the source-sliced layout built by
                                            * gz_graph_build_source_slices for this edge list, PR
                                            * (merges inside the phase) and CC (no reference
                                            * counterpart) */
    GZ_PULL_MERGE_BUFFER = 3               /* CC: EXPERIMENT_EDGE_PULL_FORCE_MERGE -- complete runs lower
                                            * labels with plain stores, unit-boundary runs go through
                                            * the merge buffer with min (phase_edge_pull_cc.asm:136-181);
                                            * the merge runs inside the phase.  PR: as GZ_PULL_SCHED_AWARE */
};

typedef struct gz_phase_config_t {
    double damping;                        /* const_damping_factor (constants.asm:75-79), 0.875 in the reference */
    uint32_t stat_mode;                    /* enum gz_stat_mode */
    uint32_t pull_variant;                 /* enum gz_pull_variant (PR; GZ_PULL_SOURCE_SLICED also CC) */
    uint32_t chunk_vectors;                /* Edge vectors per scheduling unit (one wavefront), multiple of 64 */
    uint32_t bfs_level;                    /* BFS depth written by the current edge phase (1-based) */
    uint32_t collect_stats;                /* nonzero: count vectors / valid lanes examined (ITERATION_STATS) */
    uint32_t push_frontier_skip;           /* nonzero: push phases visit only frontier sources (default 1) */
    uint32_t vectors_per_lane;             /* pull sweeps: consecutive vectors per lane per step (1, 2 or 4) */
} gz_phase_config_t;

extern gz_phase_config_t gz_phase_config;

/* Units (merge-buffer records) the pull phase writes for a list of `vector_count`
 * vectors under the current chunk length (the GPU analogue of
 * sched_pull_units_total, main.c:75-91). */
uint64_t gz_pull_units_for(uint64_t vector_count);

/* -------- common phase operators (phases.h:49-61) ------------------------ */

void phase_op_reset_global_accum(void);
void phase_op_write_global_accum_to_buf(uint64_t* reduce_buffer);
uint64_t phase_op_combine_global_var_from_buf(uint64_t* reduce_buffer);

/* -------- edge-pull engine operators (phases.h:64-70) -------------------- */

void edge_pull_op_merge_with_merge_buffer(mergeaccum_t* merge_buffer, uint64_t count, double* vertex_accumulators);

/* -------- phase control functions (phases.h:73-86) ----------------------- */

void perform_edge_pull_phase_pr(const gz_vec_t* edge_list, uint64_t edge_list_count);
void perform_edge_push_phase_pr(const gz_vec_t* edge_list, uint64_t edge_list_count);
void perform_vertex_phase_pr(uint64_t vertex_start, uint64_t vertex_count, const uint64_t* reduce_buffer);

void perform_edge_pull_phase_cc(const gz_vec_t* edge_list, uint64_t edge_list_count);
void perform_edge_push_phase_cc(const gz_vec_t* edge_list, uint64_t edge_list_count);
void perform_vertex_phase_cc(uint64_t vertex_start, uint64_t vertex_count, const uint64_t* reduce_buffer);

void perform_edge_pull_phase_bfs(const gz_vec_t* edge_list, uint64_t edge_list_count);
void perform_edge_push_phase_bfs(const gz_vec_t* edge_list, uint64_t edge_list_count);
void perform_vertex_phase_bfs(uint64_t vertex_start, uint64_t vertex_count, const uint64_t* reduce_buffer);

/* The merge buffer the pull phase writes into (graph_vertex_merge_buffer,
 * graphdata.c:1053-1072); allocated by gz_merge_buffer_allocate(). */
extern mergeaccum_t* gz_merge_buffer;
extern uint64_t gz_merge_buffer_count;
void gz_merge_buffer_allocate(uint64_t units);

/* Valid-lane / vector counters of the last edge phase when
 * gz_phase_config.collect_stats is set (graph_stat_num_*_per_thread,
 * phasehelpers.inc:403-428).  Synchronises. */
void gz_phase_stats_read(uint64_t* vectors, uint64_t* edges);

/* -------- BFS word helpers ------------------------------------------------ */

#define GZ_BFS_PARENT_NONE 0x0000ffffffffffffull
#define GZ_BFS_UNVISITED   0xffffffffffffffffull
static inline uint64_t gz_bfs_depth(uint64_t word) { return word >> 48; }
static inline int64_t gz_bfs_parent(uint64_t word)
{
    const uint64_t p = word & GZ_BFS_PARENT_NONE;
    return (p == GZ_BFS_PARENT_NONE) ? -1 : (int64_t)p;
}

/* -------- one PageRank iteration that allocates nothing ------------------
 * No counterpart in the reference.  For callers that cannot call hipMalloc:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation.  It is one iteration of the PageRank driver loop
 * (execution_pr.c:99-158) with the kernels of perform_edge_pull_phase_pr
 * (scheduler-aware), edge_pull_op_merge_with_merge_buffer,
 * phase_op_write_global_accum_to_buf and perform_vertex_phase_pr, over
 * caller-owned device arrays, on the library stream (gz_runtime_stream):
 *   pull_vectors[pull_vector_count]  dst-grouped Vector-Sparse list
 *   outdegrees[V]                    FP64 out-degrees (0 for sinks)
 *   props[V]                         rank / outdeg' in, the next iteration's out
 *   accumulators[V]                  the pull's per-destination sums; 0.0 on the
 *                                    first call for vertices without in-edges,
 *                                    which the pull never writes (as after
 *                                    gz_execution_prepare)
 * with the damping and chunk length of gz_phase_config.  Results are
 * bit-identical to those entry points' for the same list and settings.  Its
 * scratch (merge records and partial sums, 40 bytes per pull unit, for at
 * most 524,288 units, i.e. 512 x 2^20 vectors at 1024 vectors per unit; about
 * 20 MiB in all) is one file-scope device array that each GPU the library
 * runs on holds until the process ends.  Returns 0, or -1 with a message
 * when the list needs more units (raise gz_phase_config.chunk_vectors) or the
 * runtime has no scratch address.  Calls that overlap on one GPU race on the
 * shared scratch; calls on one stream are ordered. */
int gz_pagerank_iteration_noalloc(const gz_vec_t* pull_vectors, uint64_t pull_vector_count, uint64_t num_vertices,
                                  const double* outdegrees, double* props, double* accumulators);

/* -------- runtime hooks (replace threads.c / numanodes.c) ---------------- */

int gz_runtime_init(int device);               /* 0 on success; dies if no device */
void gz_runtime_shutdown(void);
int gz_runtime_device_count(void);             /* 0 when no GPU is visible; never dies */
const char* gz_runtime_device_name(void);
void* gz_device_malloc(size_t bytes);          /* zero-filled, like numanodes_malloc's mmap */
void gz_device_free(void* ptr);
void gz_copy_to_device(void* dst, const void* src, size_t bytes);
void gz_copy_to_host(void* dst, const void* src, size_t bytes);
void gz_device_zero(void* ptr, size_t bytes);  /* fill kernel, safe under graph capture */
uint64_t gz_device_allocation_events(void);    /* device allocations plus frees the library made so far */
void gz_device_copy(void* dst, const void* src, size_t bytes); /* device-to-device, stream ordered */
void gz_synchronize(void);
void* gz_runtime_stream(void);                 /* hipStream_t every phase is launched on */
uint64_t* gz_reduce_buffer_create(void);       /* device reduce buffer, 8 entries */
void gz_reduce_buffer_destroy(uint64_t* reduce_buffer);

/* Per-kernel launch timing with HIP events on the phase stream (off by
 * default).  Each phase brackets its main kernel; totals are kept per id. */
enum gz_kernel_id {
    GZ_K_PULL_PR = 0, GZ_K_PUSH_PR, GZ_K_VERTEX_PR, GZ_K_MERGE, GZ_K_PULL_CC, GZ_K_PUSH_CC,
    GZ_K_PULL_BFS_FIND, GZ_K_PULL_BFS_VISIT, GZ_K_PUSH_BFS, GZ_K_COMPACT, GZ_K_REDUCE, GZ_K_COUNT
};
void gz_kernel_timing_enable(int on);
int gz_kernel_timing_enabled(void);
void gz_kernel_timing_reset(void);
/* Synchronises; total milliseconds and launches recorded for `id`. */
void gz_kernel_timing_read(int id, double* total_ms, uint64_t* launches);

/* Stream-ordered timer built on HIP events. */
typedef struct gz_timer_t gz_timer_t;
gz_timer_t* gz_timer_create(void);
void gz_timer_destroy(gz_timer_t* t);
void gz_timer_start(gz_timer_t* t);
void gz_timer_stop(gz_timer_t* t);
float gz_timer_elapsed_ms(gz_timer_t* t);      /* synchronises on the stop event */

#ifdef __cplusplus
}
#endif

#endif /* GRAZELLE_PHASES_H */
