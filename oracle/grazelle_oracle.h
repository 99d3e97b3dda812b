/*
 * grazelle_oracle.h -- CPU restatement of the reference's hybrid push/pull
 * engine, used ONLY as a checker: by tests/, by __graft_entry__.smoke() and
 * as the timed CPU baseline in bench.py.  Nothing in the product links or
 * calls it.
 *
 * It restates, in portable C with pthreads, what the reference's x86
 * assembly and C compute (file:line citations at each function in
 * grazelle_oracle.c).  With threads = 1 it follows the reference's
 * single-thread (-n 1) order of operations exactly; with more threads it
 * schedules units the way scheduler_pull.inc / scheduler_push.inc do.
 *
 * The Vector-Sparse format is restated here independently of the product's
 * headers: 4 x uint64 lanes, bit 63 valid, bits 62:48 a 15/15/15/3-bit piece
 * of the shared vertex, bits 47:0 the neighbour.
 */
#ifndef GRAZELLE_ORACLE_H
#define GRAZELLE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_vec_t {
    uint64_t lane[4];
} orc_vec_t;

typedef struct orc_merge_t {
    uint64_t initial_vertex_id;
    uint64_t final_vertex_id;
    double final_partial_value;
    uint64_t pad;
} orc_merge_t;

/* ---- format and ingress --------------------------------------------------- */
uint64_t orc_shared_vertex(const orc_vec_t* v);
uint64_t orc_pack(const uint64_t* pairs, uint64_t num_edges, int by_dst, orc_vec_t* out);
void orc_vertex_index(const orc_vec_t* vecs, uint64_t nvec, uint64_t num_vertices, uint64_t* index, uint64_t* start,
                      uint64_t* end);
void orc_outdegrees(const uint64_t* pairs, uint64_t num_edges, uint64_t num_vertices, double* outdeg);
void orc_unit_range(uint64_t nvec, uint64_t units, uint64_t unit, uint64_t* first, uint64_t* last);
void orc_merge(orc_merge_t* mb, uint64_t count, double* acc);

/* ---- graph ---------------------------------------------------------------- */
typedef struct orc_graph_t {
    uint64_t num_vertices, num_edges;
    orc_vec_t* pull;
    uint64_t pull_count;
    orc_vec_t* push;
    uint64_t push_count;
    uint64_t* pull_index;     /* [V+64] reference vertex index */
    uint64_t* push_index;
    double* outdeg;           /* [V+8] */
    int borrowed;             /* pull / push lists belong to the caller */
} orc_graph_t;

/* Pair arrays as in the -pull / -push files (grouped by dst / src). */
orc_graph_t* orc_graph_create(uint64_t num_vertices, const uint64_t* pull_pairs, uint64_t pull_edges,
                              const uint64_t* push_pairs, uint64_t push_edges);
/* Graph over Vector-Sparse lists that are already packed (for instance a
 * device ingress's lists copied back to the host).  The lists are borrowed
 * and must outlive the graph.  E = valid lanes of the pull list (of the push
 * list when there is no pull list); outdegrees count each valid pull lane for
 * its neighbour, the source, as graphdata.c:792-821 counts the pull file's
 * source column; with a push list they are read from it instead, as each
 * source's valid lanes there, which is the same count in source order). */
orc_graph_t* orc_graph_from_vectors(uint64_t num_vertices, const orc_vec_t* pull, uint64_t pull_count,
                                    const orc_vec_t* push, uint64_t push_count);
void orc_graph_destroy(orc_graph_t* g);
uint64_t orc_graph_pull_count(const orc_graph_t* g);
uint64_t orc_graph_push_count(const orc_graph_t* g);

/* ---- algorithms ----------------------------------------------------------- */
enum { ORC_STAT_BOTH = 0, ORC_STAT_OUTDEG = 1, ORC_STAT_COUNT = 2 };
enum { ORC_ENGINE_HYBRID = 0, ORC_ENGINE_PULL = 1, ORC_ENGINE_PUSH = 2 };

typedef struct orc_options_t {
    uint32_t threads;          /* >= 1 */
    uint32_t units_per_thread; /* 32 in the reference (main.c:77) */
    uint32_t engine;
    uint32_t stat_mode;
    uint32_t traditional_pull; /* EXPERIMENT_EDGE_PULL_WITHOUT_SCHED_AWARE */
    uint32_t iterations;       /* PageRank */
    uint32_t max_iterations;   /* CC / BFS cap, 0 = none */
    double damping;
    uint64_t root;
} orc_options_t;

typedef struct orc_iteration_t {
    uint32_t engine;           /* 1 pull, 2 push */
    uint32_t pad;
    uint64_t vote;             /* converge_vote that selected the engine */
} orc_iteration_t;

typedef struct orc_result_t {
    uint64_t iterations, pull_iterations, push_iterations;
    double elapsed_ms;         /* CLOCK_MONOTONIC around the iteration loop */
} orc_result_t;

/* props[V] receives rank/outdeg' (the reference's props); ranks[V] (optional) rank. */
int orc_pagerank(const orc_graph_t* g, const orc_options_t* o, double* props, double* ranks, orc_result_t* r);
/* labels[V]; log (optional, capacity log_cap) receives one entry per iteration. */
int orc_cc(const orc_graph_t* g, const orc_options_t* o, uint64_t* labels, orc_iteration_t* log, uint64_t log_cap,
           orc_result_t* r);
/* As orc_cc with the labels starting from init[V] instead of the vertex ids
 * (execution_cc.c:87-90 starts from the ids).  A graph whose vertices were
 * renumbered starts from the original ids, so that every label is the
 * smallest original id that reaches the vertex.  init values must be below
 * 2^53 (labels are exact FP64 integers, as in the reference). */
int orc_cc_from(const orc_graph_t* g, const orc_options_t* o, const uint64_t* init, uint64_t* labels,
                orc_iteration_t* log, uint64_t log_cap, orc_result_t* r);
/* parents[V] (-1 root / unreached) and depths[V] (UINT64_MAX unreached). */
int orc_bfs(const orc_graph_t* g, const orc_options_t* o, int64_t* parents, uint64_t* depths, orc_iteration_t* log,
            uint64_t log_cap, orc_result_t* r);

/* BFS certificate over the pull list (a checker, no reference counterpart):
 * counts the vertices that violate what fixes BFS depths and a valid BFS tree.
 *   root: depth 0, parent -1;
 *   every other reached v: (parents[v], v) is an edge and
 *     depths[parents[v]] == depths[v] - 1;
 *   unreached v (depth UINT64_MAX): parent -1;
 *   every edge (u, v) with u reached: depths[v] <= depths[u] + 1.
 * The parent chains bound each depth from above by the distance and the
 * edge condition bounds it from below, so zero violations means the depths
 * are the BFS distances and the parents a BFS tree. */
uint64_t orc_bfs_certificate_violations(const orc_graph_t* g, uint64_t root, const int64_t* parents,
                                        const uint64_t* depths, uint32_t threads);

/* ---- learned selector MLP (8->32->32->2, FP32) ---------------------------- */
void orc_mlp_forward(const float* params, const float* x, float* y);
/* Full-batch gradient of mean squared error; returns the loss. */
float orc_mlp_gradient(const float* params, const float* X, const float* T, uint64_t n, float* grad);

#ifdef __cplusplus
}
#endif

#endif /* GRAZELLE_ORACLE_H */
