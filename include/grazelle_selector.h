/*
 * grazelle_selector.h -- learned push/pull engine selector.
 *
 * No counterpart in the reference, which selects the engine with a fixed
 * threshold (execution_cc.c:124-142, execution_bfs.c:150-168).  The selector
 * is an 8->32->32->2 FP32 MLP over per-iteration frontier statistics; its
 * two outputs predict log edge-phase time of the pull and push engines and
 * the faster one is chosen.  Both engines compute identical results, so the
 * choice changes time only.  When no weights are loaded the drivers fall back
 * to the fixed threshold.
 */
#ifndef GRAZELLE_SELECTOR_H
#define GRAZELLE_SELECTOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GZ_SEL_IN 8
#define GZ_SEL_H 32
#define GZ_SEL_OUT 2
/* W1[32][8] b1[32] W2[32][32] b2[32] W3[2][32] b3[2] */
#define GZ_SEL_PARAMS (GZ_SEL_H * GZ_SEL_IN + GZ_SEL_H + GZ_SEL_H * GZ_SEL_H + GZ_SEL_H + GZ_SEL_OUT * GZ_SEL_H + GZ_SEL_OUT)

typedef struct gz_selector_state_t {
    uint64_t num_vertices, num_edges;
    uint64_t visited;          /* vertices that have had information (BFS: reached) */
    uint64_t last_features_valid;
    float features[GZ_SEL_IN];
} gz_selector_state_t;

/* Feature vector of one iteration (all log1p-scaled except the flags):
 *   0 log frontier vertices       1 log frontier out-edges
 *   2 log unvisited vertices      3 log unvisited in-edges (estimated E*|U|/V)
 *   4 iteration index             5 previous engine (0 none, 1 pull, 2 push)
 *   6 log previous edge-phase ms  7 graph average degree */
void gz_selector_features(float* out, uint64_t frontier_vertices, uint64_t frontier_edges, uint64_t unvisited,
                          uint64_t num_vertices, uint64_t num_edges, uint64_t iteration, uint32_t prev_engine,
                          double prev_ms);

void gz_selector_begin(gz_selector_state_t* s, uint64_t num_vertices, uint64_t num_edges);
void gz_selector_observe(gz_selector_state_t* s, uint64_t vote_before, uint64_t vote_after);
int gz_selector_loaded(void);
int gz_selector_choose_pull(gz_selector_state_t* s, uint64_t iteration, uint64_t vote, uint32_t prev_engine,
                            double prev_ms);

/* Weights: GZ_SEL_PARAMS floats plus feature normalisation (mean[8], std[8]). */
int gz_selector_set_weights(const float* params, const float* feat_mean, const float* feat_std);
void gz_selector_clear(void);
int gz_selector_load_file(const char* path);
int gz_selector_save_file(const char* path);

#ifdef __cplusplus
}
#endif

#endif /* GRAZELLE_SELECTOR_H */
