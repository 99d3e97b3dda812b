/*
 * grazelle_selector.h -- learned push/pull engine selector.
 *
 * No counterpart in the reference, which selects the engine with a fixed
 * threshold (execution_cc.c:124-142, execution_bfs.c:150-168).  The selector
 * is an 8->32->32->2 FP32 MLP over per-iteration frontier statistics; its
 * two outputs predict the log edge-phase time (ms) of the pull and push
 * engines and the faster one is chosen.  Both engines compute identical
 * results, so the choice changes time only.  When no weights are loaded the
 * drivers fall back to the fixed threshold.
 *
 * Features of iteration i (computed on the device after iteration i-1's
 * edge phase, from the bitmap of the vertices it activated):
 *   0 log(1 + frontier vertices)      1 log(1 + frontier out-edges)
 *   2 log(1 + unvisited vertices)     3 log(1 + unvisited in-edges)
 *   4 iteration index (1-based)       5 previous engine (0 none, 1 pull, 2 push)
 *   6 log(1 + previous edge-phase ms) 7 average degree E / V
 * "Unvisited" counts vertices never in a frontier (BFS); for CC, which
 * revisits vertices, it is V and E.  The network sees (x - mean) / std.
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

int gz_selector_loaded(void);
/* Whether the loaded model uses input feature i (0..7).  A feature whose
 * normalisation std is 0 in the weights is unused: the model sees 0 there,
 * and for feature 3 (unvisited in-edges) the BFS edge phases then skip the
 * per-vertex in-degree reads that feed it. */
int gz_selector_feature_used(uint32_t i);
/* The learned engine follows the model only while the frontier statistic
 * lies in [threshold / below, above * threshold], the threshold being the
 * fixed rule's (E/5, or V/2 with counts only); outside, the fixed rule's
 * choice stands (pull above, push below), so that a model is not trusted
 * where its training graphs gave it no reason to differ from the rule.
 * Defaults 1 and 32: on the training graphs the push never beat the pull
 * above the threshold, and the pull often beat the push down to about a
 * 32nd of it.  A factor below 1 lifts that side of the band. */
void gz_selector_set_trust(double above, double below);

/* Weights: GZ_SEL_PARAMS floats plus feature normalisation (mean[8], std[8]). */
int gz_selector_set_weights(const float* params, const float* feat_mean, const float* feat_std);
void gz_selector_clear(void);
/* File: params, mean, std as little-endian FP32, nothing else. */
int gz_selector_load_file(const char* path);
int gz_selector_save_file(const char* path);

/* Device kernels (hand-written, selector_kernels.hip), host arrays in/out. */
int gz_selector_forward_batch(const float* params, const float* X, float* Y, uint64_t n);
/* Full-batch gradient of the mean squared error over n samples and both
 * outputs; returns the loss. */
float gz_selector_gradient(const float* params, const float* X, const float* T, uint64_t n, float* grad);
/* Full-batch Adam, `steps` steps on the device; params updated in place. */
int gz_selector_train(float* params, const float* X, const float* T, uint64_t n, uint32_t steps, float lr,
                      float* loss_history);

#ifdef __cplusplus
}
#endif

#endif /* GRAZELLE_SELECTOR_H */
