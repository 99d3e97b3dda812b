/*
 * selector.c -- host side of the learned engine selector: feature
 * construction, weight storage and the per-iteration decision.  The MLP
 * forward pass that makes the decision runs on the device
 * (gz_selector_device_forward, selector_kernels.hip).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grazelle_selector.h"

int gz_selector_device_forward(const float* x, float* y);
int gz_selector_device_upload(const float* params);

static int gz_sel_loaded = 0;
static float gz_sel_params[GZ_SEL_PARAMS];
static float gz_sel_mean[GZ_SEL_IN];
static float gz_sel_std[GZ_SEL_IN];

void gz_selector_features(float* out, uint64_t frontier_vertices, uint64_t frontier_edges, uint64_t unvisited,
                          uint64_t num_vertices, uint64_t num_edges, uint64_t iteration, uint32_t prev_engine,
                          double prev_ms)
{
    const double avg = num_vertices ? (double)num_edges / (double)num_vertices : 0.0;
    out[0] = (float)log1p((double)frontier_vertices);
    out[1] = (float)log1p((double)frontier_edges);
    out[2] = (float)log1p((double)unvisited);
    out[3] = (float)log1p((double)unvisited * avg);
    out[4] = (float)iteration;
    out[5] = (float)prev_engine;
    out[6] = (float)log1p(prev_ms > 0.0 ? prev_ms : 0.0);
    out[7] = (float)avg;
}

void gz_selector_begin(gz_selector_state_t* s, uint64_t num_vertices, uint64_t num_edges)
{
    memset(s, 0, sizeof(*s));
    s->num_vertices = num_vertices;
    s->num_edges = num_edges;
}

void gz_selector_observe(gz_selector_state_t* s, uint64_t vote_before, uint64_t vote_after)
{
    (void)vote_before;
    (void)vote_after;
    s->last_features_valid = 0;
}

int gz_selector_loaded(void) { return gz_sel_loaded; }

int gz_selector_set_weights(const float* params, const float* feat_mean, const float* feat_std)
{
    memcpy(gz_sel_params, params, sizeof(gz_sel_params));
    memcpy(gz_sel_mean, feat_mean, sizeof(gz_sel_mean));
    memcpy(gz_sel_std, feat_std, sizeof(gz_sel_std));
    gz_sel_loaded = 1;
    return gz_selector_device_upload(gz_sel_params);
}

void gz_selector_clear(void) { gz_sel_loaded = 0; }

int gz_selector_choose_pull(gz_selector_state_t* s, uint64_t iteration, uint64_t vote, uint32_t prev_engine,
                            double prev_ms)
{
    /* frontier statistics: the vote counts vertices plus their out-edges;
     * split it with the graph's average degree */
    const double avg = s->num_vertices ? (double)s->num_edges / (double)s->num_vertices : 0.0;
    const uint64_t fv = (uint64_t)((double)vote / (1.0 + avg));
    const uint64_t fe = vote > fv ? vote - fv : 0;
    const uint64_t unvisited = s->num_vertices > s->visited ? s->num_vertices - s->visited : 0;
    float x[GZ_SEL_IN], y[GZ_SEL_OUT];
    gz_selector_features(x, fv, fe, unvisited, s->num_vertices, s->num_edges, iteration, prev_engine, prev_ms);
    for (int i = 0; i < GZ_SEL_IN; ++i) x[i] = (x[i] - gz_sel_mean[i]) / (gz_sel_std[i] > 0.f ? gz_sel_std[i] : 1.f);
    memcpy(s->features, x, sizeof(x));
    s->last_features_valid = 1;
    s->visited += fv;
    gz_selector_device_forward(x, y);
    return y[0] < y[1];
}

int gz_selector_load_file(const char* path)
{
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    float p[GZ_SEL_PARAMS], m[GZ_SEL_IN], sd[GZ_SEL_IN];
    int ok = fread(p, sizeof(float), GZ_SEL_PARAMS, f) == GZ_SEL_PARAMS && fread(m, sizeof(float), GZ_SEL_IN, f) == GZ_SEL_IN &&
             fread(sd, sizeof(float), GZ_SEL_IN, f) == GZ_SEL_IN;
    fclose(f);
    if (!ok) return -2;
    return gz_selector_set_weights(p, m, sd);
}

int gz_selector_save_file(const char* path)
{
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    int ok = fwrite(gz_sel_params, sizeof(float), GZ_SEL_PARAMS, f) == GZ_SEL_PARAMS &&
             fwrite(gz_sel_mean, sizeof(float), GZ_SEL_IN, f) == GZ_SEL_IN &&
             fwrite(gz_sel_std, sizeof(float), GZ_SEL_IN, f) == GZ_SEL_IN;
    fclose(f);
    return ok ? 0 : -2;
}
