/*
 * selector.c -- host side of the learned engine selector: weight storage and
 * files.  Features, the forward pass and the decision run on the device
 * (selector_kernels.hip); the drivers read the decision together with the
 * converge vote.  The first BFS decision, whose features the host knows,
 * can also be taken here (gz_selector_host_decide).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grazelle_selector.h"
#include "gz_host_internal.h"

int gz_selector_device_upload(const float* params);
int gz_selector_device_upload_norm(const float* mean, const float* stdev);

static int gz_sel_loaded = 0;
static float gz_sel_params[GZ_SEL_PARAMS];
static float gz_sel_mean[GZ_SEL_IN];
static float gz_sel_std[GZ_SEL_IN];

int gz_selector_loaded(void) { return gz_sel_loaded; }

int gz_selector_set_weights(const float* params, const float* feat_mean, const float* feat_std)
{
    memcpy(gz_sel_params, params, sizeof(gz_sel_params));
    memcpy(gz_sel_mean, feat_mean, sizeof(gz_sel_mean));
    memcpy(gz_sel_std, feat_std, sizeof(gz_sel_std));
    gz_sel_loaded = 1;
    if (gz_selector_device_upload(gz_sel_params) != 0) return -1;
    return gz_selector_device_upload_norm(gz_sel_mean, gz_sel_std);
}

void gz_selector_clear(void) { gz_sel_loaded = 0; }

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

int gz_selector_feature_used(uint32_t i)
{
    return gz_sel_loaded && i < GZ_SEL_IN && gz_sel_std[i] > 0.f;
}

uint32_t gz_selector_host_decide(const float* x, float* y)
{
    if (!gz_sel_loaded) return 0;
    /* parameter layout W1 b1 W2 b2 W3 b3, as on the device */
    const float* W1 = gz_sel_params;
    const float* b1 = W1 + GZ_SEL_H * GZ_SEL_IN;
    const float* W2 = b1 + GZ_SEL_H;
    const float* b2 = W2 + GZ_SEL_H * GZ_SEL_H;
    const float* W3 = b2 + GZ_SEL_H;
    const float* b3 = W3 + GZ_SEL_OUT * GZ_SEL_H;
    float xn[GZ_SEL_IN], h1[GZ_SEL_H], h2[GZ_SEL_H];
    for (int i = 0; i < GZ_SEL_IN; ++i) {
        /* a feature with std 0 is unused by the model: its input is 0 */
        const float sd = gz_sel_std[i];
        xn[i] = sd > 0.f ? (x[i] - gz_sel_mean[i]) / sd : 0.f;
    }
    for (int j = 0; j < GZ_SEL_H; ++j) {
        float a = b1[j];
        for (int i = 0; i < GZ_SEL_IN; ++i) a += W1[j * GZ_SEL_IN + i] * xn[i];
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < GZ_SEL_H; ++j) {
        float a = b2[j];
        for (int i = 0; i < GZ_SEL_H; ++i) a += W2[j * GZ_SEL_H + i] * h1[i];
        h2[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < GZ_SEL_OUT; ++j) {
        float a = b3[j];
        for (int i = 0; i < GZ_SEL_H; ++i) a += W3[j * GZ_SEL_H + i] * h2[i];
        y[j] = a;
    }
    return (y[0] < y[1]) ? 1u : 2u;
}
