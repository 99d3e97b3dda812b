/*
 * selector.c -- host side of the learned engine selector: weight storage and
 * files.  Features, the forward pass and the decision run on the device
 * (selector_kernels.hip); the drivers read the decision together with the
 * converge vote.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grazelle_selector.h"

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
