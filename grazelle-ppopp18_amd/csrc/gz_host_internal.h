/*
 * gz_host_internal.h -- device helpers the host drivers call that are not
 * part of the reference's phase interface.
 */
#ifndef GZ_HOST_INTERNAL_H
#define GZ_HOST_INTERNAL_H

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
void gz_graph_reset_roles(void);
void gz_clear_bitmap_range(uint64_t* bitmap, uint64_t vertex_start, uint64_t vertex_count);

#ifdef __cplusplus
}
#endif

#endif /* GZ_HOST_INTERNAL_H */
