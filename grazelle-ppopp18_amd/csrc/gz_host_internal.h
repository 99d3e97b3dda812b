/*
 * gz_host_internal.h -- device helpers the host drivers call that are not
 * part of the reference's phase interface.
 */
#ifndef GZ_HOST_INTERNAL_H
#define GZ_HOST_INTERNAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

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
