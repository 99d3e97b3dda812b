/*
 * gz_pr_ops.h -- the PageRank pull gather shared by the single-graph and the
 * batched pull sweeps (phase_edge_pull_pr.asm:91-171).
 */
#ifndef GZ_PR_OPS_H
#define GZ_PR_OPS_H

#include "gz_internal.h"

/* One property gather; kStream marks it nontemporal, so that gathers known
 * to miss do not evict lines that other gathers reuse. */
template <bool kStream>
__device__ __forceinline__ double gz_gather_prop(const double* __restrict__ props, uint64_t i)
{
    if (kStream) return __builtin_nontemporal_load(props + i);
    return props[i];
}

/* Sum of the gathered properties of one vector, in the lane order of the
 * asm's vhaddpd / vextractf128 / vaddpd sequence: (g2+g3)+(g0+g1). */
template <bool kStream = false>
__device__ __forceinline__ double gz_pr_vector_sum(const gz_lanes& v, const double* __restrict__ props, uint32_t& edges,
                                                   bool in_range = true)
{
    /* branch-free: invalid lanes read element 0 and are zeroed, so the four
     * loads issue back to back */
    const bool v0 = in_range && gz_lane_valid(v.l0), v1 = in_range && gz_lane_valid(v.l1),
               v2 = in_range && gz_lane_valid(v.l2), v3 = in_range && gz_lane_valid(v.l3);
    double g0 = gz_gather_prop<kStream>(props, v0 ? gz_lane_id(v.l0) : 0);
    double g1 = gz_gather_prop<kStream>(props, v1 ? gz_lane_id(v.l1) : 0);
    double g2 = gz_gather_prop<kStream>(props, v2 ? gz_lane_id(v.l2) : 0);
    double g3 = gz_gather_prop<kStream>(props, v3 ? gz_lane_id(v.l3) : 0);
    g0 = v0 ? g0 : 0.0;
    g1 = v1 ? g1 : 0.0;
    g2 = v2 ? g2 : 0.0;
    g3 = v3 ? g3 : 0.0;
    edges += (uint32_t)v0 + (uint32_t)v1 + (uint32_t)v2 + (uint32_t)v3;
    return (g2 + g3) + (g0 + g1);
}

/* The same sum for a packed vector (empty lanes hold GZ_PACKED_NONE). */
template <bool kStream = false>
__device__ __forceinline__ double gz_pr_packed_sum(const gz_packed_vec& v, const double* __restrict__ props,
                                                   uint32_t& edges, bool in_range = true)
{
    const bool v0 = in_range && v.s0 != GZ_PACKED_NONE, v1 = in_range && v.s1 != GZ_PACKED_NONE,
               v2 = in_range && v.s2 != GZ_PACKED_NONE, v3 = in_range && v.s3 != GZ_PACKED_NONE;
    double g0 = gz_gather_prop<kStream>(props, v0 ? v.s0 : 0);
    double g1 = gz_gather_prop<kStream>(props, v1 ? v.s1 : 0);
    double g2 = gz_gather_prop<kStream>(props, v2 ? v.s2 : 0);
    double g3 = gz_gather_prop<kStream>(props, v3 ? v.s3 : 0);
    g0 = v0 ? g0 : 0.0;
    g1 = v1 ? g1 : 0.0;
    g2 = v2 ? g2 : 0.0;
    g3 = v3 ? g3 : 0.0;
    edges += (uint32_t)v0 + (uint32_t)v1 + (uint32_t)v2 + (uint32_t)v3;
    return (g2 + g3) + (g0 + g1);
}

template <bool kStream>
struct gz_pr_pull_op_t {
    const double* __restrict__ props;
    double* __restrict__ acc;
    bool sched_aware;
    double sum;
    __device__ __forceinline__ double identity() const { return 0.0; }
    __device__ __forceinline__ double combine(double a, double b) const { return a + b; }
    __device__ __forceinline__ double gather(const gz_lanes& v, bool in_range, uint32_t& edges)
    {
        const double g = gz_pr_vector_sum<kStream>(v, props, edges, in_range);
        sum += g;
        return g;
    }
    __device__ __forceinline__ double gather(const gz_packed_vec& v, bool in_range, uint32_t& edges)
    {
        const double g = gz_pr_packed_sum<kStream>(v, props, edges, in_range);
        sum += g;
        return g;
    }
    __device__ __forceinline__ void store(uint64_t d, double s)
    {
        if (sched_aware)
            acc[d] = s;
        else
            unsafeAtomicAdd(&acc[d], s);
    }
};

typedef gz_pr_pull_op_t<false> gz_pr_pull_op;

#endif /* GZ_PR_OPS_H */
