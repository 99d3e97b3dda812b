/*
 * gz_walk.h -- the wavefront chunk walker shared by the pull sweeps
 * (PageRank sum, CC min).
 *
 * A wavefront owns the vectors [b, e) of one scheduling unit.  Each step covers
 * 64*VPL vectors: lane L holds VPL consecutive vectors (one 32*VPL-byte
 * contiguous slice, so a step reads 2*VPL KiB with every byte used) and the
 * next step's slice is loaded before the current one is reduced.
 *
 * Reduction of a step, keyed by destination ("runs" of equal destinations):
 *   1. lane-local: the lane folds its VPL values; runs that start and end
 *      inside the lane (interior runs) are complete and stored at once; the
 *      lane keeps its head run (h_d, h_s) and tail run (t_d, t_s), which are
 *      the same run when the lane holds a single destination;
 *   2. cross-lane: an inclusive segmented scan of t_s keyed by t_d joins the
 *      tails of lanes with the single-run lanes that follow them; a lane whose
 *      head continues the previous lane's tail adds the previous lane's
 *      scanned value to its head;
 *   3. the run still open at the last valid lane is carried to the next step,
 *      and a carried run that the next step does not continue is stored.
 * The run open when the chunk ends is returned to the caller (merge record or
 * direct atomic), together with the chunk's first destination.
 *
 * Op supplies: T identity(); T combine(T, T); T gather(const vec&, bool
 * in_range, uint32_t& edges) (identity when !in_range) for the list's vector
 * type; void store(uint64_t d, T value) for a complete run.  The list is a
 * Vector-Sparse list (gz_vs_list) or a packed copy (gz_packed_list,
 * gz_packed24_list).
 */
#ifndef GZ_WALK_H
#define GZ_WALK_H

#include "gz_internal.h"

/* Cross-lane moves of 64-bit values through DPP (data-parallel primitives:
 * VALU moves that read another lane of the same instruction, no LDS round
 * trip).  Lanes whose source does not exist keep `old`. */
#define GZ_DPP_ROW_SHR(n) (0x110 + (n)) /* lane i <- lane i-n of its 16-lane row */
#define GZ_DPP_ROW_BCAST15 0x142        /* rows 1..3 <- lane 15 of the row before */
#define GZ_DPP_ROW_BCAST31 0x143        /* rows 2..3 <- lane 31 */
#define GZ_DPP_WAVE_SHR1 0x138          /* lane i <- lane i-1 */
#define GZ_DPP_WAVE_ROL1 0x134          /* lane i <- lane i+1, lane 63 <- lane 0 */

template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t gz_dpp_u64(uint64_t old, uint64_t x)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)old, (int)(uint32_t)x, CTRL, ROWS, 0xf,
                                                              false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(old >> 32), (int)(uint32_t)(x >> 32),
                                                              CTRL, ROWS, 0xf, false);
    return ((uint64_t)hi << 32) | lo;
}

template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ T gz_dpp(T x);

template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t gz_dpp_impl(uint64_t x)
{
    return gz_dpp_u64<CTRL, ROWS>(x, x);
}

template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ T gz_dpp(T x)
{
    if constexpr (sizeof(T) == 8 && !__is_same(T, uint64_t)) {
        return __longlong_as_double((long long)gz_dpp_impl<CTRL, ROWS>((uint64_t)__double_as_longlong(x)));
    } else {
        return gz_dpp_impl<CTRL, ROWS>(x);
    }
}

/* The value of x in lane `src` (src uniform across the wavefront). */
__device__ __forceinline__ uint64_t gz_readlane_u64(uint64_t x, int src)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

template <typename T>
__device__ __forceinline__ T gz_readlane_t(T x, int src)
{
    if constexpr (!__is_same(T, uint64_t)) {
        return __longlong_as_double((long long)gz_readlane_u64((uint64_t)__double_as_longlong(x), src));
    } else {
        return gz_readlane_u64(x, src);
    }
}

/* The same moves through LDS permutes (ds_bpermute): fewer VALU
 * instructions, a longer dependent latency.  The plain pull sweeps, which run
 * many wavefronts per SIMD, are a little faster this way; the LDS-cached
 * sliced sweep, whose LDS port also serves its gathers, uses DPP. */
template <typename T>
__device__ __forceinline__ T gz_shfl_up_t(T x, int o)
{
    return __shfl_up(x, o, GZ_WAVE);
}

template <>
__device__ __forceinline__ uint64_t gz_shfl_up_t<uint64_t>(uint64_t x, int o)
{
    return (uint64_t)__shfl_up((unsigned long long)x, o, GZ_WAVE);
}

template <typename T>
__device__ __forceinline__ T gz_shfl_t(T x, int src)
{
    return __shfl(x, src, GZ_WAVE);
}

template <>
__device__ __forceinline__ uint64_t gz_shfl_t<uint64_t>(uint64_t x, int src)
{
    return (uint64_t)__shfl((unsigned long long)x, src, GZ_WAVE);
}

/* Inclusive segmented scan across the wavefront: S_i = combine of the values
 * of lanes [seg_i, i], seg_i being the first lane of lane i's segment.  Within
 * each 16-lane row a Hillis-Steele scan over row_shr 1, 2, 4, 8 (a lane only
 * takes a partner in its own row and segment); then rows 1 and 3 take lane 15
 * and lane 47, and rows 2 and 3 take lane 31, when their segment reaches back
 * past their row or half. */
template <typename T, typename Op>
__device__ __forceinline__ T gz_segscan_dpp(T S, int lane, int seg, const Op& op)
{
    const int rl = lane & 15;
    {
        const T x = gz_dpp<GZ_DPP_ROW_SHR(1), 0xf>(S);
        if (rl >= 1 && lane - 1 >= seg) S = op.combine(S, x);
    }
    {
        const T x = gz_dpp<GZ_DPP_ROW_SHR(2), 0xf>(S);
        if (rl >= 2 && lane - 2 >= seg) S = op.combine(S, x);
    }
    {
        const T x = gz_dpp<GZ_DPP_ROW_SHR(4), 0xf>(S);
        if (rl >= 4 && lane - 4 >= seg) S = op.combine(S, x);
    }
    {
        const T x = gz_dpp<GZ_DPP_ROW_SHR(8), 0xf>(S);
        if (rl >= 8 && lane - 8 >= seg) S = op.combine(S, x);
    }
    {
        const T x = gz_dpp<GZ_DPP_ROW_BCAST15, 0xa>(S);
        if (((lane >> 4) & 1) && seg < (lane & ~15)) S = op.combine(S, x);
    }
    {
        const T x = gz_dpp<GZ_DPP_ROW_BCAST31, 0xc>(S);
        if (lane >= 32 && seg < 32) S = op.combine(S, x);
    }
    return S;
}

/* Where the walker reads its vectors.  A list supplies load(i), the
 * destination of the chunk's first vector, and the destinations of a step's
 * vectors (dsts), with a cursor carried from step to step.
 *
 * A Vector-Sparse list: each vector names its destination. */
struct gz_vs_list {
    const gz_vec_t* __restrict__ vecs;
    typedef gz_lanes vec;
    struct cursor {};
    __device__ __forceinline__ vec load(uint64_t i) const { return gz_load_vector(vecs, i); }
    __device__ __forceinline__ uint64_t first_dst(uint64_t b, cursor&) const
    {
        return gz_shared_vertex(load(b));
    }
    template <int VPL>
    __device__ __forceinline__ void dsts(const vec (&v)[VPL], uint64_t, uint64_t, uint64_t, cursor&,
                                         uint64_t (&d)[VPL]) const
    {
#pragma unroll
        for (int k = 0; k < VPL; ++k) d[k] = gz_shared_vertex(v[k]);
    }
};

/* Packed lists (gz_packed_vec, gz_internal.h) whose destinations are the
 * dense ranks 0, 1, 2, ... in order: a flag on a vector's first neighbour
 * marks the first vector of a destination, so a vector's destination is the
 * chunk's first destination plus the flags after the chunk's first vector up
 * to it.  `shared` holds every vector's destination but is read only at chunk
 * starts.  Vector i of the concatenated slices is either 16 bytes (four
 * 32-bit ids, flag in bit 31 of the first) or, inside a source window, 12
 * bytes (four 24-bit offsets from the window's first vertex, flag in bit 23 of
 * the first); the flagged first lane always holds an edge. */
#define GZ_PACKED_FIRST 0x80000000u
#define GZ_PACKED24_NONE 0xffffffu
#define GZ_PACKED24_FIRST 0x800000u

struct gz_flagged_dsts {
    const uint32_t* __restrict__ shared;
    struct cursor {
        uint64_t d;
    };
    __device__ __forceinline__ uint64_t first_dst(uint64_t b, cursor& c) const
    {
        c.d = shared[b];
        return c.d;
    }
    template <int VPL>
    __device__ __forceinline__ void dsts(const gz_packed_vec (&v)[VPL], uint64_t my_first, uint64_t b, uint64_t e,
                                         cursor& c, uint64_t (&d)[VPL]) const
    {
        /* flags counted: in range and not the chunk's first vector */
        uint32_t f[VPL], t = 0;
#pragma unroll
        for (int k = 0; k < VPL; ++k) {
            const uint64_t i = my_first + (uint64_t)k;
            f[k] = (i < e && i != b) ? v[k].d : 0u;
            t += f[k];
        }
        /* exclusive prefix of the lanes' counts (at most VPL <= 4: 3 bits) */
        const uint64_t below = (1ull << gz_lane()) - 1ull;
        uint64_t before = 0, total = 0;
#pragma unroll
        for (int bit = 0; bit < 3; ++bit) {
            if ((1 << bit) > VPL) break;
            const uint64_t bal = __ballot((t >> bit) & 1u);
            before += (uint64_t)__popcll(bal & below) << bit;
            total += (uint64_t)__popcll(bal) << bit;
        }
        uint64_t x = c.d + before;
#pragma unroll
        for (int k = 0; k < VPL; ++k) {
            x += f[k];
            d[k] = x;
        }
        c.d += total;
    }
};

/* 16-byte vectors; nbr[i - first] holds vector i. */
struct gz_packed_list : gz_flagged_dsts {
    const gz_u32x4* __restrict__ nbr;
    uint64_t first;
    typedef gz_packed_vec vec;
    __device__ __forceinline__ vec load(uint64_t i) const
    {
        const gz_u32x4 a = __builtin_nontemporal_load(nbr + (i - first));
        vec v;
        v.s0 = a.x & ~GZ_PACKED_FIRST;
        v.s1 = a.y;
        v.s2 = a.z;
        v.s3 = a.w;
        v.d = a.x >> 31; /* the flag; the destination comes from dsts() */
        return v;
    }
};

/* 12-byte vectors of one window: words w[3j .. 3j+2], j = i - first, hold
 * the offsets f0 | f1 << 24 | f2 << 48 | f3 << 72 (little-endian, 96 bits). */
struct gz_packed24_list : gz_flagged_dsts {
    const uint32_t* __restrict__ w;
    uint32_t window; /* the window's first vertex */
    uint64_t first;  /* vector index stored at w[0] */
    typedef gz_packed_vec vec;
    __device__ __forceinline__ vec load(uint64_t i) const
    {
        const uint64_t j = i - first;
        const uint32_t w0 = __builtin_nontemporal_load(w + 3 * j);
        const uint32_t w1 = __builtin_nontemporal_load(w + 3 * j + 1);
        const uint32_t w2 = __builtin_nontemporal_load(w + 3 * j + 2);
        const uint32_t f0 = w0 & 0xffffffu;
        const uint32_t f1 = (w0 >> 24) | ((w1 & 0xffffu) << 8);
        const uint32_t f2 = (w1 >> 16) | ((w2 & 0xffu) << 16);
        const uint32_t f3 = w2 >> 8;
        vec v;
        v.s0 = window + (f0 & ~GZ_PACKED24_FIRST);
        v.s1 = (f1 == GZ_PACKED24_NONE) ? GZ_PACKED_NONE : window + f1;
        v.s2 = (f2 == GZ_PACKED24_NONE) ? GZ_PACKED_NONE : window + f2;
        v.s3 = (f3 == GZ_PACKED24_NONE) ? GZ_PACKED_NONE : window + f3;
        v.d = f0 >> 23;
        return v;
    }
};

/* 8-byte vectors of one LDS block: w[i - first] holds four 16-bit offsets
 * from the block's first vertex (GZ_PACKED16_NONE for an empty lane), bit 15
 * of the first set on the first vector of each destination. */
struct gz_packed16_list : gz_flagged_dsts {
    const uint64_t* __restrict__ w;
    uint32_t window; /* the block's first vertex */
    uint64_t first;
    typedef gz_packed_vec vec;
    __device__ __forceinline__ vec load(uint64_t i) const
    {
        const uint64_t x = __builtin_nontemporal_load(w + (i - first));
        const uint32_t f0 = (uint32_t)x & 0xffffu, f1 = (uint32_t)(x >> 16) & 0xffffu,
                       f2 = (uint32_t)(x >> 32) & 0xffffu, f3 = (uint32_t)(x >> 48);
        vec v;
        v.s0 = window + (f0 & ~GZ_PACKED16_FIRST);
        v.s1 = (f1 == GZ_PACKED16_NONE) ? GZ_PACKED_NONE : window + f1;
        v.s2 = (f2 == GZ_PACKED16_NONE) ? GZ_PACKED_NONE : window + f2;
        v.s3 = (f3 == GZ_PACKED16_NONE) ? GZ_PACKED_NONE : window + f3;
        v.d = f0 >> 15;
        return v;
    }
};

template <typename L, int VPL>
struct gz_slice {
    typename L::vec v[VPL];
};

template <typename L, int VPL>
__device__ __forceinline__ void gz_load_slice(gz_slice<L, VPL>& s, const L& list, uint64_t first, uint64_t e)
{
    /* unconditional loads (indices past the chunk end re-read its last vector
     * and are masked by the caller): no branch, so the wait counts after the
     * loads stay exact */
#pragma unroll
    for (int k = 0; k < VPL; ++k) {
        const uint64_t i = first + (uint64_t)k;
        s.v[k] = list.load((i < e) ? i : e - 1);
    }
}

/* One step of the walk: reduce the slice `cur` (vectors base + lane*VPL ...)
 * and prefetch the following step's slice into `nxt`.  kPrefetchFirst issues
 * the prefetch before the gathers, for operators whose gathers end in a wait
 * for every load in flight (a select between memory and LDS values): the
 * prefetch then overlaps the gathers instead of starting after them. */
template <int VPL, typename T, bool kDpp, bool kPrefetchFirst = false, typename L, typename Op>
__device__ __forceinline__ void gz_walk_step(const L& list, uint64_t b, uint64_t e, uint64_t base, Op& op,
                                             const gz_slice<L, VPL>& cur, gz_slice<L, VPL>& nxt,
                                             typename L::cursor& cursor, uint64_t& carry_d, T& carry_v,
                                             uint32_t& edges)
{
    const int lane = gz_lane();
    const uint64_t step = (uint64_t)GZ_WAVE * VPL;
    const uint64_t my_first = base + (uint64_t)lane * VPL;
    /* issue every gather of the step first, then the next slice's loads:
     * waiting for the gathers then leaves the prefetch in flight */
    T gv[VPL];
    if constexpr (kPrefetchFirst) gz_load_slice<L, VPL>(nxt, list, my_first + step, e);
#pragma unroll
    for (int k = 0; k < VPL; ++k) gv[k] = op.gather(cur.v[k], my_first + (uint64_t)k < e, edges);
    if constexpr (!kPrefetchFirst) gz_load_slice<L, VPL>(nxt, list, my_first + step, e);
    uint64_t dd[VPL];
    list.template dsts<VPL>(cur.v, my_first, b, e, cursor, dd);
    /* 1. lane-local runs */
    int nvalid = 0;
    uint64_t h_d = ~0ull, t_d = ~0ull;
    T h_s = op.identity(), t_s = op.identity();
#pragma unroll
    for (int k = 0; k < VPL; ++k) {
        if (my_first + (uint64_t)k >= e) break;
        const uint64_t d = dd[k];
        const T g = gv[k];
        if (nvalid == 0) {
            h_d = t_d = d;
            h_s = t_s = g;
        } else if (d == t_d) {
            t_s = op.combine(t_s, g);
            if (t_d == h_d) h_s = t_s;
        } else {
            if (t_d != h_d) op.store(t_d, t_s); /* interior run, complete */
            t_d = d;
            t_s = g;
        }
        ++nvalid;
    }
    const bool has = nvalid > 0;
    const bool single = has && (h_d == t_d);
    /* a carried run the step does not continue is complete */
    const uint64_t h0 = gz_readlane_u64(h_d, 0);
    if (lane == 0 && base != b && h0 != carry_d) op.store(carry_d, carry_v);
    /* 2. cross-lane segmented scan of tails keyed by t_d */
    uint64_t prev_t;
    if constexpr (kDpp)
        prev_t = gz_dpp<GZ_DPP_WAVE_SHR1, 0xf>(t_d);
    else
        prev_t = gz_shfl_up_t<uint64_t>(t_d, 1);
    const uint64_t heads = __ballot(lane == 0 || t_d != prev_t);
    const int seg = gz_segment_start(heads, lane);
    T S = t_s;
    if constexpr (kDpp) {
        S = gz_segscan_dpp<T>(t_s, lane, seg, op);
    } else {
#pragma unroll
        for (int o = 1; o < GZ_WAVE; o <<= 1) {
            const T x = gz_shfl_up_t<T>(S, o);
            if (lane - o >= seg) S = op.combine(S, x);
        }
    }
    /* lanes whose scan segment starts at lane 0 and continues the carry */
    if (seg == 0 && t_d == carry_d && has) S = op.combine(S, carry_v);
    T S_prev;
    if constexpr (kDpp)
        S_prev = gz_dpp<GZ_DPP_WAVE_SHR1, 0xf>(S);
    else
        S_prev = gz_shfl_up_t<T>(S, 1);
    /* the head of a multi-run lane completes inside the lane */
    if (has && !single) {
        T head = h_s;
        if (lane == 0) {
            if (h_d == carry_d) head = op.combine(head, carry_v);
        } else if (prev_t == h_d) {
            head = op.combine(head, S_prev);
        }
        op.store(h_d, head);
    }
    /* 3. tails: complete when the next lane starts another destination */
    uint64_t next_h;
    if constexpr (kDpp)
        next_h = gz_dpp<GZ_DPP_WAVE_ROL1, 0xf>(h_d);
    else
        next_h = gz_shfl_t<uint64_t>(h_d, (lane + 1) & (GZ_WAVE - 1));
    const int last = __popcll(__ballot(has)) - 1; /* valid lanes form a prefix */
    if (has && lane != last && next_h != t_d) op.store(t_d, S);
    carry_d = gz_readlane_u64(t_d, last);
    carry_v = gz_readlane_t<T>(S, last);
}

template <int VPL, typename T, bool kDpp, bool kPrefetchFirst = false, typename L, typename Op>
__device__ __forceinline__ void gz_walk_list(const L& list, uint64_t b, uint64_t e, Op& op, uint64_t& first_d,
                                             uint64_t& carry_d, T& carry_v, uint32_t& edges)
{
    const int lane = gz_lane();
    const uint64_t step = (uint64_t)GZ_WAVE * VPL;
    typename L::cursor cursor;
    first_d = list.first_dst(b, cursor);
    carry_d = first_d;
    carry_v = op.identity();
    /* two slice buffers used in turn, so that no register copy forces the
     * prefetched loads to complete early */
    gz_slice<L, VPL> sa, sb;
    gz_load_slice<L, VPL>(sa, list, b + (uint64_t)lane * VPL, e);
    uint64_t base = b;
    while (true) {
        gz_walk_step<VPL, T, kDpp, kPrefetchFirst>(list, b, e, base, op, sa, sb, cursor, carry_d, carry_v, edges);
        base += step;
        if (base >= e) break;
        gz_walk_step<VPL, T, kDpp, kPrefetchFirst>(list, b, e, base, op, sb, sa, cursor, carry_d, carry_v, edges);
        base += step;
        if (base >= e) break;
    }
}

/* The walk over a Vector-Sparse list. */
template <int VPL, typename T, bool kDpp = false, typename Op>
__device__ __forceinline__ void gz_walk_chunk(const gz_vec_t* __restrict__ vecs, uint64_t b, uint64_t e, Op& op,
                                              uint64_t& first_d, uint64_t& carry_d, T& carry_v, uint32_t& edges)
{
    const gz_vs_list list{vecs};
    gz_walk_list<VPL, T, kDpp>(list, b, e, op, first_d, carry_d, carry_v, edges);
}

#endif /* GZ_WALK_H */
