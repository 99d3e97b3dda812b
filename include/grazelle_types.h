/*
 * grazelle_types.h -- in-memory formats shared by the host code, the HIP
 * kernels and the C-ABI boundary.
 *
 * Vector-Sparse edge vector (reference: source/graphdata.c:143-173 encoder,
 * source/graphdata.c:114-120 decoder, source/constants.asm:51-67 masks):
 *   four little-endian uint64 lanes, 32 bytes, 32-byte aligned.
 *   lane k bit 63      : valid bit; valid lanes form a prefix (0..count-1)
 *   lane k bits 62..48 : piece k of the shared ("top-level") vertex
 *                        piece0 = v[14:0], piece1 = v[29:15],
 *                        piece2 = v[44:30], piece3 = v[47:45] (3 bits)
 *   lane k bits 47..0  : neighbour vertex id
 *   lane 3 bits 62..51 : unused 12-bit field (graphdata.c:122-135)
 * In the pull (in-edge) list the shared vertex is the destination and the
 * neighbours are sources; in the push (out-edge) list it is the reverse.
 * Invalid lanes keep whatever neighbour bits the packer left there and must
 * be ignored by every consumer.
 *
 * Merge-buffer record (reference: include/graphtypes.h:27-35), 32 bytes.
 */
#ifndef GRAZELLE_TYPES_H
#define GRAZELLE_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GZ_LANE_VALID_BIT        0x8000000000000000ull
#define GZ_LANE_NEIGHBOUR_MASK   0x0000ffffffffffffull
#define GZ_LANE_PIECE_MASK       0x7fff000000000000ull
#define GZ_LANE3_PIECE_MASK      0x0007000000000000ull
#define GZ_LANE3_UNUSED_MASK     0x7ff8000000000000ull
#define GZ_MAX_VERTEX_ID         0x0000ffffffffffffull

/* Vertex-index markers (reference: source/graphdata.c:219-267). */
#define GZ_VINDEX_NO_EDGES       0x7fffffffffffffffull
#define GZ_VINDEX_PAST_END       0xffffffffffffffffull

typedef struct gz_vec_t {
    uint64_t lane[4];
} __attribute__((aligned(32))) gz_vec_t;

typedef struct mergeaccum_t {
    uint64_t initial_vertex_id;     /* first target vertex written by the unit   */
    uint64_t final_vertex_id;       /* last target vertex touched by the unit    */
    double   final_partial_value;   /* partial value of the last target vertex   */
    uint64_t __padding__;
} mergeaccum_t;

/* One chunk slot of the source-sliced pull (no reference counterpart): the
 * `count` vectors from `begin` of the concatenated slices; `record` is the
 * chunk's merge record with its slice in the top 7 bits.  count 0 marks an
 * empty slot. */
typedef struct gz_slice_chunk_t {
    uint64_t begin;
    uint32_t count;
    uint32_t record;
} gz_slice_chunk_t;
#define GZ_SLICE_RECORD_BITS 25

/* Recover the shared vertex from its four 15/15/15/3-bit pieces. */
static inline uint64_t gz_vec_shared_vertex(const gz_vec_t* v)
{
    return ((v->lane[0] & GZ_LANE_PIECE_MASK) >> 48) |
           ((v->lane[1] & GZ_LANE_PIECE_MASK) >> 33) |
           ((v->lane[2] & GZ_LANE_PIECE_MASK) >> 18) |
           ((v->lane[3] & GZ_LANE3_PIECE_MASK) >> 3);
}

/* Number of valid lanes (valid lanes form a prefix). */
static inline unsigned gz_vec_valid_count(const gz_vec_t* v)
{
    return (unsigned)((v->lane[0] >> 63) + (v->lane[1] >> 63) + (v->lane[2] >> 63) + (v->lane[3] >> 63));
}

/* Compose one vector from a shared vertex and 1..4 neighbours.  Lanes at or
 * beyond `count` are written invalid with neighbour id 0. */
static inline void gz_vec_encode(gz_vec_t* out, uint64_t shared, const uint64_t* nbr, unsigned count)
{
    const uint64_t piece[4] = {
        (shared >> 0) & 0x7fffull,
        (shared >> 15) & 0x7fffull,
        (shared >> 30) & 0x7fffull,
        (shared >> 45) & 0x7ull,
    };
    for (unsigned k = 0; k < 4; ++k) {
        const uint64_t valid = (k < count) ? GZ_LANE_VALID_BIT : 0ull;
        const uint64_t id = (k < count) ? (nbr[k] & GZ_LANE_NEIGHBOUR_MASK) : 0ull;
        out->lane[k] = valid | (piece[k] << 48) | id;
    }
}

#ifdef __cplusplus
}
#endif

#endif /* GRAZELLE_TYPES_H */
