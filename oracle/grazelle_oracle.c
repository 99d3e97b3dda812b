/*
 * grazelle_oracle.c -- CPU restatement of the reference engine.  TEST
 * INFRASTRUCTURE ONLY (see grazelle_oracle.h): a checker for the HIP path and
 * the timed CPU baseline, never part of the product.
 *
 * Every function cites the reference file:line whose semantics it restates.
 * Threads run SPMD exactly like the reference's workers (threads.c): each
 * runs the algorithm driver, phases are separated by barriers, thread 0 does
 * the serial steps (merge, frontier swap).
 */
#define _GNU_SOURCE
#include "grazelle_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define VALID_BIT 0x8000000000000000ull
#define NBR_MASK 0x0000ffffffffffffull
#define INDEX_NO_EDGES 0x7fffffffffffffffull
#define INDEX_PAST_END 0xffffffffffffffffull

/* ---- format (graphdata.c:114-120, :143-173) -------------------------------- */

uint64_t orc_shared_vertex(const orc_vec_t* v)
{
    return ((v->lane[0] & 0x7fff000000000000ull) >> 48) | ((v->lane[1] & 0x7fff000000000000ull) >> 33) |
           ((v->lane[2] & 0x7fff000000000000ull) >> 18) | ((v->lane[3] & 0x0007000000000000ull) >> 3);
}

static void orc_compose(orc_vec_t* out, uint64_t shared, const uint64_t* ids, unsigned count)
{
    const uint64_t piece[4] = {(shared & 0x7fffull), (shared & 0x3fff8000ull) >> 15, (shared & 0x1fffc0000000ull) >> 30,
                               (shared & 0xe00000000000ull) >> 45};
    for (unsigned k = 0; k < 4; ++k)
        out->lane[k] = ((count > k ? 1ull : 0ull) << 63) | (piece[k] << 48) | (count > k ? (ids[k] & NBR_MASK) : 0ull);
}

/* graphdata.c:618-701 (pull) and :705-788 (push): stash of up to 4
 * neighbours, flushed when the shared vertex changes or the stash is full. */
uint64_t orc_pack(const uint64_t* pairs, uint64_t num_edges, int by_dst, orc_vec_t* out)
{
    uint64_t ids[4] = {0, 0, 0, 0}, stash_vertex = 0, n = 0;
    unsigned count = 0;
    for (uint64_t i = 0; i < num_edges; ++i) {
        const uint64_t src = pairs[2 * i], dst = pairs[2 * i + 1];
        const uint64_t top = by_dst ? dst : src, other = by_dst ? src : dst;
        if ((0 != count && stash_vertex != top) || 4 == count) {
            orc_compose(&out[n++], stash_vertex, ids, count);
            count = 0;
        }
        stash_vertex = top;
        ids[count++] = other;
    }
    if (count) orc_compose(&out[n++], stash_vertex, ids, count);
    return n;
}

/* graphdata.c:219-267 */
void orc_vertex_index(const orc_vec_t* vecs, uint64_t nvec, uint64_t num_vertices, uint64_t* index, uint64_t* start,
                      uint64_t* end)
{
    const uint64_t count = num_vertices + 64;
    if (nvec == 0) {
        for (uint64_t i = 0; i < count; ++i) index[i] = INDEX_PAST_END;
        *start = *end = 0;
        return;
    }
    uint64_t cur = orc_shared_vertex(&vecs[0]), last;
    for (uint64_t i = 0; i < cur; ++i) index[i] = INDEX_NO_EDGES;
    index[cur] = 0;
    last = cur;
    *start = cur;
    for (uint64_t i = 1; i < nvec; ++i) {
        cur = orc_shared_vertex(&vecs[i]);
        if (last != cur) {
            while (last < cur - 1) index[++last] = INDEX_NO_EDGES;
            index[cur] = i;
            last = cur;
        }
    }
    *end = cur;
    for (uint64_t i = cur + 1; i < count; ++i) index[i] = INDEX_PAST_END;
}

/* graphdata.c:792-821: outdeg[src] += 1.0 over the pull file */
void orc_outdegrees(const uint64_t* pairs, uint64_t num_edges, uint64_t num_vertices, double* outdeg)
{
    for (uint64_t i = 0; i < num_vertices + 8; ++i) outdeg[i] = 0.0;
    for (uint64_t i = 0; i < num_edges; ++i) outdeg[pairs[2 * i]] += 1.0;
}

/* scheduler_pull.inc:94-138: base = q*u + min(r,u), length q + (u < r) */
void orc_unit_range(uint64_t nvec, uint64_t units, uint64_t unit, uint64_t* first, uint64_t* last)
{
    const uint64_t q = nvec / units, r = nvec % units;
    const uint64_t base = q * unit + (unit < r ? unit : r);
    *first = base;
    *last = base + q + (unit < r ? 1 : 0);
}

/* phases.c:41-70 with SCALAR_REDUCE_OP = + (execution.h:59) */
void orc_merge(orc_merge_t* mb, uint64_t count, double* acc)
{
    uint64_t i = 0, j = 0;
    while (i < count && !(mb[i].initial_vertex_id & VALID_BIT)) {
        double proposed = mb[i].final_partial_value;
        for (j = i + 1; j < count && mb[j].final_vertex_id == mb[i].final_vertex_id; ++j)
            proposed = proposed + mb[j].final_partial_value;
        if (j < count && mb[j].initial_vertex_id == mb[i].final_vertex_id) proposed = proposed + acc[mb[j].initial_vertex_id];
        acc[mb[i].final_vertex_id] = proposed;
        i = j;
    }
}

/* ---- graph ------------------------------------------------------------------ */

orc_graph_t* orc_graph_create(uint64_t num_vertices, const uint64_t* pull_pairs, uint64_t pull_edges,
                              const uint64_t* push_pairs, uint64_t push_edges)
{
    orc_graph_t* g = (orc_graph_t*)calloc(1, sizeof(*g));
    g->num_vertices = num_vertices;
    g->num_edges = pull_pairs ? pull_edges : push_edges;
    g->outdeg = (double*)malloc((num_vertices + 8) * sizeof(double));
    orc_outdegrees(pull_pairs ? pull_pairs : push_pairs, g->num_edges, num_vertices, g->outdeg);
    uint64_t s, e;
    if (pull_pairs) {
        g->pull = (orc_vec_t*)malloc((pull_edges ? pull_edges : 1) * sizeof(orc_vec_t));
        g->pull_count = orc_pack(pull_pairs, pull_edges, 1, g->pull);
        g->pull_index = (uint64_t*)malloc((num_vertices + 64) * sizeof(uint64_t));
        orc_vertex_index(g->pull, g->pull_count, num_vertices, g->pull_index, &s, &e);
    }
    if (push_pairs) {
        g->push = (orc_vec_t*)malloc((push_edges ? push_edges : 1) * sizeof(orc_vec_t));
        g->push_count = orc_pack(push_pairs, push_edges, 0, g->push);
        g->push_index = (uint64_t*)malloc((num_vertices + 64) * sizeof(uint64_t));
        orc_vertex_index(g->push, g->push_count, num_vertices, g->push_index, &s, &e);
    }
    return g;
}

orc_graph_t* orc_graph_from_vectors(uint64_t num_vertices, const orc_vec_t* pull, uint64_t pull_count,
                                    const orc_vec_t* push, uint64_t push_count)
{
    orc_graph_t* g = (orc_graph_t*)calloc(1, sizeof(*g));
    g->num_vertices = num_vertices;
    g->borrowed = 1;
    g->pull = (orc_vec_t*)pull;
    g->pull_count = pull ? pull_count : 0;
    g->push = (orc_vec_t*)push;
    g->push_count = push ? push_count : 0;
    g->outdeg = (double*)calloc(num_vertices + 8, sizeof(double));
    uint64_t s, e, pull_edges = 0, push_edges = 0;
    if (pull) {
        for (uint64_t i = 0; i < pull_count; ++i)
            for (int k = 0; k < 4; ++k) {
                if (!(pull[i].lane[k] & VALID_BIT)) continue;
                pull_edges += 1;
                if (!push) g->outdeg[pull[i].lane[k] & NBR_MASK] += 1.0;
            }
        g->pull_index = (uint64_t*)malloc((num_vertices + 64) * sizeof(uint64_t));
        orc_vertex_index(g->pull, g->pull_count, num_vertices, g->pull_index, &s, &e);
    }
    if (push) {
        /* the same per-source edge counts, read in source order */
        for (uint64_t i = 0; i < push_count; ++i) {
            const uint64_t src = orc_shared_vertex(&push[i]);
            for (int k = 0; k < 4; ++k)
                if (push[i].lane[k] & VALID_BIT) {
                    g->outdeg[src] += 1.0;
                    push_edges += 1;
                }
        }
        g->push_index = (uint64_t*)malloc((num_vertices + 64) * sizeof(uint64_t));
        orc_vertex_index(g->push, g->push_count, num_vertices, g->push_index, &s, &e);
    }
    g->num_edges = pull ? pull_edges : push_edges;
    return g;
}

void orc_graph_destroy(orc_graph_t* g)
{
    if (!g) return;
    if (!g->borrowed) {
        free(g->pull);
        free(g->push);
    }
    free(g->pull_index);
    free(g->push_index);
    free(g->outdeg);
    free(g);
}

uint64_t orc_graph_pull_count(const orc_graph_t* g) { return g->pull_count; }
uint64_t orc_graph_push_count(const orc_graph_t* g) { return g->push_count; }

/* ---- shared run state --------------------------------------------------------- */

typedef struct orc_run_t {
    const orc_graph_t* g;
    const orc_options_t* o;
    uint32_t T;
    uint64_t units;
    double* props;
    double* acc;                 /* PR accumulators, or the HasInfo* bitmap (CC/BFS) */
    uint64_t* has;
    uint64_t* wants;
    uint64_t* depth;             /* BFS visit level (derived, not in the reference) */
    uint64_t frontier_words;
    orc_merge_t* merge;
    double* sbuf;                /* per-thread PageRank partial sums (reduce buffer) */
    uint64_t* vbuf;              /* per-thread converge_vote partials */
    uint64_t counter;            /* dynamic unit counter (scheduler_pull.inc:37-52) */
    pthread_barrier_t barrier;
    uint64_t level;
    int use_pull;
    uint64_t vote;
    int done;
    orc_iteration_t* log;
    uint64_t log_cap;
    orc_result_t* res;
} orc_run_t;

static inline void orc_barrier(orc_run_t* R)
{
    if (R->T > 1) pthread_barrier_wait(&R->barrier);
}

static inline uint64_t orc_next_unit(orc_run_t* R)
{
    return __atomic_fetch_add(&R->counter, 1, __ATOMIC_RELAXED);
}

/* CAS-add on a double (phase_edge_pull_pr.asm:152-168, phase_edge_push_pr.asm lock cmpxchg) */
static inline void orc_atomic_add_double(double* p, double v, uint32_t T)
{
    if (T == 1) {
        *p = *p + v;
        return;
    }
    uint64_t* q = (uint64_t*)p;
    uint64_t old = __atomic_load_n(q, __ATOMIC_RELAXED), nw;
    do {
        double d;
        memcpy(&d, &old, 8);
        d = d + v;
        memcpy(&nw, &d, 8);
    } while (!__atomic_compare_exchange_n(q, &old, nw, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}

/* CAS-min on a non-negative double; returns 1 when the value dropped
 * (phase_edge_pull_cc.asm:183-233, phase_edge_push_cc.asm lock cmpxchg). */
static inline int orc_atomic_min_double(double* p, double v, uint32_t T)
{
    if (T == 1) {
        if (v < *p) {
            *p = v;
            return 1;
        }
        return 0;
    }
    uint64_t* q = (uint64_t*)p;
    uint64_t old = __atomic_load_n(q, __ATOMIC_RELAXED), nw;
    memcpy(&nw, &v, 8);
    while (1) {
        double d;
        memcpy(&d, &old, 8);
        if (!(v < d)) return 0;
        if (__atomic_compare_exchange_n(q, &old, nw, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return 1;
    }
}

/* byte-granular lock or / lock and (phasehelpers.inc:60-132) */
static inline void orc_bit_set(uint64_t* bm, uint64_t v) { __atomic_fetch_or(&bm[v >> 6], 1ull << (v & 63), __ATOMIC_RELAXED); }
static inline void orc_bit_clear(uint64_t* bm, uint64_t v) { __atomic_fetch_and(&bm[v >> 6], ~(1ull << (v & 63)), __ATOMIC_RELAXED); }
static inline int orc_bit_test(const uint64_t* bm, uint64_t v) { return (int)((__atomic_load_n(&bm[v >> 6], __ATOMIC_RELAXED) >> (v & 63)) & 1ull); }

/* phasehelpers.inc:388-399 */
static inline uint64_t orc_stat(const orc_run_t* R, uint64_t v)
{
    uint64_t s = 0;
    if (R->o->stat_mode != ORC_STAT_OUTDEG) s += 1;
    if (R->o->stat_mode != ORC_STAT_COUNT) s += (uint64_t)llrint(R->g->outdeg[v]);
    return s;
}

/* Strong frontier check, phasehelpers.inc:137-213.  Returns 1 when vertex
 * vid is active; otherwise sets *next to the first vector of the next active
 * vertex that owns vectors and returns 0, or returns -1 when none is left. */
static int orc_strong_check(const uint64_t* frontier, uint64_t words, const uint64_t* index, uint64_t vid,
                            uint64_t* next)
{
    uint64_t w = vid >> 6;
    unsigned bit = (unsigned)(vid & 63);
    if (w < words && orc_bit_test(frontier, vid)) return 1;
    if (w >= words) return -1;
    uint64_t word = __atomic_load_n(&frontier[w], __ATOMIC_RELAXED);
    while (1) {
        const uint64_t keep = (bit == 63) ? 0ull : ~((2ull << bit) - 1ull);
        word &= keep;
        while (word == 0) {
            w += 1;
            if (w >= words) return -1;
            word = __atomic_load_n(&frontier[w], __ATOMIC_RELAXED);
        }
        bit = (unsigned)__builtin_ctzll(word);
        const uint64_t cand = w * 64 + bit;
        const uint64_t idx = index[cand];
        if (idx & 0x8000000000000000ull) return -1;
        if (idx & 0x4000000000000000ull) continue;
        *next = idx;
        return 0;
    }
}

/* ---- PageRank phases -------------------------------------------------------- */

/* phase_edge_pull_pr.asm:45-171 (one vector) / :189-307 (unit loop) */
static void orc_pull_pr(orc_run_t* R, uint32_t tid)
{
    const orc_vec_t* L = R->g->pull;
    const uint64_t N = R->g->pull_count;
    double S = 0.0;
    for (uint64_t u = tid; u < R->units; u = orc_next_unit(R)) {
        uint64_t first, last;
        orc_unit_range(N, R->units, u, &first, &last);
        if (first >= last) continue;
        uint64_t prev = orc_shared_vertex(&L[first]);
        if (!R->o->traditional_pull) R->merge[u].initial_vertex_id = prev;
        double acc_run = 0.0;
        for (uint64_t i = first; i < last; ++i) {
            const orc_vec_t* v = &L[i];
            double gv[4] = {0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < 4; ++k)
                if (v->lane[k] & VALID_BIT) gv[k] = R->props[v->lane[k] & NBR_MASK];
            const double g = (gv[2] + gv[3]) + (gv[0] + gv[1]);
            const uint64_t d = orc_shared_vertex(v);
            if (!R->o->traditional_pull) {
                if (prev != d) {
                    R->acc[prev] = acc_run;
                    S = S + acc_run;
                    acc_run = 0.0;
                }
                prev = d;
                acc_run = g + acc_run;
            } else {
                orc_atomic_add_double(&R->acc[d], g, R->T);
                S = S + g;
            }
        }
        if (!R->o->traditional_pull) {
            S = S + acc_run;
            R->merge[u].final_vertex_id = prev;
            R->merge[u].final_partial_value = acc_run;
        }
    }
    R->sbuf[tid] = S;
}

/* phase_edge_push_pr.asm:42-354 / :360-425: msg = props[src]; every valid
 * lane CAS-adds msg into acc[dst] and adds msg to S. */
static void orc_push_pr(orc_run_t* R, uint32_t tid)
{
    const orc_vec_t* L = R->g->push;
    const uint64_t N = R->g->push_count;
    const uint64_t units = (uint64_t)R->T << 5;
    double S = 0.0;
    for (uint64_t u = tid; u < units; u = orc_next_unit(R)) {
        uint64_t first, last;
        orc_unit_range(N, units, u, &first, &last);
        for (uint64_t i = first; i < last; ++i) {
            const orc_vec_t* v = &L[i];
            const double msg = R->props[orc_shared_vertex(v)];
            for (int k = 0; k < 4; ++k) {
                if (v->lane[k] & VALID_BIT) {
                    orc_atomic_add_double(&R->acc[v->lane[k] & NBR_MASK], msg, R->T);
                    S = S + msg;
                }
            }
        }
    }
    R->sbuf[tid] = S;
}

/* phase_vertex_pr.asm:27-72 (S from the reduce buffer, 4 lanes then
 * vhaddpd) and :75-134 (per vertex, separate add/mul/add/div). */
static void orc_vertex_pr(orc_run_t* R, uint32_t tid, int zero_acc)
{
    const uint64_t V = R->g->num_vertices;
    const double nv = (double)V;
    double lanes[4] = {0.0, 0.0, 0.0, 0.0};
    const uint32_t loads = 1 + (R->T - 1) / 4;
    for (uint32_t k = 0; k < loads; ++k)
        for (int j = 0; j < 4; ++j) {
            const uint32_t t = 4 * k + (uint32_t)j;
            lanes[j] = lanes[j] + (t < R->T ? R->sbuf[t] : 0.0);
        }
    const double S = (lanes[2] + lanes[3]) + (lanes[0] + lanes[1]);
    const double c = (1.0 - S) / nv;
    const double d = R->o->damping;
    const double k = (1.0 - d) / nv;
    const uint64_t first = V * tid / R->T, last = V * (tid + 1) / R->T;
    for (uint64_t v = first; v < last; ++v) {
        double od = R->g->outdeg[v];
        if (od == 0.0) od = nv;
        double a = R->acc[v];
        if (zero_acc) R->acc[v] = 0.0;
        a = a + c;
        a = a * d;
        a = a + k;
        R->props[v] = a / od;
    }
}

/* ---- CC phases ---------------------------------------------------------------- */

static inline uint64_t orc_frontier_limit(uint64_t index_end, uint64_t words)
{
    const uint64_t lim = ((index_end + 1) >> 6) + 1;
    return lim < words ? lim : words;
}

/* phase_edge_pull_cc.asm:50-325 (vector step), :357-392 (final write),
 * :398-529 (unit loop); default build: CAS at every destination change. */
static void orc_pull_cc(orc_run_t* R, uint32_t tid)
{
    const orc_vec_t* L = R->g->pull;
    const uint64_t N = R->g->pull_count;
    uint64_t vote = 0;
    for (uint64_t u = tid; u < R->units; u = orc_next_unit(R)) {
        uint64_t first, last;
        orc_unit_range(N, R->units, u, &first, &last);
        if (first >= last) continue;
        uint64_t prev = orc_shared_vertex(&L[first]);
        double acc_run = INFINITY;
        for (uint64_t i = first; i < last; ++i) {
            const orc_vec_t* v = &L[i];
            double gmin = INFINITY;
            for (int k = 0; k < 4; ++k)
                if (v->lane[k] & VALID_BIT) {
                    const double x = R->props[v->lane[k] & NBR_MASK];
                    if (x < gmin) gmin = x;
                }
            const uint64_t d = orc_shared_vertex(v);
            if (prev != d) {
                if (orc_atomic_min_double(&R->props[prev], acc_run, R->T)) {
                    orc_bit_set((uint64_t*)R->acc, prev);
                    vote += orc_stat(R, prev);
                }
                acc_run = INFINITY;
            }
            prev = d;
            if (gmin < acc_run) acc_run = gmin;
        }
        if (orc_atomic_min_double(&R->props[prev], acc_run, R->T)) {
            orc_bit_set((uint64_t*)R->acc, prev);
            vote += orc_stat(R, prev);
        }
    }
    R->vbuf[tid] = vote;
}

/* phase_edge_push_cc.asm:51-573 (vector step), :579-673 (unit loop with the
 * strong frontier check and resume at max(previous position, unit base)). */
static void orc_push_cc(orc_run_t* R, uint32_t tid)
{
    const orc_vec_t* L = R->g->push;
    const uint64_t N = R->g->push_count;
    const uint64_t units = (uint64_t)R->T << 5;
    const uint64_t end = N ? orc_shared_vertex(&L[N - 1]) : 0;
    const uint64_t words = orc_frontier_limit(end, R->frontier_words);
    uint64_t vote = 0, pos = 0;
    for (uint64_t u = tid; u < units; u = orc_next_unit(R)) {
        uint64_t first, last;
        orc_unit_range(N, units, u, &first, &last);
        if (pos > last) continue;
        uint64_t i = pos > first ? pos : first;
        for (; i < last; ++i) {
            const orc_vec_t* v = &L[i];
            const uint64_t s = orc_shared_vertex(v);
            uint64_t next;
            const int c = orc_strong_check(R->has, words, R->g->push_index, s, &next);
            if (c < 0) {
                R->vbuf[tid] = vote;
                return;
            }
            if (c == 0) {
                i = next - 1;
                continue;
            }
            const double msg = R->props[s];
            for (int k = 0; k < 4; ++k) {
                if (!(v->lane[k] & VALID_BIT)) continue;
                const uint64_t d = v->lane[k] & NBR_MASK;
                if (orc_atomic_min_double(&R->props[d], msg, R->T)) {
                    orc_bit_set(R->has, d);                 /* asynchronous HasInfo update */
                    orc_bit_set((uint64_t*)R->acc, d);      /* HasInfo* */
                    vote += orc_stat(R, d);
                }
            }
        }
        pos = i;
    }
    R->vbuf[tid] = vote;
}

/* ---- BFS phases --------------------------------------------------------------- */

/* phase_edge_pull_bfs.asm:53-211 (step) and :217-311 (unit loop). */
static void orc_pull_bfs(orc_run_t* R, uint32_t tid)
{
    const orc_vec_t* L = R->g->pull;
    const uint64_t N = R->g->pull_count;
    /* r_frontiercount from graph_vertex_gather_index_end (phase_edge_pull_bfs.asm:270-279) */
    const uint64_t end = N ? orc_shared_vertex(&L[N - 1]) : 0;
    const uint64_t words = orc_frontier_limit(end, R->frontier_words);
    uint64_t vote = 0, pos = 0;
    for (uint64_t u = tid; u < R->units; u = orc_next_unit(R)) {
        uint64_t first, last;
        orc_unit_range(N, R->units, u, &first, &last);
        if (pos > last) continue;
        uint64_t i = pos > first ? pos : first;
        for (; i < last; ++i) {
            const orc_vec_t* v = &L[i];
            const uint64_t d = orc_shared_vertex(v);
            uint64_t next;
            const int c = orc_strong_check(R->wants, words, R->g->pull_index, d, &next);
            if (c < 0) {
                R->vbuf[tid] = vote;
                return;
            }
            if (c == 0) {
                i = next - 1;
                continue;
            }
            int any = 0;
            uint64_t parent = 0;
            for (int k = 0; k < 4; ++k) {
                if (!(v->lane[k] & VALID_BIT)) continue;
                const uint64_t s = v->lane[k] & NBR_MASK;
                if (orc_bit_test(R->has, s)) {
                    any = 1;
                    if (s > parent) parent = s;
                }
            }
            if (!any) continue;
            R->props[d] = (double)(int64_t)parent;
            R->depth[d] = R->level;
            orc_bit_set((uint64_t*)R->acc, d);
            orc_bit_clear(R->wants, d);
            vote += orc_stat(R, d);
        }
        pos = i;
    }
    R->vbuf[tid] = vote;
}

/* phase_edge_push_bfs.asm:53-202 (step) and :208-302 (unit loop). */
static void orc_push_bfs(orc_run_t* R, uint32_t tid)
{
    const orc_vec_t* L = R->g->push;
    const uint64_t N = R->g->push_count;
    const uint64_t units = (uint64_t)R->T << 5;
    const uint64_t end = N ? orc_shared_vertex(&L[N - 1]) : 0;
    const uint64_t words = orc_frontier_limit(end, R->frontier_words);
    uint64_t vote = 0, pos = 0;
    for (uint64_t u = tid; u < units; u = orc_next_unit(R)) {
        uint64_t first, last;
        orc_unit_range(N, units, u, &first, &last);
        if (pos > last) continue;
        uint64_t i = pos > first ? pos : first;
        for (; i < last; ++i) {
            const orc_vec_t* v = &L[i];
            const uint64_t s = orc_shared_vertex(v);
            uint64_t next;
            const int c = orc_strong_check(R->has, words, R->g->push_index, s, &next);
            if (c < 0) {
                R->vbuf[tid] = vote;
                return;
            }
            if (c == 0) {
                i = next - 1;
                continue;
            }
            int want[4] = {0, 0, 0, 0};
            for (int k = 0; k < 4; ++k)
                if (v->lane[k] & VALID_BIT) want[k] = orc_bit_test(R->wants, v->lane[k] & NBR_MASK);
            for (int k = 0; k < 4; ++k) {
                if (!want[k]) continue;
                const uint64_t d = v->lane[k] & NBR_MASK;
                orc_bit_clear(R->wants, d);
                orc_bit_set((uint64_t*)R->acc, d);
                R->props[d] = (double)(int64_t)s;
                R->depth[d] = R->level;
                vote += orc_stat(R, d);
            }
        }
        pos = i;
    }
    R->vbuf[tid] = vote;
}

/* phase_vertex_cc.asm / phase_vertex_bfs.asm:64-170: zero HasInfo* */
static void orc_vertex_clear(orc_run_t* R, uint32_t tid)
{
    uint64_t* bm = (uint64_t*)R->acc;
    const uint64_t W = R->frontier_words;
    const uint64_t first = W * tid / R->T, last = W * (tid + 1) / R->T;
    for (uint64_t w = first; w < last; ++w) bm[w] = 0;
}

/* ---- drivers -------------------------------------------------------------------- */

static double orc_now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

typedef struct orc_worker_t {
    orc_run_t* R;
    uint32_t tid;
    int algorithm;
} orc_worker_t;

/* scheduler_pull.inc:37-52: local thread 0 resets the counter to T behind a barrier */
static void orc_sched_init(orc_run_t* R, uint32_t tid)
{
    if (tid == 0) R->counter = R->T;
    orc_barrier(R);
}

/* execution_pr.c:92-168 */
static void orc_driver_pr(orc_run_t* R, uint32_t tid)
{
    const int push = (R->o->engine == ORC_ENGINE_PUSH);
    for (uint32_t it = 0; it < R->o->iterations; ++it) {
        orc_sched_init(R, tid);
        if (!push) {
            orc_pull_pr(R, tid);
            orc_barrier(R);
            if (tid == 0 && !R->o->traditional_pull) orc_merge(R->merge, R->units, R->acc);
        } else {
            orc_push_pr(R, tid);
        }
        orc_barrier(R);
        orc_vertex_pr(R, tid, push || R->o->traditional_pull);
        orc_barrier(R);
    }
    if (tid == 0) {
        R->res->iterations = R->o->iterations;
        R->res->pull_iterations = push ? 0 : R->o->iterations;
        R->res->push_iterations = push ? R->o->iterations : 0;
    }
}

/* execution_cc.c:94-268 and execution_bfs.c:120-276 */
static void orc_driver_hybrid(orc_run_t* R, uint32_t tid, int bfs)
{
    const uint64_t V = R->g->num_vertices, E = R->g->num_edges;
    const uint64_t threshold = (R->o->stat_mode == ORC_STAT_COUNT) ? V / 2 : E / 5;
    uint64_t ctr = 0;
    while (1) {
        ctr += 1;
        if (tid == 0) {
            int pull;
            if (R->o->engine == ORC_ENGINE_PULL)
                pull = 1;
            else if (R->o->engine == ORC_ENGINE_PUSH)
                pull = 0;
            else
                pull = R->vote > threshold;
            if (R->o->max_iterations && ctr > R->o->max_iterations) R->done = 1;
            R->use_pull = pull;
            R->level = ctr;
            if (!R->done) {
                if (ctr - 1 < R->log_cap && R->log) {
                    R->log[ctr - 1].engine = pull ? 1 : 2;
                    R->log[ctr - 1].vote = R->vote;
                }
                if (pull)
                    R->res->pull_iterations += 1;
                else
                    R->res->push_iterations += 1;
            }
            R->counter = R->T;
        }
        orc_barrier(R);
        if (R->done) {
            ctr -= 1;
            break;
        }
        if (R->use_pull) {
            if (bfs)
                orc_pull_bfs(R, tid);
            else
                orc_pull_cc(R, tid);
        } else {
            if (bfs)
                orc_push_bfs(R, tid);
            else
                orc_push_cc(R, tid);
        }
        orc_barrier(R);
        if (tid == 0) {
            /* phases.c:27-37 */
            uint64_t v = 0;
            for (uint32_t t = 0; t < R->T; ++t) v += R->vbuf[t];
            R->vote = v;
            if (v == 0) {
                R->done = 2;
            } else {
                uint64_t* t = R->has;
                R->has = (uint64_t*)R->acc;
                R->acc = (double*)t;
            }
        }
        orc_barrier(R);
        if (R->done == 2) break;
        orc_vertex_clear(R, tid);
        orc_barrier(R);
    }
    if (tid == 0) R->res->iterations = ctr;
}

static void* orc_worker_main(void* arg)
{
    orc_worker_t* w = (orc_worker_t*)arg;
    if (w->algorithm == 0)
        orc_driver_pr(w->R, w->tid);
    else
        orc_driver_hybrid(w->R, w->tid, w->algorithm == 2);
    return NULL;
}

static void orc_spawn(orc_run_t* R, int algorithm)
{
    const uint32_t T = R->T;
    if (T > 1) pthread_barrier_init(&R->barrier, NULL, T);
    orc_worker_t* w = (orc_worker_t*)calloc(T, sizeof(*w));
    pthread_t* th = (pthread_t*)calloc(T, sizeof(*th));
    const double t0 = orc_now_ms();
    for (uint32_t t = 0; t < T; ++t) {
        w[t].R = R;
        w[t].tid = t;
        w[t].algorithm = algorithm;
        if (t > 0) pthread_create(&th[t], NULL, orc_worker_main, &w[t]);
    }
    orc_worker_main(&w[0]);
    for (uint32_t t = 1; t < T; ++t) pthread_join(th[t], NULL);
    R->res->elapsed_ms = orc_now_ms() - t0;
    if (T > 1) pthread_barrier_destroy(&R->barrier);
    free(w);
    free(th);
}

static void orc_run_init(orc_run_t* R, const orc_graph_t* g, const orc_options_t* o, orc_result_t* r)
{
    memset(R, 0, sizeof(*R));
    memset(r, 0, sizeof(*r));
    R->g = g;
    R->o = o;
    R->T = o->threads ? o->threads : 1;
    R->units = (uint64_t)(o->units_per_thread ? o->units_per_thread : 32) * R->T;
    R->res = r;
    const uint64_t V = g->num_vertices;
    R->frontier_words = (V >> 6) + ((V & 63) ? 1 : 0);
    R->props = (double*)calloc(V + 8, sizeof(double));
    R->acc = (double*)calloc(V + 8, sizeof(double));
    R->has = (uint64_t*)calloc(R->frontier_words + 8, sizeof(uint64_t));
    R->wants = (uint64_t*)calloc(R->frontier_words + 8, sizeof(uint64_t));
    R->sbuf = (double*)calloc(R->T + 8, sizeof(double));
    R->vbuf = (uint64_t*)calloc(R->T + 8, sizeof(uint64_t));
    R->merge = (orc_merge_t*)malloc(R->units * sizeof(orc_merge_t));
    for (uint64_t u = 0; u < R->units; ++u) {
        R->merge[u].initial_vertex_id = ~0ull;
        R->merge[u].final_vertex_id = ~0ull;
        R->merge[u].final_partial_value = 0.0;
        R->merge[u].pad = 0;
    }
}

static void orc_run_free(orc_run_t* R)
{
    free(R->props);
    free(R->acc);
    free(R->has);
    free(R->wants);
    free(R->sbuf);
    free(R->vbuf);
    free(R->merge);
    free(R->depth);
}

int orc_pagerank(const orc_graph_t* g, const orc_options_t* o, double* props, double* ranks, orc_result_t* r)
{
    orc_run_t R;
    orc_run_init(&R, g, o, r);
    const uint64_t V = g->num_vertices;
    const double nv = (double)V;
    /* execution_pr.c:78-88 */
    for (uint64_t v = 0; v < V; ++v) R.props[v] = (1.0 / nv) / (0.0 == g->outdeg[v] ? nv : g->outdeg[v]);
    orc_spawn(&R, 0);
    for (uint64_t v = 0; v < V; ++v) {
        if (props) props[v] = R.props[v];
        if (ranks) ranks[v] = R.props[v] * (0.0 == g->outdeg[v] ? nv : g->outdeg[v]);
    }
    orc_run_free(&R);
    return 0;
}

int orc_cc_from(const orc_graph_t* g, const orc_options_t* o, const uint64_t* init, uint64_t* labels,
                orc_iteration_t* log, uint64_t log_cap, orc_result_t* r)
{
    orc_run_t R;
    orc_run_init(&R, g, o, r);
    const uint64_t V = g->num_vertices, E = g->num_edges;
    for (uint64_t v = 0; v < V; ++v) R.props[v] = (double)(init ? init[v] : v); /* execution_cc.c:87-90 */
    for (uint64_t w = 0; w < R.frontier_words; ++w) R.has[w] = R.wants[w] = ~0ull;
    R.vote = 0;
    if (o->stat_mode != ORC_STAT_COUNT) R.vote += E;                  /* execution_cc.c:106-118 */
    if (o->stat_mode != ORC_STAT_OUTDEG) R.vote += V;
    R.log = log;
    R.log_cap = log_cap;
    orc_spawn(&R, 1);
    for (uint64_t v = 0; v < V; ++v) labels[v] = (uint64_t)R.props[v];
    orc_run_free(&R);
    return 0;
}

int orc_cc(const orc_graph_t* g, const orc_options_t* o, uint64_t* labels, orc_iteration_t* log, uint64_t log_cap,
           orc_result_t* r)
{
    return orc_cc_from(g, o, NULL, labels, log, log_cap, r);
}

int orc_bfs(const orc_graph_t* g, const orc_options_t* o, int64_t* parents, uint64_t* depths, orc_iteration_t* log,
            uint64_t log_cap, orc_result_t* r)
{
    orc_run_t R;
    orc_run_init(&R, g, o, r);
    const uint64_t V = g->num_vertices, root = o->root;
    if (root >= V) {
        orc_run_free(&R);
        return -1;
    }
    R.depth = (uint64_t*)malloc(V * sizeof(uint64_t));
    for (uint64_t v = 0; v < V; ++v) {
        R.props[v] = -1.0;                                            /* execution_bfs.c:113-116 */
        R.depth[v] = UINT64_MAX;
    }
    R.depth[root] = 0;
    for (uint64_t w = 0; w < R.frontier_words; ++w) {                 /* execution_bfs.c:70-102 */
        R.has[w] = 0;
        R.wants[w] = ~0ull;
    }
    R.has[root >> 6] = 1ull << (root & 63);
    R.wants[root >> 6] = ~(1ull << (root & 63));
    R.vote = 0;                                                       /* execution_bfs.c:132-144 */
    if (o->stat_mode != ORC_STAT_COUNT) R.vote += (uint64_t)g->outdeg[root];
    if (o->stat_mode != ORC_STAT_OUTDEG) R.vote += 1;
    R.log = log;
    R.log_cap = log_cap;
    orc_spawn(&R, 2);
    for (uint64_t v = 0; v < V; ++v) {
        if (parents) parents[v] = (int64_t)R.props[v];
        if (depths) depths[v] = R.depth[v];
    }
    orc_run_free(&R);
    return 0;
}

/* ---- BFS certificate --------------------------------------------------------- */

typedef struct orc_cert_t {
    const orc_graph_t* g;
    const int64_t* parents;
    const uint64_t* depths;
    uint8_t* bad;
    uint8_t* found;
    uint64_t first, last;
} orc_cert_t;

static void* orc_cert_edges(void* arg)
{
    orc_cert_t* c = (orc_cert_t*)arg;
    for (uint64_t i = c->first; i < c->last; ++i) {
        const orc_vec_t* v = &c->g->pull[i];
        const uint64_t d = orc_shared_vertex(v);
        for (int k = 0; k < 4; ++k) {
            if (!(v->lane[k] & VALID_BIT)) continue;
            const uint64_t u = v->lane[k] & NBR_MASK;
            /* a destination's vectors can straddle two threads' ranges */
            if (c->parents[d] >= 0 && (uint64_t)c->parents[d] == u) __atomic_store_n(&c->found[d], 1, __ATOMIC_RELAXED);
            if (c->depths[u] != UINT64_MAX && (c->depths[d] == UINT64_MAX || c->depths[d] > c->depths[u] + 1))
                __atomic_store_n(&c->bad[d], 1, __ATOMIC_RELAXED);
        }
    }
    return NULL;
}

uint64_t orc_bfs_certificate_violations(const orc_graph_t* g, uint64_t root, const int64_t* parents,
                                        const uint64_t* depths, uint32_t threads)
{
    const uint64_t V = g->num_vertices;
    const uint32_t T = threads ? threads : 1;
    uint8_t* bad = (uint8_t*)calloc(V ? V : 1, 1);
    uint8_t* found = (uint8_t*)calloc(V ? V : 1, 1);
    orc_cert_t* c = (orc_cert_t*)calloc(T, sizeof(*c));
    pthread_t* th = (pthread_t*)calloc(T, sizeof(*th));
    for (uint32_t t = 0; t < T; ++t) {
        c[t] = (orc_cert_t){g, parents, depths, bad, found, g->pull_count * t / T, g->pull_count * (t + 1) / T};
        if (t > 0) pthread_create(&th[t], NULL, orc_cert_edges, &c[t]);
    }
    orc_cert_edges(&c[0]);
    for (uint32_t t = 1; t < T; ++t) pthread_join(th[t], NULL);
    uint64_t n = 0;
    for (uint64_t v = 0; v < V; ++v) {
        int ok = !bad[v];
        if (v == root)
            ok = ok && depths[v] == 0 && parents[v] == -1;
        else if (depths[v] == UINT64_MAX)
            ok = ok && parents[v] == -1;
        else
            ok = ok && found[v] && parents[v] >= 0 && (uint64_t)parents[v] < V &&
                 depths[parents[v]] != UINT64_MAX && depths[parents[v]] + 1 == depths[v];
        n += ok ? 0 : 1;
    }
    free(c);
    free(th);
    free(bad);
    free(found);
    return n;
}

/* ---- selector MLP (restated independently of the HIP kernels) ----------------- */

#define MI 8
#define MH 32
#define MO 2
#define P_W1 0
#define P_B1 (P_W1 + MH * MI)
#define P_W2 (P_B1 + MH)
#define P_B2 (P_W2 + MH * MH)
#define P_W3 (P_B2 + MH)
#define P_B3 (P_W3 + MO * MH)
#define P_N (P_B3 + MO)

static void orc_mlp_fwd(const float* p, const float* x, double* h1, double* h2, double* y)
{
    for (int j = 0; j < MH; ++j) {
        double a = p[P_B1 + j];
        for (int i = 0; i < MI; ++i) a += (double)p[P_W1 + j * MI + i] * x[i];
        h1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < MH; ++j) {
        double a = p[P_B2 + j];
        for (int i = 0; i < MH; ++i) a += (double)p[P_W2 + j * MH + i] * h1[i];
        h2[j] = a > 0 ? a : 0;
    }
    for (int k = 0; k < MO; ++k) {
        double a = p[P_B3 + k];
        for (int i = 0; i < MH; ++i) a += (double)p[P_W3 + k * MH + i] * h2[i];
        y[k] = a;
    }
}

void orc_mlp_forward(const float* params, const float* x, float* y)
{
    double h1[MH], h2[MH], yy[MO];
    orc_mlp_fwd(params, x, h1, h2, yy);
    for (int k = 0; k < MO; ++k) y[k] = (float)yy[k];
}

float orc_mlp_gradient(const float* p, const float* X, const float* T, uint64_t n, float* grad)
{
    double* g = (double*)calloc(P_N, sizeof(double));
    double loss = 0.0;
    for (uint64_t s = 0; s < n; ++s) {
        double h1[MH], h2[MH], y[MO], dy[MO], dh2[MH], dh1[MH];
        orc_mlp_fwd(p, X + s * MI, h1, h2, y);
        for (int k = 0; k < MO; ++k) {
            const double r = y[k] - T[s * MO + k];
            loss += r * r;
            dy[k] = 2.0 * r / (double)(n * MO);
        }
        for (int i = 0; i < MH; ++i) dh2[i] = 0;
        for (int k = 0; k < MO; ++k) {
            g[P_B3 + k] += dy[k];
            for (int i = 0; i < MH; ++i) {
                g[P_W3 + k * MH + i] += dy[k] * h2[i];
                dh2[i] += dy[k] * p[P_W3 + k * MH + i];
            }
        }
        for (int i = 0; i < MH; ++i) dh1[i] = 0;
        for (int j = 0; j < MH; ++j) {
            const double da = h2[j] > 0 ? dh2[j] : 0;
            g[P_B2 + j] += da;
            for (int i = 0; i < MH; ++i) {
                g[P_W2 + j * MH + i] += da * h1[i];
                dh1[i] += da * p[P_W2 + j * MH + i];
            }
        }
        for (int j = 0; j < MH; ++j) {
            const double da = h1[j] > 0 ? dh1[j] : 0;
            g[P_B1 + j] += da;
            for (int i = 0; i < MI; ++i) g[P_W1 + j * MI + i] += da * X[s * MI + i];
        }
    }
    for (int i = 0; i < P_N; ++i) grad[i] = (float)g[i];
    free(g);
    return (float)(loss / (double)(n * MO));
}
