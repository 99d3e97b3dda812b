/*
 * execution.c -- the algorithm drivers, kept as host C code.  One host thread
 * runs the iteration loop and calls the per-phase C ABI once per phase.
 *
 *   PageRank   execution_pr.c:92-168
 *   CC         execution_cc.c:94-268
 *   BFS        execution_bfs.c:120-276
 *
 * The reference's thread barriers become stream order.  The only host/device
 * synchronisation per CC/BFS iteration is phase_op_combine_global_var_from_buf,
 * which returns converge_vote for the termination check and the engine choice,
 * as in the reference.  PageRank never synchronises inside its loop: S stays
 * in the device reduce buffer between the edge and vertex phases.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "grazelle_host.h"
#include "grazelle_phases.h"
#include "grazelle_selector.h"
#include "gz_host_internal.h"

static uint64_t* gz_reduce_buffer = NULL;
static gz_iteration_record_t* gz_records = NULL;
static uint64_t gz_records_count = 0, gz_records_cap = 0;

void gz_run_options_default(gz_run_options_t* o)
{
    memset(o, 0, sizeof(*o));
    o->algorithm = GZ_ALG_PR;
    o->engine = GZ_ENGINE_HYBRID;
    o->iterations = 1;
    o->stat_mode = GZ_STAT_COUNT_AND_OUTDEG;
    o->pull_variant = GZ_PULL_SCHED_AWARE;
    o->damping = 0.875;
    o->root = 0;
}

static double gz_now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

static void gz_record(const gz_iteration_record_t* r)
{
    if (gz_records_count == gz_records_cap) {
        gz_records_cap = gz_records_cap ? 2 * gz_records_cap : 256;
        gz_records = (gz_iteration_record_t*)realloc(gz_records, gz_records_cap * sizeof(*gz_records));
    }
    gz_records[gz_records_count++] = *r;
}

uint64_t gz_execution_records(gz_iteration_record_t* out, uint64_t capacity)
{
    const uint64_t n = gz_records_count < capacity ? gz_records_count : capacity;
    if (out && n) memcpy(out, gz_records, n * sizeof(*out));
    return gz_records_count;
}

void gz_execution_prepare(const gz_run_options_t* o)
{
    const uint64_t V = gz_graph.num_vertices;
    const uint64_t fwords = (V + 63) / 64;
    gz_phase_config.damping = o->damping;
    gz_phase_config.stat_mode = o->stat_mode;
    gz_phase_config.pull_variant = o->pull_variant;
    gz_phase_config.collect_stats = o->collect_stats;
    gz_graph_reset_roles();
    if (gz_reduce_buffer == NULL) gz_reduce_buffer = gz_reduce_buffer_create();
    gz_device_zero(gz_reduce_buffer, 8 * sizeof(uint64_t));
    gz_reduce_mirror_stale();
    if (gz_graph.pull_vectors != NULL) gz_merge_buffer_allocate(gz_pull_units_for(gz_graph.pull_vector_count));

    switch (o->algorithm) {
    case GZ_ALG_PR:
        gz_dev_init_pr((double*)gz_graph.vertex_props, (double*)gz_graph.vertex_accumulators,
                       gz_graph.vertex_outdegrees, V);
        gz_pr_clear_fixed();
        gz_dev_fill_u64(gz_graph.frontier_has_info, ~0ull, fwords);
        gz_dev_fill_u64(gz_graph.frontier_wants_info, ~0ull, fwords);
        break;
    case GZ_ALG_CC:
        /* labels start at the original ids, also on a renumbered graph */
        if (gz_graph_original_ids_device() != NULL)
            gz_device_copy(gz_graph.vertex_props, gz_graph_original_ids_device(), V * sizeof(uint64_t));
        else
            gz_dev_iota_u64((uint64_t*)gz_graph.vertex_props, V);
        gz_dev_fill_u64((uint64_t*)gz_graph.vertex_accumulators, 0, fwords);
        gz_dev_fill_u64(gz_graph.frontier_has_info, ~0ull, fwords);
        gz_dev_fill_u64(gz_graph.frontier_wants_info, ~0ull, fwords);
        break;
    case GZ_ALG_BFS: {
        if (o->root >= V) {
            fprintf(stderr, "grazelle: BFS root %llu is not a vertex of a %llu-vertex graph\n",
                    (unsigned long long)o->root, (unsigned long long)V);
            exit(66);
        }
        const uint64_t r = gz_graph_device_vertex(o->root);
        gz_dev_bfs_reset((uint64_t*)gz_graph.vertex_props, V, (uint64_t*)gz_graph.vertex_accumulators,
                         gz_graph.frontier_has_info, gz_graph.frontier_wants_info, fwords, r);
        break;
    }
    default:
        fprintf(stderr, "grazelle: unknown algorithm %u\n", o->algorithm);
        exit(67);
    }
    gz_records_count = 0;
}

/* ---- CC / BFS helpers ------------------------------------------------------ */

static void gz_edge_phase(int bfs, int pull)
{
    if (pull) {
        if (bfs)
            perform_edge_pull_phase_bfs(gz_graph.pull_vectors, gz_graph.pull_vector_count);
        else
            perform_edge_pull_phase_cc(gz_graph.pull_vectors, gz_graph.pull_vector_count);
    } else {
        if (bfs)
            perform_edge_push_phase_bfs(gz_graph.push_vectors, gz_graph.push_vector_count);
        else
            perform_edge_push_phase_cc(gz_graph.push_vectors, gz_graph.push_vector_count);
    }
}

/* What the selector left next to the vote (gz_sel_dev_step). */
typedef struct gz_selector_reading_t {
    uint32_t choice;            /* 1 pull, 2 push, 0 no model */
    float features[8];
    float predicted[2];
} gz_selector_reading_t;

static void gz_selector_read(gz_selector_reading_t* s)
{
    const uint64_t* w = gz_reduce_words();
    s->choice = (uint32_t)w[1];
    memcpy(s->features, &w[2], sizeof(s->features));
    memcpy(s->predicted, &w[6], sizeof(s->predicted));
}

/* The learned BFS engine's decision on the host (gz_sel_dev_sums left the
 * frontier's sums in the reduce words): the features in the decision
 * kernel's formulas, with the unvisited counts kept here, and the forward
 * pass of gz_selector_host_decide. */
static void gz_selector_host_step(gz_selector_reading_t* s, uint64_t* unvisited, uint64_t* unvisited_in,
                                  uint64_t V, uint64_t E, float iteration, float prev_engine)
{
    const uint64_t* w = gz_reduce_words();
    const uint64_t cv = w[2], ce = w[3], ci = w[4];
    const uint32_t msb = (uint32_t)w[5];
    float ms;
    memcpy(&ms, &msb, sizeof(ms));
    *unvisited -= cv;
    *unvisited_in -= ci;
    s->features[0] = logf(1.f + (float)cv);
    s->features[1] = logf(1.f + (float)ce);
    s->features[2] = logf(1.f + (float)*unvisited);
    s->features[3] = logf(1.f + (float)*unvisited_in);
    s->features[4] = iteration;
    s->features[5] = prev_engine;
    s->features[6] = logf(1.f + ms);
    s->features[7] = V ? (float)E / (float)V : 0.f;
    s->choice = gz_selector_host_decide(s->features, s->predicted);
}

/* Everything a CC / BFS edge phase writes: vertex words and the three
 * bitmaps (HasInfo may be updated by CC push, WantsInfo by BFS, HasInfo* by
 * both). */
typedef struct gz_snapshot_t {
    void* props;
    void* has;
    void* wants;
    void* next;
    size_t vbytes, fbytes;
} gz_snapshot_t;

static void gz_snapshot_alloc(gz_snapshot_t* s)
{
    const uint64_t V = gz_graph.num_vertices;
    s->vbytes = V * sizeof(uint64_t);
    s->fbytes = ((V + 63) / 64) * sizeof(uint64_t);
    s->props = gz_device_malloc(s->vbytes ? s->vbytes : 8);
    s->has = gz_device_malloc(s->fbytes ? s->fbytes : 8);
    s->wants = gz_device_malloc(s->fbytes ? s->fbytes : 8);
    s->next = gz_device_malloc(s->fbytes ? s->fbytes : 8);
}

static void gz_snapshot_free(gz_snapshot_t* s)
{
    gz_device_free(s->props);
    gz_device_free(s->has);
    gz_device_free(s->wants);
    gz_device_free(s->next);
}

static void gz_snapshot_take(gz_snapshot_t* s)
{
    gz_device_copy(s->props, gz_graph.vertex_props, s->vbytes);
    gz_device_copy(s->has, gz_graph.frontier_has_info, s->fbytes);
    gz_device_copy(s->wants, gz_graph.frontier_wants_info, s->fbytes);
    gz_device_copy(s->next, gz_graph.vertex_accumulators, s->fbytes);
}

static void gz_snapshot_restore(const gz_snapshot_t* s)
{
    gz_device_copy(gz_graph.vertex_props, s->props, s->vbytes);
    gz_device_copy(gz_graph.frontier_has_info, s->has, s->fbytes);
    gz_device_copy(gz_graph.frontier_wants_info, s->wants, s->fbytes);
    gz_device_copy(gz_graph.vertex_accumulators, s->next, s->fbytes);
}

/* Best of three timed runs of one engine's edge phase, each from the
 * snapshot; the state afterwards is that engine's result.  With `stamp` the
 * last run is also bracketed by the selector's device clock, so the next
 * iteration's "previous edge-phase time" feature is measured exactly as the
 * learned engine measures it. */
/* A push over the frontier that a vote of `vote` built may skip the
 * compaction when the vote is small: tell the push phase. */
static void gz_push_hint(int bfs, int pull, uint64_t vote)
{
    if (pull) return;
    if (bfs)
        gz_bfs_push_vote_hint(vote);
    else
        gz_cc_push_vote_hint(vote);
}

static double gz_probe_engine(int bfs, int pull, const gz_snapshot_t* s, gz_timer_t* t, int stamp, uint64_t vote)
{
    double best = 0.0;
    for (int rep = 0; rep < 3; ++rep) {
        gz_snapshot_restore(s);
        if (stamp && rep == 2) gz_sel_dev_stamp(bfs && gz_phase_config.push_frontier_skip);
        gz_timer_start(t);
        phase_op_reset_global_accum();
        gz_push_hint(bfs, pull, vote);
        gz_edge_phase(bfs, pull);
        gz_timer_stop(t);
        const double ms = gz_timer_elapsed_ms(t);
        if (rep == 0 || ms < best) best = ms;
    }
    return best;
}

static int gz_probe_wide_on = 0;

/* The learned engine decides only while threshold / gz_sel_below <= vote <=
 * gz_sel_above * threshold; outside, the fixed rule's choice stands (pull
 * above, push below).  A factor of 0 lifts that side of the band. */
static double gz_sel_above = 1.0, gz_sel_below = 32.0;

void gz_selector_set_trust(double above, double below)
{
    gz_sel_above = (above >= 1.0) ? above : 0.0;
    gz_sel_below = (below >= 1.0) ? below : 0.0;
}

void gz_probe_wide(int on) { gz_probe_wide_on = on ? 1 : 0; }

/* As gz_probe_engine for an edge-phase variant: the frontier skip (the
 * compaction) on or off and the push's vote hint (~0: no direct bitmap
 * read) given; the configuration is restored afterwards. */
static double gz_probe_variant(int bfs, int pull, const gz_snapshot_t* s, gz_timer_t* t, int skip, uint64_t hint)
{
    const uint32_t saved = gz_phase_config.push_frontier_skip;
    gz_phase_config.push_frontier_skip = (uint32_t)skip;
    double best = 0.0;
    for (int rep = 0; rep < 3; ++rep) {
        gz_snapshot_restore(s);
        gz_timer_start(t);
        phase_op_reset_global_accum();
        if (!pull) {
            if (bfs)
                gz_bfs_push_vote_hint(hint);
            else
                gz_cc_push_vote_hint(hint);
        }
        gz_edge_phase(bfs, pull);
        gz_timer_stop(t);
        const double ms = gz_timer_elapsed_ms(t);
        if (rep == 0 || ms < best) best = ms;
    }
    gz_phase_config.push_frontier_skip = saved;
    return best;
}

static void gz_swap_frontiers(void)
{
    /* execution_cc.c:245-250 / execution_bfs.c:253-258 */
    void* t = gz_graph.frontier_has_info;
    gz_graph.frontier_has_info = (uint64_t*)gz_graph.vertex_accumulators;
    gz_graph.vertex_accumulators = t;
}

static int gz_pr_replay_on = 1;

void gz_pr_graph_replay(int on) { gz_pr_replay_on = on ? 1 : 0; }

typedef struct gz_pr_iteration_t {
    int use_push;
    uint32_t pull_variant;
    int fused_tail; /* merge, sum and vertex phase in one launch (gz_pr_vertex_merged) */
} gz_pr_iteration_t;

/* One PageRank iteration's launches, as the loop below makes them. */
static void gz_pr_enqueue_iteration(void* p)
{
    const gz_pr_iteration_t* c = (const gz_pr_iteration_t*)p;
    phase_op_reset_global_accum();
    if (c->fused_tail) {
        perform_edge_pull_phase_pr(gz_graph.pull_vectors, gz_graph.pull_vector_count);
        gz_pr_vertex_merged(gz_reduce_buffer);
        return;
    }
    if (!c->use_push) {
        perform_edge_pull_phase_pr(gz_graph.pull_vectors, gz_graph.pull_vector_count);
        if (c->pull_variant == GZ_PULL_SCHED_AWARE || c->pull_variant == GZ_PULL_MERGE_BUFFER)
            edge_pull_op_merge_with_merge_buffer(gz_merge_buffer, gz_pull_units_for(gz_graph.pull_vector_count),
                                                 (double*)gz_graph.vertex_accumulators);
    } else {
        perform_edge_push_phase_pr(gz_graph.push_vectors, gz_graph.push_vector_count);
    }
    phase_op_write_global_accum_to_buf(gz_reduce_buffer);
    perform_vertex_phase_pr(0, gz_graph.num_vertices, gz_reduce_buffer);
}

/* Everything the launches of one iteration depend on besides the device
 * allocations: the engine, the phase configuration, the graph record and
 * the sliced pull's settings. */
static size_t gz_pr_iteration_key(const gz_pr_iteration_t* c, uint64_t* k)
{
    size_t n = 0;
    uint64_t damping;
    memcpy(&damping, &gz_phase_config.damping, sizeof(damping));
    uint32_t sl[3];
    gz_slices_settings(sl);
    k[n++] = (uint64_t)c->use_push;
    k[n++] = c->pull_variant;
    k[n++] = (uint64_t)c->fused_tail;
    k[n++] = damping;
    k[n++] = gz_phase_config.stat_mode;
    k[n++] = gz_phase_config.pull_variant;
    k[n++] = gz_phase_config.chunk_vectors;
    k[n++] = gz_phase_config.collect_stats;
    k[n++] = gz_phase_config.push_frontier_skip;
    k[n++] = gz_phase_config.vectors_per_lane;
    k[n++] = gz_graph.num_vertices;
    k[n++] = (uint64_t)(uintptr_t)gz_graph.pull_vectors;
    k[n++] = gz_graph.pull_vector_count;
    k[n++] = (uint64_t)(uintptr_t)gz_graph.push_vectors;
    k[n++] = gz_graph.push_vector_count;
    k[n++] = (uint64_t)(uintptr_t)gz_graph.vertex_outdegrees;
    k[n++] = (uint64_t)(uintptr_t)gz_graph.vertex_props;
    k[n++] = (uint64_t)(uintptr_t)gz_graph.vertex_accumulators;
    k[n++] = (uint64_t)(uintptr_t)gz_reduce_buffer;
    k[n++] = (uint64_t)(uintptr_t)gz_merge_buffer;
    k[n++] = gz_merge_buffer_count;
    k[n++] = sl[0];
    k[n++] = sl[1];
    k[n++] = sl[2];
    return n;
}

static void gz_run_pr(const gz_run_options_t* o, gz_run_result_t* r)
{
    const uint64_t V = gz_graph.num_vertices;
    int use_push = (o->engine == GZ_ENGINE_PUSH);
    if (use_push && gz_graph.push_vectors == NULL) use_push = 0;
    if (!use_push && gz_graph.pull_vectors == NULL) use_push = 1;
    gz_timer_t* t = o->profile ? gz_timer_create() : NULL;
    /* with nothing read back per iteration, iterations 2.. run as one graph
     * of their launches (the first runs eagerly and makes every allocation
     * the phases need); a configuration is captured when it runs again */
    const int replay = gz_pr_replay_on && t == NULL && !o->collect_stats && !o->edge_only && !o->vertex_only &&
                       !gz_kernel_timing_enabled() && o->iterations > 1 && o->iterations - 1 <= 0xffffffffull;
    const int plain = t == NULL && !use_push && !o->collect_stats && !o->edge_only && !o->vertex_only &&
                      !gz_kernel_timing_enabled() &&
                      (o->pull_variant == GZ_PULL_SCHED_AWARE || o->pull_variant == GZ_PULL_MERGE_BUFFER);
    /* with few enough chunks, the merge, the sum of S and the vertex phase
     * run as one launch (gz_pr_vertex_merged) */
    const int fused_tail = plain && gz_pr_fused_tail_applies();
    /* every iteration in one launch (gz_pr_run_persistent), when switched on */
    if (plain) {
        const int one = gz_pr_run_persistent(o->iterations, gz_reduce_buffer);
        if (one > 0) {
            r->pull_iterations = o->iterations;
            r->iterations = o->iterations;
            return;
        }
        if (one < 0) /* a barrier gave up mid-run: start again from the initial ranks */
            gz_dev_init_pr((double*)gz_graph.vertex_props, (double*)gz_graph.vertex_accumulators,
                           gz_graph.vertex_outdegrees, V);
    }
    uint64_t ctr;
    for (ctr = 0; ctr < o->iterations; ++ctr) {
        if (replay && ctr == 1) {
            gz_pr_iteration_t c = {use_push, o->pull_variant, fused_tail};
            uint64_t key[32];
            const size_t n = gz_pr_iteration_key(&c, key);
            const uint64_t rest = o->iterations - 1;
            if (gz_graph_replay_cached(key, n, gz_pr_enqueue_iteration, &c, (uint32_t)rest)) {
                if (use_push)
                    r->push_iterations += rest;
                else
                    r->pull_iterations += rest;
                ctr = o->iterations;
                break;
            }
        }
        if (fused_tail) {
            phase_op_reset_global_accum();
            perform_edge_pull_phase_pr(gz_graph.pull_vectors, gz_graph.pull_vector_count);
            if (ctr == 0) gz_pr_run_heads_build(); /* the first pull's records fix the runs */
            gz_pr_vertex_merged(gz_reduce_buffer);
            r->pull_iterations += 1;
            continue;
        }
        gz_iteration_record_t rec;
        memset(&rec, 0, sizeof(rec));
        rec.iteration = ctr + 1;
        if (!o->vertex_only) {
            if (t) gz_timer_start(t);
            phase_op_reset_global_accum();
            if (!use_push) {
                r->pull_iterations += 1;
                rec.engine = 1;
                perform_edge_pull_phase_pr(gz_graph.pull_vectors, gz_graph.pull_vector_count);
                if (o->pull_variant == GZ_PULL_SCHED_AWARE || o->pull_variant == GZ_PULL_MERGE_BUFFER)
                    edge_pull_op_merge_with_merge_buffer(gz_merge_buffer, gz_pull_units_for(gz_graph.pull_vector_count),
                                                         (double*)gz_graph.vertex_accumulators);
            } else {
                r->push_iterations += 1;
                rec.engine = 2;
                perform_edge_push_phase_pr(gz_graph.push_vectors, gz_graph.push_vector_count);
            }
            phase_op_write_global_accum_to_buf(gz_reduce_buffer);
            if (t) {
                gz_timer_stop(t);
                rec.edge_ms = gz_timer_elapsed_ms(t);
            }
            if (o->collect_stats) gz_phase_stats_read(&rec.vectors, &rec.edges);
        }
        if (!o->edge_only) perform_vertex_phase_pr(0, V, gz_reduce_buffer);
        if (o->profile || o->collect_stats) gz_record(&rec);
    }
    r->iterations = ctr;
    if (t) gz_timer_destroy(t);
}

/* ---- CC with the engine chosen on the device --------------------------------
 * A CC run that records nothing per iteration and uses the fixed threshold
 * (or a forced engine) over the plain pull and the compacted push keeps its
 * vote and engine choice in device control words (gz_cc_loop_enqueue), so
 * its iterations can run as a captured HIP graph: GZ_CC_LOOP_ROUND
 * iterations per graph, relaunched until the engine word says stop.  As with
 * PageRank, a configuration is captured the second time in a row it runs
 * (gz_graph_replay_cached); a first request runs the host loop below.  Off
 * by default: every captured iteration also dispatches the kernels of the
 * engines it does not run, which cost more than the one host round trip per
 * iteration they save on the configurations measured (CC on R-MAT s22 with
 * trees, the 1024-graph batch). */
#define GZ_CC_LOOP_ROUND 4
#define GZ_CC_LOOP_LOG 4096
static int gz_cc_replay_on = 0;
static uint64_t* gz_cc_ctrl = NULL;

void gz_cc_graph_replay(int on) { gz_cc_replay_on = on ? 1 : 0; }

static void gz_cc_loop_round(void* p)
{
    gz_cc_loop_enqueue((uint64_t*)p);
    gz_swap_frontiers();
}

static int gz_cc_device_loop(const gz_run_options_t* o, uint64_t vote, uint64_t threshold, gz_run_result_t* r)
{
    const uint64_t V = gz_graph.num_vertices;
    if (!gz_cc_replay_on || o->algorithm != GZ_ALG_CC || o->profile || o->collect_stats || gz_kernel_timing_enabled() ||
        (o->engine != GZ_ENGINE_HYBRID && o->engine != GZ_ENGINE_PULL && o->engine != GZ_ENGINE_PUSH) ||
        o->pull_variant != GZ_PULL_SCHED_AWARE || !gz_phase_config.push_frontier_skip ||
        gz_graph.pull_vectors == NULL || gz_graph.push_vectors == NULL || gz_graph.push_vertex_offsets == NULL ||
        V == 0)
        return 0;
    uint64_t pull_waves, push_waves;
    const uint64_t waves = gz_cc_loop_waves(&pull_waves, &push_waves);
    /* every buffer the launches use exists before any capture */
    gz_scratch_reserve_public(waves, V);
    if (gz_cc_ctrl == NULL) gz_cc_ctrl = (uint64_t*)gz_device_malloc(GZ_LOOP_WORDS(GZ_CC_LOOP_LOG) * sizeof(uint64_t));
    uint64_t ctrl[16];
    memset(ctrl, 0, sizeof(ctrl));
    const int first_pull = o->engine == GZ_ENGINE_PULL ? 1 : o->engine == GZ_ENGINE_PUSH ? 0 : vote > threshold;
    /* out-degrees in the vote bound a frontier's vectors (gz_push_sparse_eligible) */
    const uint64_t sparse_ok = o->stat_mode != GZ_STAT_COUNT_ONLY;
    ctrl[GZ_LOOP_ENGINE] = gz_loop_engine(vote, first_pull, sparse_ok);
    ctrl[GZ_LOOP_VOTE] = vote;
    ctrl[GZ_LOOP_THRESHOLD] = threshold;
    ctrl[GZ_LOOP_MODE] = o->engine;
    ctrl[GZ_LOOP_MAX_ITERATIONS] = o->max_iterations;
    ctrl[GZ_LOOP_LOG_CAP] = GZ_CC_LOOP_LOG;
    ctrl[GZ_LOOP_SPARSE_OK] = sparse_ok;
    gz_copy_to_device(gz_cc_ctrl, ctrl, sizeof(ctrl));
    uint64_t key[24];
    size_t n = 0;
    key[n++] = 0x43434c4f4f50ull; /* distinct from the PageRank keys */
    key[n++] = V;
    key[n++] = (uint64_t)(uintptr_t)gz_graph.pull_vectors;
    key[n++] = gz_graph.pull_vector_count;
    key[n++] = (uint64_t)(uintptr_t)gz_graph.push_vectors;
    key[n++] = gz_graph.push_vector_count;
    key[n++] = (uint64_t)(uintptr_t)gz_graph.push_vertex_offsets;
    key[n++] = (uint64_t)(uintptr_t)gz_graph.vertex_props;
    key[n++] = (uint64_t)(uintptr_t)gz_graph.vertex_accumulators;
    key[n++] = (uint64_t)(uintptr_t)gz_graph.frontier_has_info;
    key[n++] = (uint64_t)(uintptr_t)gz_graph.vertex_outdegrees;
    key[n++] = (uint64_t)(uintptr_t)gz_cc_ctrl;
    key[n++] = gz_phase_config.chunk_vectors;
    key[n++] = gz_phase_config.vectors_per_lane;
    key[n++] = gz_phase_config.stat_mode;
    uint64_t launches = 0;
    while (1) {
        if (!gz_graph_replay_cached_slot(1, key, n, gz_cc_loop_round, gz_cc_ctrl, GZ_CC_LOOP_ROUND)) {
            if (launches == 0) return 0; /* a first request: the host loop runs */
            fprintf(stderr, "grazelle: the captured CC loop was dropped mid-run\n");
            exit(72);
        }
        launches += 1;
        gz_copy_to_host(ctrl, gz_cc_ctrl, sizeof(ctrl));
        if (ctrl[GZ_LOOP_ENGINE] == 0) break;
        /* labels only decrease, so a CC run ends within V + 1 iterations */
        if (ctrl[GZ_LOOP_ITERATIONS] > V + 1) {
            fprintf(stderr, "grazelle: the CC loop ran %llu iterations on %llu vertices\n",
                    (unsigned long long)ctrl[GZ_LOOP_ITERATIONS], (unsigned long long)V);
            exit(72);
        }
    }
    r->iterations = ctrl[GZ_LOOP_ITERATIONS];
    r->pull_iterations = ctrl[GZ_LOOP_PULLS];
    r->push_iterations = ctrl[GZ_LOOP_PUSHES];
    return 1;
}

static uint64_t gz_cc_tail_rounds = 0;

uint64_t gz_cc_push_tail_rounds(void) { return gz_cc_tail_rounds; }

/* Shared loop of execution_cc.c:94-268 and execution_bfs.c:120-276. */
static void gz_run_hybrid(const gz_run_options_t* o, gz_run_result_t* r)
{
    const int bfs = (o->algorithm == GZ_ALG_BFS);
    const uint64_t V = gz_graph.num_vertices, E = gz_graph.num_edges;
    uint64_t vote = 0;
    if (bfs) {
        const double* od = gz_graph_outdegrees_host();
        if (o->stat_mode != GZ_STAT_COUNT_ONLY) vote += (uint64_t)od[gz_graph_device_vertex(o->root)];
        if (o->stat_mode != GZ_STAT_OUTDEG_ONLY) vote += 1;
    } else {
        if (o->stat_mode != GZ_STAT_COUNT_ONLY) vote += E;
        if (o->stat_mode != GZ_STAT_OUTDEG_ONLY) vote += V;
    }
    const uint64_t threshold = (o->stat_mode == GZ_STAT_COUNT_ONLY) ? V / 2 : E / 5;
    if (!bfs && gz_cc_device_loop(o, vote, threshold, r)) return;
    const int learned = (o->engine == GZ_ENGINE_LEARNED) && gz_selector_loaded();
    const int probe = (o->engine == GZ_ENGINE_PROBE) && gz_graph.pull_vectors != NULL && gz_graph.push_vectors != NULL;
    const int features = learned || probe;
    gz_timer_t* t = (o->profile || probe) ? gz_timer_create() : NULL;
    /* the in-degrees feed the "unvisited in-edges" feature; a learned model
     * that does not use it (std 0 in its weights) skips their reads */
    const uint32_t* indeg =
        (probe || (learned && gz_selector_feature_used(3))) ? gz_graph_indegrees_device() : NULL;
    gz_snapshot_t snap;
    memset(&snap, 0, sizeof(snap));
    if (probe) gz_snapshot_alloc(&snap);
    gz_selector_reading_t sel;
    memset(&sel, 0, sizeof(sel));
    /* learned BFS from a root: the host takes every decision, from sums the
     * device leaves next to the vote, and keeps the unvisited counts */
    int host_decides = 0;
    uint64_t sel_unvisited = V, sel_unvisited_in = E;
    if (features) {
        /* features of the first iteration: the initial frontier */
        const uint64_t root = bfs ? gz_graph_device_vertex(o->root) : V;
        const int host_first = learned && bfs && root < V;
        /* with the first decision on the host, the device counts start past
         * the root and no first step runs */
        gz_sel_dev_begin(V, E, root, indeg, host_first);
        if (!host_first)
            gz_sel_dev_step(gz_graph.frontier_has_info, indeg, gz_reduce_buffer, learned, bfs, 1, 0, 0.0f,
                            bfs ? GZ_SEL_FROM_ROOT : GZ_SEL_FROM_BITMAP, 0);
        if (host_first) {
            /* the first frontier is the root alone, so the host knows the
             * features and takes the first decision itself */
            const double* od = gz_graph_outdegrees_host();
            sel.features[0] = logf(1.f + 1.f);
            sel.features[1] = logf(1.f + (float)(uint64_t)od[root]);
            sel.features[2] = logf(1.f + (float)(V - 1));
            sel.features[3] = gz_selector_feature_used(3) ? logf(1.f + (float)(E - gz_graph_indegree_host(root))) : 0.f;
            sel.features[4] = 1.f;
            sel.features[5] = 0.f;
            sel.features[6] = 0.f;
            sel.features[7] = V ? (float)E / (float)V : 0.f;
            sel.choice = gz_selector_host_decide(sel.features, sel.predicted);
            host_decides = 1;
            sel_unvisited = V - 1;
            sel_unvisited_in = E - (gz_selector_feature_used(3) ? gz_graph_indegree_host(root) : 0);
        } else {
            (void)gz_combine_vote(gz_reduce_buffer);
            gz_selector_read(&sel);
        }
        /* BFS activates each vertex once: its edge phases fold the statistics
         * of the next frontier into their vote pass */
        if (bfs) gz_features_fold(1, indeg);
    }
    uint64_t ctr = 0;
    while (1) {
        ctr += 1;
        if (o->max_iterations && ctr > o->max_iterations) {
            ctr -= 1;
            break;
        }
        int use_pull;
        if (o->engine == GZ_ENGINE_PULL)
            use_pull = 1;
        else if (o->engine == GZ_ENGINE_PUSH)
            use_pull = 0;
        else if (learned) {
            use_pull = (sel.choice == 1);
            if (gz_sel_above > 0.0 && (double)vote > gz_sel_above * (double)threshold)
                use_pull = 1;
            else if (gz_sel_below > 0.0 && (double)vote * gz_sel_below < (double)threshold)
                use_pull = 0;
        }
        else
            use_pull = vote > threshold;
        if (use_pull && gz_graph.pull_vectors == NULL) use_pull = 0;
        if (!use_pull && gz_graph.push_vectors == NULL) use_pull = 1;

        gz_iteration_record_t rec;
        memset(&rec, 0, sizeof(rec));
        rec.iteration = ctr;
        rec.engine = use_pull ? 1 : 2;
        rec.frontier_stat = vote;
        memcpy(rec.features, sel.features, sizeof(rec.features));
        memcpy(rec.predicted, sel.predicted, sizeof(rec.predicted));
        gz_phase_config.bfs_level = (uint32_t)ctr;
        float prev_ms = -1.0f;
        /* a small CC frontier: its push iterations run on the device back to
         * back (gz_cc_launch_push_tail); the vote comes back with their count */
        const int tail = !bfs && !use_pull && !features && !o->profile && !o->collect_stats &&
                         (o->engine == GZ_ENGINE_PUSH || o->engine == GZ_ENGINE_HYBRID ||
                          o->engine == GZ_ENGINE_LEARNED) &&
                         gz_cc_push_tail_applies(vote);
        if (probe) {
            /* the chosen engine runs last, so its result is what continues */
            gz_snapshot_take(&snap);
            if (gz_probe_wide_on && bfs) {
                rec.variant_ms[0] = gz_probe_variant(bfs, 1, &snap, t, 0, ~0ull);
                rec.variant_ms[1] = gz_probe_variant(bfs, 0, &snap, t, 1, ~0ull);
                rec.variant_ms[2] = gz_probe_variant(bfs, 0, &snap, t, 0, ~0ull);
            }
            const double other = gz_probe_engine(bfs, !use_pull, &snap, t, 0, vote);
            const double chosen = gz_probe_engine(bfs, use_pull, &snap, t, 1, vote);
            rec.pull_ms = use_pull ? chosen : other;
            rec.push_ms = use_pull ? other : chosen;
            rec.edge_ms = chosen;
        } else {
            if (t) gz_timer_start(t);
            if (learned) gz_sel_dev_stamp(bfs && gz_phase_config.push_frontier_skip);
            phase_op_reset_global_accum();
            if (tail) {
                const uint64_t cap = o->engine == GZ_ENGINE_PUSH ? ~0ull : threshold;
                const uint64_t rounds = o->max_iterations ? o->max_iterations - ctr + 1 : ~0ull;
                gz_cc_launch_push_tail(cap, rounds, gz_reduce_buffer);
            } else {
                gz_push_hint(bfs, use_pull, vote);
                gz_edge_phase(bfs, use_pull);
            }
        }
        if (use_pull)
            r->pull_iterations += 1;
        else
            r->push_iterations += 1;
        /* with BFS statistics folded into the edge phase, the selector's
         * decision kernel also sums the vote */
        const int vote_in_step = features && bfs;
        if (!vote_in_step && !tail) phase_op_write_global_accum_to_buf(gz_reduce_buffer);
        if (t && !probe) gz_timer_stop(t);
        if (host_decides) {
            gz_sel_dev_sums(gz_reduce_buffer, indeg);
        } else if (features) {
            /* the vertices this phase activated are the next frontier */
            gz_sel_dev_step((const uint64_t*)gz_graph.vertex_accumulators, indeg, gz_reduce_buffer, learned, bfs,
                            ctr + 1, rec.engine, prev_ms, bfs ? GZ_SEL_FROM_EDGE_PHASE : GZ_SEL_FROM_BITMAP,
                            vote_in_step);
        }
        const uint64_t next_vote = gz_combine_vote(gz_reduce_buffer);
        if (tail) {
            const uint64_t rounds = gz_reduce_words()[7];
            ctr += rounds - 1;
            r->push_iterations += rounds - 1;
            gz_cc_tail_rounds += rounds;
        }
        if (host_decides)
            gz_selector_host_step(&sel, &sel_unvisited, &sel_unvisited_in, V, E, (float)(ctr + 1), (float)rec.engine);
        else if (features)
            gz_selector_read(&sel);
        if (t && !probe) rec.edge_ms = gz_timer_elapsed_ms(t);
        if (o->collect_stats) gz_phase_stats_read(&rec.vectors, &rec.edges);
        if (o->profile || o->collect_stats || features) gz_record(&rec);
        vote = next_vote;
        if (0 == vote) break;
        gz_swap_frontiers();
        if (bfs)
            perform_vertex_phase_bfs(0, V, NULL);
        else
            perform_vertex_phase_cc(0, V, NULL);
    }
    r->iterations = ctr;
    gz_features_fold(0, NULL);
    if (t) gz_timer_destroy(t);
    if (probe) gz_snapshot_free(&snap);
}

void gz_execution_run(const gz_run_options_t* o, gz_run_result_t* r)
{
    memset(r, 0, sizeof(*r));
    gz_records_count = 0;
    gz_synchronize();
    const double t0 = gz_now_ms();
    if (o->algorithm == GZ_ALG_PR)
        gz_run_pr(o, r);
    else
        gz_run_hybrid(o, r);
    gz_synchronize();
    r->elapsed_ms = gz_now_ms() - t0;
}
