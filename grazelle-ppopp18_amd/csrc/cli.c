/*
 * cli.c -- command line and program flow (main.c:35-167, cmdline.c:46-492).
 *
 * Reference options keep their meaning and error behaviour:
 *   -i base   required; reads base-pull and base-push          (cmdline.c:319-324)
 *   -n num    worker threads: parsed and validated, ignored (one host thread drives the GPU)
 *   -N num    PageRank iterations, >= 1                         (default 1)
 *   -o file   output file, "%llu %.5le" (PR) / "%llu %.0lf"    (graphdata.c:1119-1138)
 *   -s num    vectors per scheduling unit, >= 1 (rounded up to a multiple of 64)
 *   -u list   NUMA nodes: parsed and validated against one node
 *   -V / -h   version / usage
 * Exit codes: 1 unknown option, 2 invalid value, 3 missing value, 4 value not
 * accepted, 5 missing required option, 6 incompatible options
 * (cmdline.c:120-171), 255 unreadable input (graphdata.c:963-967).
 *
 * Options the reference fixed at compile time (ALGORITHM=, EXPERIMENTS=,
 * constants) are runtime switches here:
 *   -a pr|cc|bfs                algorithm (default pr)
 *   -e hybrid|pull|push|learned engine selection (EDGE_FORCE_PULL / _PUSH)
 *   -d damping                  PageRank damping (reference constant 0.875)
 *   -r root                     BFS root (reference constant 0)
 *   -t both|outdeg|count        frontier statistic (THRESHOLD_WITHOUT_COUNT / _OUTDEGREES)
 *   -T                          traditional pull (EDGE_PULL_WITHOUT_SCHED_AWARE)
 *   -M                          CC pull through the merge buffer (EDGE_PULL_FORCE_MERGE)
 *   -E / -X                     edge phase only / vertex phase only
 *   -p                          per-iteration profile CSV on stderr (ITERATION_PROFILE)
 *   -S                          per-iteration vector statistics on stderr (ITERATION_STATS)
 *   -w file                     selector weights for -e learned
 *   -D index                    HIP device
 *   -R                          device ingress of base-pull only, vertices renumbered
 *                               by out-degree (self-loops and duplicate edges dropped;
 *                               results and -r keep original ids)
 *   -L k[,t[,g]]                -R, plus the source-sliced pull layout (windows of 2^k
 *                               sources, t tiers, g LDS blocks) and the source-sliced pull
 *                               for PR and CC, with the 20480-source LDS cache (no
 *                               reference counterpart; chunks of 2048 vectors unless -s)
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grazelle_host.h"
#include "grazelle_phases.h"
#include "grazelle_selector.h"

#define GZ_PROGRAM_NAME "grazelle"
#define GZ_PROGRAM_VERSION "1.0-mi355x"

typedef struct gz_cli_t {
    char gather[1024];
    char scatter[1024];
    const char* output;
    const char* weights;
    uint32_t threads;
    uint32_t iterations;
    uint64_t sched_granularity;
    int device;
    int renumber;
    uint32_t slice_log2, slice_tiers, slice_blocks;
    gz_run_options_t run;
} gz_cli_t;

static int gz_is_switch(char c) { return c == '-'; }

static int gz_recognized(char c) { return strchr("hinNosVuaedrtTMEXpSwDRL", c) != NULL; }

static int gz_requires_value(char c) { return strchr("inNousaedrtwDL", c) != NULL; }

static void gz_err(const char* argv0, const char* fmt, const char* a, const char* b, int code)
{
    printf(fmt, argv0, a, b);
    printf("Try `%s -h' for more information.\n", argv0);
    exit(code);
}

static void gz_usage(const char* argv0)
{
    printf("Usage: %s [options] -i input-graph\n", argv0);
    printf("       %s -h\n       %s -V\n\n", argv0, argv0);
    printf("Required:\n  -i input-graph\n        Path of the input graph; reads input-graph-pull and input-graph-push.\n\n");
    printf("Options:\n");
    printf("  -h\n        Prints this information and exits.\n");
    printf("  -n num-threads\n        Accepted for compatibility; one host thread drives the GPU.\n");
    printf("  -N num-iterations\n        PageRank iterations. Defaults to 1.\n");
    printf("  -o output-file\n        Path of the file to write as output.\n");
    printf("  -s vectors-per-unit\n        Pull/push scheduling unit in vectors (rounded up to 64). Default 512.\n");
    printf("  -u node1[,node2...]\n        Accepted for compatibility; only node 0 exists on one GPU.\n");
    printf("  -a pr|cc|bfs\n        Algorithm. Defaults to pr.\n");
    printf("  -e hybrid|pull|push|learned\n        Engine selection. Defaults to hybrid.\n");
    printf("  -d damping\n        PageRank damping factor. Defaults to 0.875.\n");
    printf("  -r root\n        BFS search root. Defaults to 0.\n");
    printf("  -t both|outdeg|count\n        Frontier statistic for the engine threshold. Defaults to both.\n");
    printf("  -T\n        Traditional (atomic) pull instead of scheduler-aware merge buffers.\n");
    printf("  -M\n        CC: boundary runs of the pull through the merge buffer (plain stores elsewhere).\n");
    printf("  -E, -X\n        Edge phase only / vertex phase only (PageRank).\n");
    printf("  -p\n        Per-iteration profile CSV on stderr.\n");
    printf("  -S\n        Per-iteration vector statistics CSV on stderr.\n");
    printf("  -w weights\n        Selector weights for -e learned.\n");
    printf("  -D device\n        HIP device index. Defaults to 0.\n");
    printf("  -R\n        Build the graph on the device from input-graph-pull alone, vertices renumbered\n"
           "        by out-degree (drops self-loops and duplicate edges; output keeps original ids).\n");
    printf("  -L k[,tiers[,blocks]]\n        As -R, plus the source-sliced pull for PageRank and CC: windows of 2^k sources\n"
           "        (1..23), 1..3 tiers, 0..6 LDS blocks of %u sources (tiers + blocks <= 7).\n",
           GZ_SLICE_BLOCK_VERTICES);
    printf("  -V\n        Prints version information and exits.\n");
    exit(0);
}

static uint64_t gz_parse_u64(const char* argv0, const char* opt, const char* val, uint64_t min)
{
    char* end = NULL;
    if (!isdigit((unsigned char)val[0])) gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
    const unsigned long long x = strtoull(val, &end, 10);
    if (*end != '\0' || x < min) gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
    return (uint64_t)x;
}

static void gz_parse_option(gz_cli_t* c, const char* argv0, const char* opt, const char* val)
{
    if (!gz_is_switch(opt[0]) || opt[1] == '\0' || opt[2] != '\0' || !gz_recognized(opt[1]))
        gz_err(argv0, "%s: Unrecognized option `%s'.%s\n", opt, "", 1);
    if (gz_requires_value(opt[1]) && val == NULL) gz_err(argv0, "%s: Missing argument for option `%s'.%s\n", opt, "", 3);
    if (!gz_requires_value(opt[1]) && val != NULL && opt[1] != 'h')
        gz_err(argv0, "%s: Option `%s' does not accept an argument.%s\n", opt, "", 4);
    switch (opt[1]) {
    case 'h':
        gz_usage(argv0);
        break;
    case 'V':
        printf("%s v%s for HIP/gfx950\n", GZ_PROGRAM_NAME, GZ_PROGRAM_VERSION);
        exit(0);
    case 'i':
        snprintf(c->gather, sizeof(c->gather), "%s-pull", val);
        snprintf(c->scatter, sizeof(c->scatter), "%s-push", val);
        break;
    case 'n':
        c->threads = (uint32_t)gz_parse_u64(argv0, opt, val, 0);
        break;
    case 'N':
        c->iterations = (uint32_t)gz_parse_u64(argv0, opt, val, 1);
        break;
    case 'o':
        c->output = val;
        break;
    case 's':
        c->sched_granularity = gz_parse_u64(argv0, opt, val, 1);
        break;
    case 'u': {
        /* only NUMA node 0 exists: every listed node must be 0 */
        const char* p = val;
        while (*p) {
            char* end = NULL;
            if (!isdigit((unsigned char)*p)) gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
            const long node = strtol(p, &end, 10);
            if (node != 0 || (*end != '\0' && *end != ',')) gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
            p = (*end == ',' && end[1] != '\0') ? end + 1 : end + (*end == ',' ? 1 : 0);
        }
        break;
    }
    case 'a':
        if (!strcmp(val, "pr"))
            c->run.algorithm = GZ_ALG_PR;
        else if (!strcmp(val, "cc"))
            c->run.algorithm = GZ_ALG_CC;
        else if (!strcmp(val, "bfs"))
            c->run.algorithm = GZ_ALG_BFS;
        else
            gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
        break;
    case 'e':
        if (!strcmp(val, "hybrid"))
            c->run.engine = GZ_ENGINE_HYBRID;
        else if (!strcmp(val, "pull"))
            c->run.engine = GZ_ENGINE_PULL;
        else if (!strcmp(val, "push"))
            c->run.engine = GZ_ENGINE_PUSH;
        else if (!strcmp(val, "learned"))
            c->run.engine = GZ_ENGINE_LEARNED;
        else
            gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
        break;
    case 'd': {
        char* end = NULL;
        const double d = strtod(val, &end);
        if (*end != '\0' || !(d > 0.0 && d < 1.0)) gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
        c->run.damping = d;
        break;
    }
    case 'r':
        c->run.root = gz_parse_u64(argv0, opt, val, 0);
        break;
    case 't':
        if (!strcmp(val, "both"))
            c->run.stat_mode = GZ_STAT_COUNT_AND_OUTDEG;
        else if (!strcmp(val, "outdeg"))
            c->run.stat_mode = GZ_STAT_OUTDEG_ONLY;
        else if (!strcmp(val, "count"))
            c->run.stat_mode = GZ_STAT_COUNT_ONLY;
        else
            gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
        break;
    case 'T':
        c->run.pull_variant = GZ_PULL_TRADITIONAL;
        break;
    case 'M':
        c->run.pull_variant = GZ_PULL_MERGE_BUFFER;
        break;
    case 'E':
        c->run.edge_only = 1;
        break;
    case 'X':
        c->run.vertex_only = 1;
        break;
    case 'p':
        c->run.profile = 1;
        break;
    case 'S':
        c->run.collect_stats = 1;
        break;
    case 'w':
        c->weights = val;
        break;
    case 'D':
        c->device = (int)gz_parse_u64(argv0, opt, val, 0);
        break;
    case 'R':
        c->renumber = 1;
        break;
    case 'L': {
        /* k[,tiers[,blocks]] */
        unsigned long long f[3] = {0, 1, 0};
        int n = 0;
        const char* p = val;
        while (n < 3) {
            if (!isdigit((unsigned char)*p)) gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
            char* end = NULL;
            f[n++] = strtoull(p, &end, 10);
            p = end;
            if (*p == '\0') break;
            if (*p != ',' || n == 3) gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
            p += 1;
        }
        if (f[0] < 1 || f[0] > 23 || f[1] < 1 || f[1] > 3 || f[2] > 6 || f[1] + f[2] > 7)
            gz_err(argv0, "%s: Invalid value `%s' for option `%s'.\n", val, opt, 2);
        c->slice_log2 = (uint32_t)f[0];
        c->slice_tiers = (uint32_t)f[1];
        c->slice_blocks = (uint32_t)f[2];
        c->renumber = 1;
        break;
    }
    default:
        gz_err(argv0, "%s: Unrecognized option `%s'.%s\n", opt, "", 1);
    }
}

int gz_main(int argc, char** argv)
{
    gz_cli_t c;
    memset(&c, 0, sizeof(c));
    gz_run_options_default(&c.run);
    c.iterations = 1;
    const char* argv0 = argc > 0 ? argv[0] : GZ_PROGRAM_NAME;
    for (int i = 1; i < argc; ++i) {
        const char* opt = argv[i];
        const char* val = NULL;
        if (i + 1 < argc && isprint((unsigned char)argv[i + 1][0]) && !gz_is_switch(argv[i + 1][0])) {
            val = argv[i + 1];
            i += 1;
        }
        gz_parse_option(&c, argv0, opt, val);
    }
    if (c.gather[0] == '\0') gz_err(argv0, "%s: Missing required option `-%s'.%s\n", "i", "", 5);
    if (c.run.edge_only && c.run.vertex_only) gz_err(argv0, "%s: Specified option combination is incompatible.%s%s\n", "", "", 6);
    if (c.run.engine == GZ_ENGINE_LEARNED && c.weights == NULL)
        gz_err(argv0, "%s: Specified option combination is incompatible.%s%s\n", "", "", 6);
    if (c.slice_log2 && (c.run.pull_variant == GZ_PULL_TRADITIONAL || c.run.pull_variant == GZ_PULL_MERGE_BUFFER))
        gz_err(argv0, "%s: Specified option combination is incompatible.%s%s\n", "", "", 6);
    c.run.iterations = c.iterations;

    /* unreadable inputs exit 255 before the device is touched (graphdata.c:963-967) */
    const int need_push = !c.renumber && !(c.run.algorithm == GZ_ALG_PR && c.run.engine != GZ_ENGINE_PUSH);
    const char* files[2] = {c.gather, need_push ? c.scatter : NULL};
    for (int k = 0; k < 2; ++k) {
        if (files[k] == NULL) continue;
        FILE* f = fopen(files[k], "rb");
        uint64_t hdr[2];
        const int ok = f != NULL && fread(hdr, sizeof(uint64_t), 2, f) == 2;
        if (f) fclose(f);
        if (!ok) {
            fprintf(stderr, "Error: unable to read file \"%s\"\n", files[k]);
            exit(255);
        }
    }
    gz_runtime_init(c.device);
    if (c.sched_granularity)
        gz_phase_config.chunk_vectors = (uint32_t)((c.sched_granularity + 63) / 64 * 64);
    else if (c.slice_log2)
        gz_phase_config.chunk_vectors = 2048;
    if (c.renumber) {
        uint64_t nv = 0, ne = 0, *pairs = NULL;
        if (gz_edge_list_read(c.gather, &nv, &ne, &pairs) != 0) {
            fprintf(stderr, "Error: unable to read file \"%s\"\n", c.gather);
            exit(255);
        }
        uint32_t flags = GZ_INGRESS_RELABEL_BY_OUTDEGREE;
        if (c.slice_log2)
            flags |= GZ_INGRESS_SOURCE_SLICES(c.slice_log2) | GZ_INGRESS_SLICE_TIERS(c.slice_tiers) |
                     GZ_INGRESS_SLICE_BLOCKS(c.slice_blocks);
        gz_graph_build_pairs_device(nv, ne, pairs, 0, flags);
        free(pairs);
        printf("Device ingress: %llu vertices, %llu edges, %llu pull / %llu push vectors, renumbered by out-degree\n",
               (unsigned long long)gz_graph_num_vertices(), (unsigned long long)gz_graph_num_edges(),
               (unsigned long long)gz_graph_pull_vector_count(), (unsigned long long)gz_graph_push_vector_count());
        if (c.slice_log2) {
            uint64_t S = 0, H = 0;
            uint32_t T = 0, G = 0;
            gz_graph_source_slices_info(NULL, &S, &T, &G, &H);
            c.run.pull_variant = GZ_PULL_SOURCE_SLICED;
            gz_slices_set_hot_cache(GZ_SLICE_BLOCK_VERTICES);
            printf("Source-sliced pull: windows of %llu sources, %u tier(s), %u LDS block(s), LDS cache %u\n",
                   (unsigned long long)S, T, G, GZ_SLICE_BLOCK_VERTICES);
        }
    } else {
        graph_data_read_from_file(c.gather, need_push ? c.scatter : NULL, 1, NULL);
    }
    graph_data_allocate_merge_buffers(c.threads, 1, NULL);
    printf("Scheduler: total units = %llu, vectors per unit = %u\n",
           (unsigned long long)gz_pull_units_for(gz_graph_pull_vector_count()), gz_phase_config.chunk_vectors);
    if (c.weights && gz_selector_load_file(c.weights) != 0) {
        fprintf(stderr, "Error: unable to read selector weights \"%s\"\n", c.weights);
        exit(255);
    }

    gz_execution_prepare(&c.run);
    printf("Starting execution.\n");
    gz_run_result_t r;
    gz_execution_run(&c.run, &r);
    printf("Execution completed.\n");
    const double E = (double)gz_graph_num_edges();
    printf("\n------------ EXECUTION STATISTICS ------------\n");
    printf("%-25s = %.2lfms\n", "Running Time", r.elapsed_ms);
    if (c.run.algorithm == GZ_ALG_PR) {
        printf("%-25s = %.0lf Medges/sec\n", "Processing Rate", E * (double)c.run.iterations / r.elapsed_ms / 1000.0);
        const uint64_t n = gz_graph_num_vertices();
        double* ranks = (double*)malloc(n * sizeof(double));
        gz_result_pr_ranks(ranks);
        double sum = 0.0;
        for (uint64_t i = 0; i < n; ++i) sum += ranks[i];
        free(ranks);
        printf("%-25s = %.10lf\n", "PageRank Sum", sum);
    } else {
        printf("%-25s = %.0lf Medges/sec\n", "Effective Processing Rate", E * (double)r.iterations / r.elapsed_ms / 1000.0);
    }
    printf("%-25s = %llu\n", "Total Iterations", (unsigned long long)r.iterations);
    printf("%-25s = %llu\n", "Pull-Based Iterations", (unsigned long long)r.pull_iterations);
    printf("%-25s = %llu\n", "Push-Based Iterations", (unsigned long long)r.push_iterations);
    printf("----------------------------------------------\n");
    if (c.run.profile || c.run.collect_stats) {
        const uint64_t n = gz_execution_records(NULL, 0);
        gz_iteration_record_t* recs = (gz_iteration_record_t*)malloc((n ? n : 1) * sizeof(*recs));
        gz_execution_records(recs, n);
        const double cmp = (c.run.stat_mode == GZ_STAT_COUNT_ONLY) ? (double)gz_graph_num_vertices() : E;
        if (c.run.profile) {
            fprintf(stderr, "Iteration,Selected Engine,Edge Phase Execution Time (ms),Frontier / Comparator\n");
            for (uint64_t i = 0; i < n; ++i)
                fprintf(stderr, "%llu,%s,%.6lf,%.10lf\n", (unsigned long long)recs[i].iteration,
                        recs[i].engine == 1 ? "Pull" : "Push", recs[i].edge_ms, (double)recs[i].frontier_stat / cmp);
        }
        if (c.run.collect_stats) {
            fprintf(stderr, "Iteration,# Vectors,Packing Efficiency\n");
            for (uint64_t i = 0; i < n; ++i)
                fprintf(stderr, "%llu,%llu,%lf\n", (unsigned long long)recs[i].iteration,
                        (unsigned long long)recs[i].vectors,
                        recs[i].vectors ? (double)recs[i].edges / (4.0 * (double)recs[i].vectors) : 0.0);
        }
        free(recs);
    }
    if (c.output) graph_data_write_ranks_to_file(c.output, c.run.algorithm);
    gz_graph_release();
    gz_runtime_shutdown();
    return 0;
}
