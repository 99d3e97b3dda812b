/*
 * rmat.c -- seeded R-MAT edge generator (the reference ships no generator;
 * its inputs are binary -pull / -push edge lists, README.md:50-62).
 *
 * Edge e draws `scale` quadrant choices, most significant bit first.  The
 * random stream of edge e is x_0 = splitmix64(seed ^ splitmix64(offset + e)),
 * x_{k+1} = splitmix64(x_k); each 64-bit word supplies four 16-bit draws u,
 * and u < A picks (src bit 0, dst bit 0), u < A+B picks (0, 1), u < A+B+C
 * picks (1, 0), otherwise (1, 1), with A, B, C the probabilities scaled to
 * 65536 and truncated.  The stream depends only on (seed, offset + e), so any
 * thread count produces the same edges.  Ids at or above num_vertices (when
 * it is not 2^scale) are folded by modulo.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "grazelle_host.h"

uint64_t gz_splitmix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

typedef struct gz_rmat_job_t {
    const gz_rmat_params_t* p;
    uint64_t* pairs;
    uint64_t first, last;
} gz_rmat_job_t;

static void* gz_rmat_worker(void* arg)
{
    const gz_rmat_job_t* j = (const gz_rmat_job_t*)arg;
    const gz_rmat_params_t* p = j->p;
    const uint32_t ta = (uint32_t)(p->a * 65536.0);
    const uint32_t tb = (uint32_t)((p->a + p->b) * 65536.0);
    const uint32_t tc = (uint32_t)((p->a + p->b + p->c) * 65536.0);
    const uint64_t nv = p->num_vertices ? p->num_vertices : (1ull << p->scale);
    for (uint64_t e = j->first; e < j->last; ++e) {
        uint64_t x = gz_splitmix64(p->seed ^ gz_splitmix64(p->edge_offset + e));
        uint64_t s = 0, d = 0;
        for (uint32_t l = 0; l < p->scale; ++l) {
            if (l && (l & 3) == 0) x = gz_splitmix64(x);
            const uint32_t u = (uint32_t)((x >> (16 * (l & 3))) & 0xffff);
            const uint32_t sb = (u >= tb) ? 1u : 0u;
            const uint32_t db = (u >= ta && u < tb) || (u >= tc) ? 1u : 0u;
            s = (s << 1) | sb;
            d = (d << 1) | db;
        }
        if (s >= nv) s %= nv;
        if (d >= nv) d %= nv;
        j->pairs[2 * e] = s;
        j->pairs[2 * e + 1] = d;
    }
    return NULL;
}

void gz_rmat_draw(const gz_rmat_params_t* p, uint64_t* pairs)
{
    const uint64_t n = p->num_edges;
    int threads = 8;
    if (n < 1000000) threads = 1;
    pthread_t tid[64];
    gz_rmat_job_t job[64];
    for (int t = 0; t < threads; ++t) {
        job[t].p = p;
        job[t].pairs = pairs;
        job[t].first = n * (uint64_t)t / (uint64_t)threads;
        job[t].last = n * (uint64_t)(t + 1) / (uint64_t)threads;
    }
    if (threads == 1) {
        gz_rmat_worker(&job[0]);
        return;
    }
    for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, gz_rmat_worker, &job[t]);
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}
