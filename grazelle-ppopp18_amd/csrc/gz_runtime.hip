/*
 * gz_runtime.hip -- the thin C-ABI runtime that replaces threads.c and
 * numanodes.c (threads.h:37-82, numanodes.h:27-61), plus the common phase
 * operators of phases.h:49-70:
 *   phase_op_reset_global_accum          (phases.asm:24-30)
 *   phase_op_write_global_accum_to_buf   (phases.asm:34-44)
 *   phase_op_combine_global_var_from_buf (phases.c:27-37)
 *   edge_pull_op_merge_with_merge_buffer (phases.c:41-70)
 * One host thread drives one device through one stream; work that the
 * reference split over threads is split over wavefronts.
 */
#include "gz_internal.h"
#include "gz_frontier.h"

#include <string.h>

gz_phase_graph_t gz_graph;
gz_phase_config_t gz_phase_config = {
    0.875,                    /* damping: const_damping_factor, constants.asm:75-79 */
    GZ_STAT_COUNT_AND_OUTDEG, /* stat_mode */
    GZ_PULL_SCHED_AWARE,      /* pull_variant */
    512,                      /* chunk_vectors */
    1,                        /* bfs_level */
    0,                        /* collect_stats */
    1,                        /* push_frontier_skip */
    1,                        /* vectors_per_lane */
};
mergeaccum_t* gz_merge_buffer = nullptr;
uint64_t gz_merge_buffer_count = 0;

gz_scratch_t gz_scratch;
uint32_t gz_push_queue_cap = GZ_PUSH_QUEUE_PIECES;

extern "C" int gz_push_queue_capacity(uint32_t pieces)
{
    if (pieces == 0 || pieces > GZ_PUSH_QUEUE_PIECES) return -1;
    gz_push_queue_cap = pieces;
    return 0;
}
uint32_t gz_cu_count = 1;
hipStream_t gz_stream = nullptr;
static int gz_device_id = -1;
static char gz_device_name_buf[256];
static void gz_noalloc_setup(int device);

/* ---- runtime hooks -------------------------------------------------------- */

extern "C" int gz_runtime_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int gz_runtime_init(int device)
{
    if (gz_stream != nullptr) return 0;
    int n = gz_runtime_device_count();
    if (n <= 0) {
        fprintf(stderr, "grazelle: no HIP device is visible\n");
        exit(70);
    }
    if (device < 0 || device >= n) device = 0;
    GZ_CHECK(hipSetDevice(device));
    GZ_CHECK(hipStreamCreateWithFlags(&gz_stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    GZ_CHECK(hipGetDeviceProperties(&prop, device));
    /* the marketing name can be empty: the architecture, CU count and
     * memory size then name the device */
    snprintf(gz_device_name_buf, sizeof(gz_device_name_buf), "%s%s%s, %d CUs, %.0f GiB", prop.name,
             prop.name[0] ? " " : "", prop.gcnArchName, prop.multiProcessorCount,
             (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    gz_cu_count = prop.multiProcessorCount > 0 ? (uint32_t)prop.multiProcessorCount : 1u;
    gz_device_id = device;
    gz_noalloc_setup(device);
    memset(&gz_scratch, 0, sizeof(gz_scratch));
    return 0;
}

extern "C" const char* gz_runtime_device_name(void)
{
    return gz_device_id >= 0 ? gz_device_name_buf : "none";
}

static void gz_require_runtime(void)
{
    if (gz_stream == nullptr) gz_runtime_init(0);
}

extern "C" void* gz_runtime_stream(void)
{
    gz_require_runtime();
    return (void*)gz_stream;
}

/* Bumped by every device allocation and free: a captured launch sequence is
 * stale once any buffer may have moved. */
static uint64_t gz_alloc_generation = 0;

extern "C" void* gz_device_malloc(size_t bytes)
{
    gz_require_runtime();
    ++gz_alloc_generation;
    void* p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t err = hipMalloc(&p, bytes);
    if (err != hipSuccess) {
        fprintf(stderr, "grazelle: device allocation of %zu bytes failed: %s\n", bytes, hipGetErrorString(err));
        exit(71);
    }
    gz_device_zero(p, bytes);
    return p;
}

extern "C" void gz_device_free(void* ptr)
{
    if (ptr != nullptr) {
        ++gz_alloc_generation;
        GZ_CHECK(hipFree(ptr));
    }
}

extern "C" uint64_t gz_device_allocation_events(void) { return gz_alloc_generation; }

/* ---- replay of a captured launch sequence ----------------------------------
 * One cached executable graph: `times` rounds of the launches enqueue() makes
 * on the phase stream, captured into one graph (one graph launch costs more
 * here than the few kernel launches of a round, so the rounds are not
 * replayed one launch each) and launched again while the caller's key, the
 * round count and the allocation generation stay the same.  A request is
 * captured the second time in a row it comes: a run that never repeats (one
 * run per graph of a batch) pays for no capture it never replays, and a long
 * run gains nothing from one (its kernels hide the launches; the first
 * capture in a process costs milliseconds).  enqueue() must neither allocate
 * nor synchronise (the capture would fail, and the process exits with the
 * HIP error). */
#define GZ_REPLAY_KEY_WORDS 64
struct gz_replay_key_t {
    uint64_t generation;
    uint32_t times;
    size_t n;
    uint64_t key[GZ_REPLAY_KEY_WORDS];
};
/* One cached graph per slot: GZ_REPLAY_SLOTS kinds of loops (PageRank
 * iterations, CC iterations) keep their captures apart. */
#define GZ_REPLAY_SLOTS 2
static struct gz_replay_slot_t {
    hipGraphExec_t exec;
    gz_replay_key_t captured; /* what exec holds */
    gz_replay_key_t seen;     /* the last configuration asked for and not captured */
    bool seen_valid;
} gz_replay[GZ_REPLAY_SLOTS];
static uint64_t gz_replay_captures, gz_replay_launches;

extern "C" void gz_graph_replay_counts(uint64_t* captures, uint64_t* launches)
{
    if (captures) *captures = gz_replay_captures;
    if (launches) *launches = gz_replay_launches;
}

static void gz_replay_release(void)
{
    for (auto& r : gz_replay)
        if (r.exec != nullptr) GZ_CHECK(hipGraphExecDestroy(r.exec));
    memset(gz_replay, 0, sizeof(gz_replay));
}

static bool gz_replay_key_equal(const gz_replay_key_t& a, const uint64_t* key, size_t n, uint32_t times)
{
    return a.generation == gz_alloc_generation && a.times == times && a.n == n &&
           memcmp(a.key, key, n * sizeof(uint64_t)) == 0;
}

extern "C" int gz_graph_replay_cached_slot(uint32_t slot, const uint64_t* key, size_t n, void (*enqueue)(void*),
                                           void* ctx, uint32_t times)
{
    gz_require_runtime();
    if (slot >= GZ_REPLAY_SLOTS) {
        fprintf(stderr, "grazelle: replay slot %u (there are %d)\n", slot, GZ_REPLAY_SLOTS);
        exit(72);
    }
    gz_replay_slot_t& R = gz_replay[slot];
    if (times == 0) return 1;
    if (n > GZ_REPLAY_KEY_WORDS) {
        fprintf(stderr, "grazelle: replay key of %zu words (at most %d)\n", n, GZ_REPLAY_KEY_WORDS);
        exit(72);
    }
    const bool hit = R.exec != nullptr && gz_replay_key_equal(R.captured, key, n, times);
    if (!hit) {
        const bool repeat = R.seen_valid && gz_replay_key_equal(R.seen, key, n, times);
        R.seen.generation = gz_alloc_generation;
        R.seen.times = times;
        R.seen.n = n;
        memcpy(R.seen.key, key, n * sizeof(uint64_t));
        R.seen_valid = true;
        /* a first request: the caller runs the rounds itself */
        if (!repeat) return 0;
        if (R.exec != nullptr) GZ_CHECK(hipGraphExecDestroy(R.exec));
        R.exec = nullptr;
        hipGraph_t g = nullptr;
        GZ_CHECK(hipStreamBeginCapture(gz_stream, hipStreamCaptureModeThreadLocal));
        for (uint32_t i = 0; i < times; ++i) enqueue(ctx);
        GZ_CHECK(hipStreamEndCapture(gz_stream, &g));
        GZ_CHECK(hipGraphInstantiate(&R.exec, g, nullptr, nullptr, 0));
        GZ_CHECK(hipGraphDestroy(g));
        R.captured = R.seen;
        R.seen_valid = false;
        gz_replay_captures += 1;
    }
    GZ_CHECK(hipGraphLaunch(R.exec, gz_stream));
    gz_replay_launches += 1;
    return 1;
}

extern "C" int gz_graph_replay_cached(const uint64_t* key, size_t n, void (*enqueue)(void*), void* ctx,
                                      uint32_t times)
{
    return gz_graph_replay_cached_slot(0, key, n, enqueue, ctx, times);
}

extern "C" void gz_copy_to_device(void* dst, const void* src, size_t bytes)
{
    gz_require_runtime();
    if (bytes == 0) return;
    GZ_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, gz_stream));
    GZ_CHECK(hipStreamSynchronize(gz_stream));
}

extern "C" void gz_copy_to_host(void* dst, const void* src, size_t bytes)
{
    gz_require_runtime();
    if (bytes == 0) return;
    GZ_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, gz_stream));
    GZ_CHECK(hipStreamSynchronize(gz_stream));
}

__global__ void gz_k_zero(uint64_t* __restrict__ p, uint64_t words)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void gz_k_zero_bytes(unsigned char* __restrict__ p, uint64_t bytes)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

extern "C" void gz_device_zero(void* ptr, size_t bytes)
{
    gz_require_runtime();
    if (bytes == 0 || ptr == nullptr) return;
    const uint64_t words = bytes / 8;
    if (((uintptr_t)ptr & 7) == 0 && words > 0) {
        uint64_t blocks = (words + GZ_BLOCK - 1) / GZ_BLOCK;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(gz_k_zero, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, (uint64_t*)ptr, words);
        GZ_LAUNCH_CHECK();
        const uint64_t tail = bytes - words * 8;
        if (tail) {
            hipLaunchKernelGGL(gz_k_zero_bytes, dim3(1), dim3(64), 0, gz_stream, (unsigned char*)ptr + words * 8, tail);
            GZ_LAUNCH_CHECK();
        }
    } else {
        uint64_t blocks = (bytes + GZ_BLOCK - 1) / GZ_BLOCK;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(gz_k_zero_bytes, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, (unsigned char*)ptr, (uint64_t)bytes);
        GZ_LAUNCH_CHECK();
    }
}

__global__ __launch_bounds__(GZ_BLOCK) void gz_k_copy(gz_u64x2* __restrict__ dst, const gz_u64x2* __restrict__ src,
                                                      uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" void gz_device_copy(void* dst, const void* src, size_t bytes)
{
    gz_require_runtime();
    if (bytes == 0) return;
    if ((((uintptr_t)dst | (uintptr_t)src | bytes) & 15) == 0) {
        const uint64_t n = bytes / 16;
        uint64_t blocks = (n + GZ_BLOCK - 1) / GZ_BLOCK;
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(gz_k_copy, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, (gz_u64x2*)dst,
                           (const gz_u64x2*)src, n);
        GZ_LAUNCH_CHECK();
    } else {
        GZ_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, gz_stream));
    }
}

extern "C" void gz_synchronize(void)
{
    gz_require_runtime();
    GZ_CHECK(hipStreamSynchronize(gz_stream));
}

extern "C" uint64_t* gz_reduce_buffer_create(void)
{
    return (uint64_t*)gz_device_malloc(8 * sizeof(uint64_t));
}

extern "C" void gz_reduce_buffer_destroy(uint64_t* reduce_buffer)
{
    gz_device_free(reduce_buffer);
}

struct gz_timer_t {
    hipEvent_t start, stop;
};

extern "C" gz_timer_t* gz_timer_create(void)
{
    gz_require_runtime();
    gz_timer_t* t = (gz_timer_t*)malloc(sizeof(gz_timer_t));
    GZ_CHECK(hipEventCreate(&t->start));
    GZ_CHECK(hipEventCreate(&t->stop));
    return t;
}

extern "C" void gz_timer_destroy(gz_timer_t* t)
{
    if (!t) return;
    GZ_CHECK(hipEventDestroy(t->start));
    GZ_CHECK(hipEventDestroy(t->stop));
    free(t);
}

extern "C" void gz_timer_start(gz_timer_t* t) { GZ_CHECK(hipEventRecord(t->start, gz_stream)); }
extern "C" void gz_timer_stop(gz_timer_t* t) { GZ_CHECK(hipEventRecord(t->stop, gz_stream)); }

extern "C" float gz_timer_elapsed_ms(gz_timer_t* t)
{
    float ms = 0.f;
    GZ_CHECK(hipEventSynchronize(t->stop));
    GZ_CHECK(hipEventElapsedTime(&ms, t->start, t->stop));
    return ms;
}

/* ---- per-kernel timing ---------------------------------------------------- */

#define GZ_KT_POOL 2048
static int gz_kt_on = 0;
static hipEvent_t* gz_kt_ev = nullptr; /* [GZ_KT_POOL][2] */
static int gz_kt_ids[GZ_KT_POOL];
static int gz_kt_used = 0;
static int gz_kt_open = -1;
static double gz_kt_ms[GZ_K_COUNT];
static uint64_t gz_kt_launches[GZ_K_COUNT];

static void gz_kt_flush(void)
{
    for (int i = 0; i < gz_kt_used; ++i) {
        float ms = 0.f;
        GZ_CHECK(hipEventSynchronize(gz_kt_ev[2 * i + 1]));
        GZ_CHECK(hipEventElapsedTime(&ms, gz_kt_ev[2 * i], gz_kt_ev[2 * i + 1]));
        gz_kt_ms[gz_kt_ids[i]] += ms;
        gz_kt_launches[gz_kt_ids[i]] += 1;
    }
    gz_kt_used = 0;
}

extern "C" int gz_kernel_timing_enabled(void) { return gz_kt_on; }

extern "C" void gz_features_fold(int on, const uint32_t* indeg)
{
    gz_scratch.features_on = on ? 1u : 0u;
    gz_scratch.features_indeg = on ? indeg : nullptr;
}

extern "C" void gz_kernel_timing_enable(int on)
{
    gz_require_runtime();
    if (on && gz_kt_ev == nullptr) {
        gz_kt_ev = (hipEvent_t*)malloc(2 * GZ_KT_POOL * sizeof(hipEvent_t));
        for (int i = 0; i < 2 * GZ_KT_POOL; ++i) GZ_CHECK(hipEventCreate(&gz_kt_ev[i]));
    }
    if (!on && gz_kt_ev != nullptr) gz_kt_flush();
    gz_kt_on = on;
}

extern "C" void gz_kernel_timing_reset(void)
{
    if (gz_kt_ev != nullptr) gz_kt_flush();
    for (int i = 0; i < GZ_K_COUNT; ++i) {
        gz_kt_ms[i] = 0.0;
        gz_kt_launches[i] = 0;
    }
}

extern "C" void gz_kernel_timing_read(int id, double* total_ms, uint64_t* launches)
{
    if (gz_kt_ev != nullptr) gz_kt_flush();
    *total_ms = (id >= 0 && id < GZ_K_COUNT) ? gz_kt_ms[id] : 0.0;
    *launches = (id >= 0 && id < GZ_K_COUNT) ? gz_kt_launches[id] : 0;
}

void gz_ktimer_begin(int id)
{
    if (!gz_kt_on) return;
    if (gz_kt_used == GZ_KT_POOL) gz_kt_flush();
    gz_kt_open = id;
    GZ_CHECK(hipEventRecord(gz_kt_ev[2 * gz_kt_used], gz_stream));
}

void gz_ktimer_end(int id)
{
    if (!gz_kt_on || gz_kt_open != id) return;
    GZ_CHECK(hipEventRecord(gz_kt_ev[2 * gz_kt_used + 1], gz_stream));
    gz_kt_ids[gz_kt_used++] = id;
    gz_kt_open = -1;
}

/* ---- scratch -------------------------------------------------------------- */

uint32_t gz_vectors_per_lane(void)
{
    const uint32_t v = gz_phase_config.vectors_per_lane;
    return (v >= 4) ? 4u : (v >= 2 ? 2u : 1u);
}

uint64_t gz_chunk_len(void)
{
    /* a multiple of one wavefront step (64 lanes x vectors per lane) */
    const uint64_t step = (uint64_t)GZ_WAVE * gz_vectors_per_lane();
    uint64_t ch = gz_phase_config.chunk_vectors;
    if (ch < step) ch = step;
    return (ch + step - 1) / step * step;
}

uint64_t gz_chunks_for(uint64_t n)
{
    const uint64_t ch = gz_chunk_len();
    return (n + ch - 1) / ch;
}

extern "C" uint64_t gz_pull_units_for(uint64_t vector_count) { return gz_chunks_for(vector_count); }

void gz_scratch_reserve(uint64_t waves, uint64_t num_vertices)
{
    gz_require_runtime();
    if (waves < (uint64_t)GZ_PERSISTENT_BLOCKS * GZ_WAVES_PER_BLOCK) waves = (uint64_t)GZ_PERSISTENT_BLOCKS * GZ_WAVES_PER_BLOCK;
    if (waves > gz_scratch.partials_capacity) {
        gz_device_free(gz_scratch.partials);
        gz_device_free(gz_scratch.stats);
        gz_device_free(gz_scratch.features);
        gz_scratch.partials = gz_device_malloc(waves * sizeof(uint64_t));
        gz_scratch.stats = (uint64_t*)gz_device_malloc(waves * 2 * sizeof(uint64_t));
        gz_scratch.features = (uint64_t*)gz_device_malloc(waves * 3 * sizeof(uint64_t));
        gz_scratch.partials_capacity = waves;
    }
    if (gz_scratch.reduce_tmp == nullptr) gz_scratch.reduce_tmp = (uint64_t*)gz_device_malloc(2048 * sizeof(uint64_t));
    if (gz_scratch.push_queue == nullptr) {
        gz_scratch.push_pieces = (uint64_t*)gz_device_malloc((size_t)GZ_PUSH_QUEUE_PIECES * 3 * sizeof(uint64_t));
        gz_scratch.push_queue = (uint64_t*)gz_device_malloc(2 * sizeof(uint64_t));
        GZ_CHECK(hipMemsetAsync(gz_scratch.push_queue, 0, 2 * sizeof(uint64_t), gz_stream));
        gz_scratch.push_queue_parity = 0;
    }
    if (num_vertices > gz_scratch.scratch_vertices) {
        gz_device_free(gz_scratch.active_vertices);
        gz_device_free(gz_scratch.active_prefix);
        gz_device_free(gz_scratch.bfs_first_vector);
        gz_scratch.active_vertices = (uint64_t*)gz_device_malloc(num_vertices * sizeof(uint64_t));
        gz_scratch.active_prefix = (uint64_t*)gz_device_malloc((num_vertices + 1) * sizeof(uint64_t));
        gz_scratch.bfs_first_vector = (uint64_t*)gz_device_malloc(num_vertices * sizeof(uint64_t));
        GZ_CHECK(hipMemsetAsync(gz_scratch.bfs_first_vector, 0xff, num_vertices * sizeof(uint64_t), gz_stream));
        const uint64_t blocks = gz_compact_blocks(num_vertices) + 1;
        gz_device_free(gz_scratch.active_block_sums);
        gz_scratch.active_block_sums = (uint64_t*)gz_device_malloc((2 * blocks + 8) * sizeof(uint64_t));
        if (gz_scratch.active_totals == nullptr) gz_scratch.active_totals = (uint64_t*)gz_device_malloc(8 * sizeof(uint64_t));
        gz_scratch.scratch_vertices = num_vertices;
        gz_scratch.active_capacity = num_vertices;
    }
}

extern "C" void gz_scratch_reserve_public(uint64_t waves, uint64_t num_vertices) { gz_scratch_reserve(waves, num_vertices); }

extern "C" void gz_merge_buffer_allocate(uint64_t units)
{
    gz_require_runtime();
    if (units == 0) units = 1;
    if (units > gz_merge_buffer_count) {
        gz_device_free(gz_merge_buffer);
        gz_merge_buffer = (mergeaccum_t*)gz_device_malloc(units * sizeof(mergeaccum_t));
    }
    gz_merge_buffer_count = units;
    /* graph_data_allocate_merge_buffers initialises every record to ~0 / 0.0 */
    GZ_CHECK(hipMemsetAsync(gz_merge_buffer, 0xff, units * sizeof(mergeaccum_t), gz_stream));
}

/* ---- reductions ----------------------------------------------------------- */

/* Two-stage deterministic reduction: block b sums a fixed slice, the last
 * block to finish sums the block results in index order. */
template <typename T>
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_reduce(const T* __restrict__ in, uint64_t count, T* __restrict__ slots,
                                                        unsigned int* __restrict__ done, uint64_t* __restrict__ out)
{
    __shared__ T lds[GZ_WAVES_PER_BLOCK];
    __shared__ bool last;
    const uint64_t per = (count + gridDim.x - 1) / gridDim.x;
    const uint64_t b = (uint64_t)blockIdx.x * per;
    const uint64_t e = (b + per < count) ? b + per : count;
    T x = 0;
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) x += in[i];
    T r = gz_block_sum<T>(x, lds);
    if (threadIdx.x == 0) {
        slots[blockIdx.x] = r;
        __threadfence();
        const unsigned int prev = atomicAdd(done, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    T y = 0;
    for (uint64_t i = threadIdx.x; i < gridDim.x; i += blockDim.x) y += ((volatile T*)slots)[i];
    T total = gz_block_sum<T>(y, lds);
    if (threadIdx.x == 0) {
        T t = total;
        out[0] = *reinterpret_cast<uint64_t*>(&t);
        *done = 0u;
    }
}

/* Integer partials (votes) up to GZ_REDUCE_ONE_MAX: one workgroup sums them.
 * An integer sum does not depend on its order, so this gives gz_k_reduce's
 * result without its second stage, release fence and ticket, which cost
 * more than the loads at a few thousand partials. */
#define GZ_REDUCE_ONE_MAX 65536
#define GZ_REDUCE_ONE_THREADS 1024
__global__ __launch_bounds__(GZ_REDUCE_ONE_THREADS) void gz_k_reduce_one_u64(const unsigned long long* __restrict__ in,
                                                                             uint64_t count, uint64_t* __restrict__ out,
                                                                             uint64_t* __restrict__ mirror)
{
    constexpr uint64_t T = GZ_REDUCE_ONE_THREADS;
    __shared__ unsigned long long lds[GZ_REDUCE_ONE_THREADS / GZ_WAVE];
    unsigned long long x0 = 0, x1 = 0, x2 = 0, x3 = 0;
    uint64_t i = threadIdx.x;
    for (; i + 3 * T < count; i += 4 * T) {
        x0 += in[i];
        x1 += in[i + T];
        x2 += in[i + 2 * T];
        x3 += in[i + 3 * T];
    }
    for (; i < count; i += T) x0 += in[i];
    const unsigned long long r = gz_block_sum<unsigned long long>((x0 + x1) + (x2 + x3), lds);
    if (threadIdx.x == 0) {
        out[0] = r;
        gz_mirror_words(mirror, out);
    }
}

uint64_t gz_reduce_blocks(uint64_t count)
{
    uint64_t blocks = (count + 4095) / 4096;
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    return blocks;
}

void gz_reduce_f64_into(const double* in, uint64_t count, double* slots, unsigned int* done, uint64_t* out)
{
    gz_reduce_mirror_stale();
    hipLaunchKernelGGL(gz_k_reduce<double>, dim3((unsigned)gz_reduce_blocks(count)), dim3(GZ_BLOCK), 0, gz_stream,
                       in, count, slots, done, out);
    GZ_LAUNCH_CHECK();
}

void gz_reduce_partials(uint64_t* dst)
{
    const uint64_t count = gz_scratch.partials_count;
    const uint64_t blocks = gz_reduce_blocks(count);
    unsigned int* done = reinterpret_cast<unsigned int*>(gz_scratch.reduce_tmp + 2040);
    gz_ktimer_begin(GZ_K_REDUCE);
    if (gz_scratch.partials_is_double) {
        gz_reduce_f64_into((const double*)gz_scratch.partials, count, (double*)gz_scratch.reduce_tmp, done, dst);
    } else if (count <= GZ_REDUCE_ONE_MAX) {
        uint64_t* mirror = gz_reduce_mirror_target(dst);
        hipLaunchKernelGGL(gz_k_reduce_one_u64, dim3(1), dim3(GZ_REDUCE_ONE_THREADS), 0, gz_stream,
                           (const unsigned long long*)gz_scratch.partials, count, dst, mirror);
        GZ_LAUNCH_CHECK();
    } else {
        gz_reduce_mirror_stale();
        hipLaunchKernelGGL(gz_k_reduce<unsigned long long>, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream,
                           (const unsigned long long*)gz_scratch.partials, count, (unsigned long long*)gz_scratch.reduce_tmp,
                           done, dst);
        GZ_LAUNCH_CHECK();
    }
    gz_ktimer_end(GZ_K_REDUCE);
}

/* ---- common phase operators ----------------------------------------------- */

extern "C" void phase_op_reset_global_accum(void)
{
    gz_require_runtime();
    gz_scratch.partials_count = 0;
    gz_scratch.stats_count = 0;
}

extern "C" void phase_op_write_global_accum_to_buf(uint64_t* reduce_buffer)
{
    gz_require_runtime();
    gz_reduce_partials(reduce_buffer);
}

/* Pinned host copy of the whole 8-word reduce buffer: the vote is word 0,
 * and the learned selector leaves its decision and features in words 1-6,
 * so one small copy serves both.  The kernels that end a CC or BFS iteration
 * (the one-workgroup vote sum, the selector's decision and sums, the CC push
 * tail) write the eight words into it themselves: `gz_mirror_holds` names
 * the reduce buffer it then holds once the stream drains, and the drivers'
 * read (gz_combine_vote) needs no copy.  Every other writer of a reduce
 * buffer marks it stale. */
static uint64_t* gz_reduce_mirror = nullptr;     /* host address */
static uint64_t* gz_reduce_mirror_dev = nullptr; /* its device address */
static const uint64_t* gz_mirror_holds = nullptr;

static void gz_reduce_mirror_alloc(void)
{
    if (gz_reduce_mirror != nullptr) return;
    GZ_CHECK(hipHostMalloc((void**)&gz_reduce_mirror, 8 * sizeof(uint64_t), hipHostMallocMapped | hipHostMallocCoherent));
    GZ_CHECK(hipHostGetDevicePointer((void**)&gz_reduce_mirror_dev, gz_reduce_mirror, 0));
}

extern "C" uint64_t* gz_reduce_mirror_target(const uint64_t* reduce_buffer)
{
    gz_reduce_mirror_alloc();
    gz_mirror_holds = reduce_buffer;
    return gz_reduce_mirror_dev;
}

extern "C" void gz_reduce_mirror_stale(void) { gz_mirror_holds = nullptr; }

/* The phase API's read copies the buffer: a caller may have changed it
 * since the kernels that wrote the pinned copy. */
extern "C" uint64_t phase_op_combine_global_var_from_buf(uint64_t* reduce_buffer)
{
    gz_require_runtime();
    gz_reduce_mirror_alloc();
    GZ_CHECK(hipMemcpyAsync(gz_reduce_mirror, reduce_buffer, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, gz_stream));
    gz_mirror_holds = nullptr;
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    return gz_reduce_mirror[0];
}

/* The library's own drivers, which write the reduce buffer only through
 * the launches above: no copy when the pinned copy holds this buffer. */
extern "C" uint64_t gz_combine_vote(uint64_t* reduce_buffer)
{
    gz_require_runtime();
    gz_reduce_mirror_alloc();
    if (gz_mirror_holds != reduce_buffer)
        GZ_CHECK(hipMemcpyAsync(gz_reduce_mirror, reduce_buffer, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, gz_stream));
    gz_mirror_holds = nullptr;
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    return gz_reduce_mirror[0];
}

extern "C" const uint64_t* gz_reduce_words(void)
{
    return gz_reduce_mirror;
}

/* Parallel form of the serial merge in phases.c:41-70, one thread per record
 * (gz_merge_sum_record). */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_merge_sum(const mergeaccum_t* __restrict__ mb, uint64_t count,
                                                           double* __restrict__ acc)
{
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= count) return;
    gz_merge_sum_record(mb, count, u, acc);
}

void gz_launch_merge_sum(const mergeaccum_t* merge, uint64_t count, double* acc)
{
    if (count == 0) return;
    const uint64_t blocks = (count + GZ_BLOCK - 1) / GZ_BLOCK;
    hipLaunchKernelGGL(gz_k_merge_sum, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, merge, count, acc);
    GZ_LAUNCH_CHECK();
}

extern "C" void edge_pull_op_merge_with_merge_buffer(mergeaccum_t* merge_buffer, uint64_t count, double* vertex_accumulators)
{
    gz_require_runtime();
    if (count == 0) return;
    gz_ktimer_begin(GZ_K_MERGE);
    gz_launch_merge_sum(merge_buffer, count, vertex_accumulators);
    gz_ktimer_end(GZ_K_MERGE);
}

/* ---- one PageRank iteration that allocates nothing ----------------------------
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Its scratch (one merge record and one
 * partial sum per pull unit, the reduction's block slots, its done counter
 * and the sum S) lives in this file-scope array; every call on a device
 * shares that device's copy.  The host uses the address looked up per device
 * at init, never the array's name. */
#define GZ_NOALLOC_MAX_UNITS (1ull << 19)
#define GZ_NOALLOC_MERGE_WORDS (4 * GZ_NOALLOC_MAX_UNITS)
#define GZ_NOALLOC_PARTIALS_WORD GZ_NOALLOC_MERGE_WORDS
#define GZ_NOALLOC_SLOTS_WORD (GZ_NOALLOC_PARTIALS_WORD + GZ_NOALLOC_MAX_UNITS)
#define GZ_NOALLOC_DONE_WORD (GZ_NOALLOC_SLOTS_WORD + 1024)
#define GZ_NOALLOC_SUM_WORD (GZ_NOALLOC_DONE_WORD + 1)
#define GZ_NOALLOC_WORDS (GZ_NOALLOC_SUM_WORD + 63)
#define GZ_MAX_DEVICES 64
static __device__ __align__(256) uint64_t scratch[GZ_NOALLOC_WORDS];
static uint64_t* gz_noalloc_scratch[GZ_MAX_DEVICES];

/* With `device` current: its copy of the scratch array. */
static void gz_noalloc_setup(int device)
{
    if (device < 0 || device >= GZ_MAX_DEVICES || gz_noalloc_scratch[device] != nullptr) return;
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || p == nullptr) {
        fprintf(stderr, "grazelle: no address for the allocation-free scratch on device %d: %s\n", device,
                hipGetErrorString(err));
        exit(70);
    }
    gz_noalloc_scratch[device] = (uint64_t*)p;
}

extern "C" int gz_pagerank_iteration_noalloc(const gz_vec_t* pull_vectors, uint64_t pull_vector_count,
                                             uint64_t num_vertices, const double* outdegrees, double* props,
                                             double* accumulators)
{
    gz_require_runtime();
    const uint64_t units = gz_chunks_for(pull_vector_count);
    if (units > GZ_NOALLOC_MAX_UNITS) {
        fprintf(stderr,
                "grazelle: gz_pagerank_iteration_noalloc: %llu vectors make %llu pull units at %llu vectors per "
                "unit, above its limit of %llu units\n",
                (unsigned long long)pull_vector_count, (unsigned long long)units,
                (unsigned long long)gz_chunk_len(), (unsigned long long)GZ_NOALLOC_MAX_UNITS);
        return -1;
    }
    if (num_vertices == 0) return 0;
    uint64_t* base = (gz_device_id >= 0 && gz_device_id < GZ_MAX_DEVICES) ? gz_noalloc_scratch[gz_device_id] : nullptr;
    if (base == nullptr) {
        fprintf(stderr, "grazelle: gz_pagerank_iteration_noalloc: no scratch on device %d\n", gz_device_id);
        return -1;
    }
    mergeaccum_t* merge = (mergeaccum_t*)base;
    double* partials = (double*)(base + GZ_NOALLOC_PARTIALS_WORD);
    double* slots = (double*)(base + GZ_NOALLOC_SLOTS_WORD);
    unsigned int* done = (unsigned int*)(base + GZ_NOALLOC_DONE_WORD);
    uint64_t* sum = base + GZ_NOALLOC_SUM_WORD;
    /* the reduction's done counter must start at 0 (a fill kernel: no memset
     * node if the caller captures this); S is 0 for an empty list */
    gz_device_zero(done, 2 * sizeof(uint64_t));
    gz_launch_pr_pull_explicit(pull_vectors, pull_vector_count, gz_chunk_len(), units, props, accumulators, merge,
                               partials);
    gz_launch_merge_sum(merge, units, accumulators);
    if (units > 0) gz_reduce_f64_into(partials, units, slots, done, sum);
    gz_launch_pr_vertex_explicit(props, accumulators, outdegrees, num_vertices, sum, gz_phase_config.damping);
    return 0;
}

/* ---- statistics ----------------------------------------------------------- */

extern "C" void gz_phase_stats_read(uint64_t* vectors, uint64_t* edges)
{
    gz_require_runtime();
    *vectors = 0;
    *edges = 0;
    const uint64_t n = gz_scratch.stats_count;
    if (n == 0) return;
    uint64_t* h = (uint64_t*)malloc(n * 2 * sizeof(uint64_t));
    gz_copy_to_host(h, gz_scratch.stats, n * 2 * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) {
        *vectors += h[2 * i];
        *edges += h[2 * i + 1];
    }
    free(h);
}

extern "C" void gz_runtime_shutdown(void)
{
    if (gz_stream == nullptr) return;
    GZ_CHECK(hipStreamSynchronize(gz_stream));
    gz_replay_release();
    gz_device_free(gz_scratch.partials);
    gz_device_free(gz_scratch.stats);
    gz_device_free(gz_scratch.features);
    gz_device_free(gz_scratch.reduce_tmp);
    gz_device_free(gz_scratch.active_vertices);
    gz_device_free(gz_scratch.active_prefix);
    gz_device_free(gz_scratch.active_block_sums);
    gz_device_free(gz_scratch.active_totals);
    gz_device_free(gz_scratch.bfs_first_vector);
    gz_device_free(gz_scratch.push_pieces);
    gz_device_free(gz_scratch.push_queue);
    gz_device_free(gz_merge_buffer);
    gz_merge_buffer = nullptr;
    gz_merge_buffer_count = 0;
    memset(&gz_scratch, 0, sizeof(gz_scratch));
    if (gz_reduce_mirror != nullptr) GZ_CHECK(hipHostFree(gz_reduce_mirror));
    gz_reduce_mirror = nullptr;
    GZ_CHECK(hipStreamDestroy(gz_stream));
    gz_stream = nullptr;
    gz_device_id = -1;
}
