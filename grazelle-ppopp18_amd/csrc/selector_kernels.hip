/*
 * selector_kernels.hip -- forward, backward and Adam kernels of the learned
 * engine selector (8->32->32->2 FP32 MLP, ReLU hidden layers, linear output).
 * No counterpart in the reference.
 *
 * Parameter layout (GZ_SEL_PARAMS floats):
 *   W1[32][8]  b1[32]  W2[32][32]  b2[32]  W3[2][32]  b3[2]
 * Training minimises the mean over samples and outputs of the squared error
 * against measured log edge-phase times.  The gradient of one sample is
 * computed by one thread; a block folds its threads' gradients in LDS and
 * writes one partial per block, and the Adam kernel sums the block partials
 * in index order.
 */
#include <string.h>

#include "gz_internal.h"
#include "gz_host_internal.h"
#include "grazelle_selector.h"

#define O_W1 0
#define O_B1 (O_W1 + GZ_SEL_H * GZ_SEL_IN)
#define O_W2 (O_B1 + GZ_SEL_H)
#define O_B2 (O_W2 + GZ_SEL_H * GZ_SEL_H)
#define O_W3 (O_B2 + GZ_SEL_H)
#define O_B3 (O_W3 + GZ_SEL_OUT * GZ_SEL_H)

__device__ __forceinline__ void gz_sel_forward(const float* __restrict__ p, const float* x, float* h1, float* h2,
                                               float* y)
{
    for (int j = 0; j < GZ_SEL_H; ++j) {
        float a = p[O_B1 + j];
        for (int i = 0; i < GZ_SEL_IN; ++i) a += p[O_W1 + j * GZ_SEL_IN + i] * x[i];
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < GZ_SEL_H; ++j) {
        float a = p[O_B2 + j];
        for (int i = 0; i < GZ_SEL_H; ++i) a += p[O_W2 + j * GZ_SEL_H + i] * h1[i];
        h2[j] = a > 0.f ? a : 0.f;
    }
    for (int k = 0; k < GZ_SEL_OUT; ++k) {
        float a = p[O_B3 + k];
        for (int i = 0; i < GZ_SEL_H; ++i) a += p[O_W3 + k * GZ_SEL_H + i] * h2[i];
        y[k] = a;
    }
}

__global__ void gz_k_sel_forward(const float* __restrict__ p, const float* __restrict__ X, float* __restrict__ Y,
                                 uint64_t n)
{
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[GZ_SEL_IN], h1[GZ_SEL_H], h2[GZ_SEL_H], y[GZ_SEL_OUT];
    for (int i = 0; i < GZ_SEL_IN; ++i) x[i] = X[s * GZ_SEL_IN + i];
    gz_sel_forward(p, x, h1, h2, y);
    for (int k = 0; k < GZ_SEL_OUT; ++k) Y[s * GZ_SEL_OUT + k] = y[k];
}

/* Gradient of L = 1/(n*2) * sum_s sum_k (y_k - t_k)^2 w.r.t. every parameter. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_sel_grad(const float* __restrict__ p, const float* __restrict__ X,
                                                          const float* __restrict__ T, uint64_t n,
                                                          float* __restrict__ block_grads, float* __restrict__ block_loss)
{
    __shared__ float g[GZ_SEL_PARAMS];
    __shared__ float lsum[GZ_BLOCK];
    for (int i = threadIdx.x; i < GZ_SEL_PARAMS; i += blockDim.x) g[i] = 0.f;
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float loss = 0.f;
    if (s < n) {
        float x[GZ_SEL_IN], h1[GZ_SEL_H], h2[GZ_SEL_H], y[GZ_SEL_OUT];
        for (int i = 0; i < GZ_SEL_IN; ++i) x[i] = X[s * GZ_SEL_IN + i];
        gz_sel_forward(p, x, h1, h2, y);
        const float scale = 1.0f / (float)(n * GZ_SEL_OUT);
        float dy[GZ_SEL_OUT];
        for (int k = 0; k < GZ_SEL_OUT; ++k) {
            const float r = y[k] - T[s * GZ_SEL_OUT + k];
            loss += r * r;
            dy[k] = 2.f * r * scale;
        }
        float dh2[GZ_SEL_H];
        for (int i = 0; i < GZ_SEL_H; ++i) dh2[i] = 0.f;
        for (int k = 0; k < GZ_SEL_OUT; ++k) {
            atomicAdd(&g[O_B3 + k], dy[k]);
            for (int i = 0; i < GZ_SEL_H; ++i) {
                atomicAdd(&g[O_W3 + k * GZ_SEL_H + i], dy[k] * h2[i]);
                dh2[i] += dy[k] * p[O_W3 + k * GZ_SEL_H + i];
            }
        }
        float dh1[GZ_SEL_H];
        for (int i = 0; i < GZ_SEL_H; ++i) dh1[i] = 0.f;
        for (int j = 0; j < GZ_SEL_H; ++j) {
            const float da = h2[j] > 0.f ? dh2[j] : 0.f;
            if (da == 0.f) continue;
            atomicAdd(&g[O_B2 + j], da);
            for (int i = 0; i < GZ_SEL_H; ++i) {
                atomicAdd(&g[O_W2 + j * GZ_SEL_H + i], da * h1[i]);
                dh1[i] += da * p[O_W2 + j * GZ_SEL_H + i];
            }
        }
        for (int j = 0; j < GZ_SEL_H; ++j) {
            const float da = h1[j] > 0.f ? dh1[j] : 0.f;
            if (da == 0.f) continue;
            atomicAdd(&g[O_B1 + j], da);
            for (int i = 0; i < GZ_SEL_IN; ++i) atomicAdd(&g[O_W1 + j * GZ_SEL_IN + i], da * x[i]);
        }
    }
    lsum[threadIdx.x] = loss;
    __syncthreads();
    for (int i = threadIdx.x; i < GZ_SEL_PARAMS; i += blockDim.x) block_grads[(uint64_t)blockIdx.x * GZ_SEL_PARAMS + i] = g[i];
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int i = 0; i < (int)blockDim.x; ++i) t += lsum[i];
        block_loss[blockIdx.x] = t;
    }
}

__global__ void gz_k_sel_sum_grads(const float* __restrict__ block_grads, uint32_t blocks, float* __restrict__ grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= GZ_SEL_PARAMS) return;
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) s += block_grads[(uint64_t)b * GZ_SEL_PARAMS + i];
    grad[i] = s;
}

__global__ void gz_k_sel_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                              const float* __restrict__ grad, float lr, float beta1, float beta2, float eps,
                              float bc1, float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= GZ_SEL_PARAMS) return;
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

/* ---- per-iteration decision on the device ---------------------------------
 * The drivers read one block of words from the reduce buffer after every
 * edge phase (the converge vote).  The selector rides on that read: right
 * after the edge phase, gz_k_sel_stats folds the statistics of the next
 * frontier (the vertices the phase activated) into per-block partials, and
 * gz_k_sel_decide (one block) sums them in a fixed order, builds the feature
 * vector, runs the MLP and writes the decision and the features next to the
 * vote.  Choosing the engine costs two small launches and no extra host
 * synchronisation; there are no atomics, so the features are exact and the
 * same on every run. */

#define GZ_SEL_WORDS_PER_WAVE 64

struct gz_sel_dev_state {
    uint64_t unvisited;          /* BFS: vertices never in a frontier */
    uint64_t unvisited_in_edges; /* BFS: their in-degrees */
    uint64_t t_begin;            /* wall clock at the start of the last edge phase */
    uint64_t pad;
};

__global__ void gz_k_sel_stamp(gz_sel_dev_state* st)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) st->t_begin = (uint64_t)wall_clock64();
}

/* One wavefront per 64 bitmap words: lane l holds word l of the wave's range
 * for the population count, then the wave walks the words one by one with
 * lane l testing vertex 64w + l, so the degree reads of a dense frontier are
 * coalesced.  partials[b], partials[B + b], partials[2B + b] = (vertices,
 * out-edges, in-edges) of block b of the grid's B. */
__global__ __launch_bounds__(GZ_BLOCK) void gz_k_sel_stats(const uint64_t* __restrict__ frontier,
                                                           const double* __restrict__ outdeg,
                                                           const uint32_t* __restrict__ indeg, uint64_t num_vertices,
                                                           uint64_t words, uint64_t* __restrict__ partials)
{
    __shared__ uint64_t red[3][GZ_WAVES_PER_BLOCK];
    const int lane = gz_lane();
    const int wib = threadIdx.x / GZ_WAVE;
    const uint64_t w0 = ((uint64_t)blockIdx.x * GZ_WAVES_PER_BLOCK + wib) * GZ_SEL_WORDS_PER_WAVE;
    uint64_t mine = (w0 + lane < words) ? frontier[w0 + lane] : 0ull;
    if (w0 + lane == words - 1 && (num_vertices & 63)) mine &= (1ull << (num_vertices & 63)) - 1ull;
    uint64_t fv = (uint64_t)__popcll(mine), fe = 0, fi = 0;
    uint64_t any = __ballot(mine != 0);
    while (any) {
        const int k = __ffsll((unsigned long long)any) - 1;
        any &= any - 1;
        const uint64_t word = gz_shfl_u64(mine, k);
        if ((word >> lane) & 1ull) {
            const uint64_t v = (w0 + (uint64_t)k) * 64 + (uint64_t)lane;
            fe += (uint64_t)outdeg[v];
            fi += indeg ? (uint64_t)indeg[v] : 0ull;
        }
    }
    fv = gz_wave_sum_u64(fv);
    fe = gz_wave_sum_u64(fe);
    fi = gz_wave_sum_u64(fi);
    if (lane == 0) {
        red[0][wib] = fv;
        red[1][wib] = fe;
        red[2][wib] = fi;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t s = 0;
        for (int i = 0; i < GZ_WAVES_PER_BLOCK; ++i) s += red[threadIdx.x][i];
        partials[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

/* out[1] = decision (1 pull, 2 push, 0 no model), out[2..5] = the raw
 * features as 8 floats, out[6] = the two outputs as floats.  The unvisited
 * counts advance by the frontier (BFS marks each vertex once). */
/* partials: `blocks` slots of (vertices, out-edges, in-edges), component k
 * of slot c at partials[k * stride + c].
 * out_from_vote: 1 when the vote counts every activated vertex plus its
 * out-degree (stat mode both), so out-edges = vote - vertices; 2 when it
 * counts out-degrees only (out-edges = vote); 0: read the out-edge partials.
 * with_in: read the in-edge partials (the folds write 0 there without
 * in-degrees). */
__global__ __launch_bounds__(1024) void gz_k_sel_decide(const uint64_t* __restrict__ partials, uint64_t blocks,
                                                            uint64_t stride, const uint64_t* __restrict__ votes,
                                                            gz_sel_dev_state* st, const float* __restrict__ p,
                                                            const float* __restrict__ norm, int have_model, int bfs,
                                                            uint64_t num_vertices, uint64_t num_edges,
                                                            float iteration, float prev_engine, float prev_ms,
                                                            float clock_khz, uint64_t* __restrict__ out,
                                                            int out_from_vote, int with_in,
                                                            uint64_t* __restrict__ mirror)
{
    __shared__ uint64_t red[4][1024 / GZ_WAVE];
    __shared__ float x[GZ_SEL_IN], xn[GZ_SEL_IN], h1[GZ_SEL_H], h2[GZ_SEL_H], y[GZ_SEL_OUT];
    __shared__ float sp[GZ_SEL_PARAMS], snorm[2 * GZ_SEL_IN];
    const int j = threadIdx.x;
    const uint64_t t_end = (uint64_t)wall_clock64();
    /* the run state is read up front too (thread 0 updates it below) */
    uint64_t st_unvisited = 0, st_unvisited_in = 0, st_t_begin = 0;
    if (j == 0) {
        st_unvisited = st->unvisited;
        st_unvisited_in = st->unvisited_in_edges;
        st_t_begin = st->t_begin;
    }
    /* the weights and normalisation go to LDS first: their loads overlap the
     * partial sums instead of following them */
    if (have_model)
        for (int i = j; i < GZ_SEL_PARAMS; i += (int)blockDim.x) sp[i] = p[i];
    for (int i = j; i < 2 * GZ_SEL_IN; i += (int)blockDim.x) snorm[i] = norm[i];
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    /* four slots per step, so that sixteen loads per thread are in flight */
    const uint64_t T = blockDim.x;
    uint64_t b = j;
    for (; b + 3 * T < blocks; b += 4 * T) {
        uint64_t q[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t c = b + (uint64_t)u * T;
            q[u][0] = partials[c];
            q[u][1] = out_from_vote ? 0 : partials[stride + c];
            q[u][2] = with_in ? partials[2 * stride + c] : 0;
            q[u][3] = votes ? votes[c] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            s0 += q[u][0];
            s1 += q[u][1];
            s2 += q[u][2];
            s3 += q[u][3];
        }
    }
    for (; b < blocks; b += T) {
        s0 += partials[b];
        if (!out_from_vote) s1 += partials[stride + b];
        if (with_in) s2 += partials[2 * stride + b];
        if (votes) s3 += votes[b];
    }
    s0 = gz_wave_sum_u64(s0);
    s1 = gz_wave_sum_u64(s1);
    s2 = gz_wave_sum_u64(s2);
    s3 = gz_wave_sum_u64(s3);
    if (gz_lane() == 0) {
        red[0][j / GZ_WAVE] = s0;
        red[1][j / GZ_WAVE] = s1;
        red[2][j / GZ_WAVE] = s2;
        red[3][j / GZ_WAVE] = s3;
    }
    __syncthreads();
    if (j == 0) {
        uint64_t cv = 0, ce = 0, ci = 0, vote = 0;
        for (int i = 0; i < (int)blockDim.x / GZ_WAVE; ++i) {
            cv += red[0][i];
            ce += red[1][i];
            ci += red[2][i];
            vote += red[3][i];
        }
        if (votes) out[0] = vote;
        if (out_from_vote == 1) ce = vote - cv;
        if (out_from_vote == 2) ce = vote;
        uint64_t uv = num_vertices, ue = num_edges;
        if (bfs) {
            uv = st_unvisited - cv;
            ue = st_unvisited_in - ci;
            st->unvisited = uv;
            st->unvisited_in_edges = ue;
        }
        float ms = prev_ms;
        if (ms < 0.f) ms = (t_end > st_t_begin) ? (float)(t_end - st_t_begin) / clock_khz : 0.f;
        x[0] = logf(1.f + (float)cv);
        x[1] = logf(1.f + (float)ce);
        x[2] = logf(1.f + (float)uv);
        x[3] = logf(1.f + (float)ue);
        x[4] = iteration;
        x[5] = prev_engine;
        x[6] = logf(1.f + ms);
        x[7] = num_vertices ? (float)num_edges / (float)num_vertices : 0.f;
        for (int i = 0; i < GZ_SEL_IN; ++i) {
            /* a feature with std 0 is unused by the model: its input is 0 */
            const float sd = snorm[GZ_SEL_IN + i];
            xn[i] = sd > 0.f ? (x[i] - snorm[i]) / sd : 0.f;
        }
    }
    __syncthreads();
    if (have_model) {
        if (j < GZ_SEL_H) {
            float a = sp[O_B1 + j];
            for (int i = 0; i < GZ_SEL_IN; ++i) a += sp[O_W1 + j * GZ_SEL_IN + i] * xn[i];
            h1[j] = a > 0.f ? a : 0.f;
        }
        __syncthreads();
        if (j < GZ_SEL_H) {
            float a = sp[O_B2 + j];
            for (int i = 0; i < GZ_SEL_H; ++i) a += sp[O_W2 + j * GZ_SEL_H + i] * h1[i];
            h2[j] = a > 0.f ? a : 0.f;
        }
        __syncthreads();
        if (j < GZ_SEL_OUT) {
            float a = sp[O_B3 + j];
            for (int i = 0; i < GZ_SEL_H; ++i) a += sp[O_W3 + j * GZ_SEL_H + i] * h2[i];
            y[j] = a;
        }
        __syncthreads();
    }
    if (j == 0) {
        float* fo = (float*)&out[2];
        for (int i = 0; i < GZ_SEL_IN; ++i) fo[i] = x[i];
        float* yo = (float*)&out[6];
        yo[0] = have_model ? y[0] : 0.f;
        yo[1] = have_model ? y[1] : 0.f;
        out[1] = have_model ? ((y[0] < y[1]) ? 1u : 2u) : 0u;
        gz_mirror_words(mirror, out);
    }
}

/* The learned BFS engine's per-level sums when the host takes the decision:
 * the edge phase's per-wave vote and feature partials summed by one
 * workgroup into reduce words 0 (the vote), 2 (the activated vertices), 3
 * (their out-edges), 4 (their in-edges) and 5 (the edge phase's time on the
 * device clock, in ms, as a float), and copied to the pinned host copy.
 * out_from_vote and with_in as for gz_k_sel_decide. */
__global__ __launch_bounds__(1024) void gz_k_sel_sums(const uint64_t* __restrict__ partials, uint64_t blocks,
                                                          uint64_t stride, const uint64_t* __restrict__ votes,
                                                          const gz_sel_dev_state* __restrict__ st, float clock_khz,
                                                          int out_from_vote, int with_in, uint64_t* __restrict__ out,
                                                          uint64_t* __restrict__ mirror)
{
    __shared__ uint64_t red[4][1024 / GZ_WAVE];
    const int j = threadIdx.x;
    const uint64_t t_end = (uint64_t)wall_clock64();
    const uint64_t t_begin = j == 0 ? st->t_begin : 0;
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    const uint64_t T = blockDim.x;
    uint64_t b = j;
    for (; b + 3 * T < blocks; b += 4 * T) {
        uint64_t q[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t c = b + (uint64_t)u * T;
            q[u][0] = partials[c];
            q[u][1] = out_from_vote ? 0 : partials[stride + c];
            q[u][2] = with_in ? partials[2 * stride + c] : 0;
            q[u][3] = votes[c];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            s0 += q[u][0];
            s1 += q[u][1];
            s2 += q[u][2];
            s3 += q[u][3];
        }
    }
    for (; b < blocks; b += T) {
        s0 += partials[b];
        if (!out_from_vote) s1 += partials[stride + b];
        if (with_in) s2 += partials[2 * stride + b];
        s3 += votes[b];
    }
    s0 = gz_wave_sum_u64(s0);
    s1 = gz_wave_sum_u64(s1);
    s2 = gz_wave_sum_u64(s2);
    s3 = gz_wave_sum_u64(s3);
    if (gz_lane() == 0) {
        red[0][j / GZ_WAVE] = s0;
        red[1][j / GZ_WAVE] = s1;
        red[2][j / GZ_WAVE] = s2;
        red[3][j / GZ_WAVE] = s3;
    }
    __syncthreads();
    if (j == 0) {
        uint64_t cv = 0, ce = 0, ci = 0, vote = 0;
        for (int i = 0; i < (int)blockDim.x / GZ_WAVE; ++i) {
            cv += red[0][i];
            ce += red[1][i];
            ci += red[2][i];
            vote += red[3][i];
        }
        if (out_from_vote == 1) ce = vote - cv;
        if (out_from_vote == 2) ce = vote;
        const float ms = (t_end > t_begin) ? (float)(t_end - t_begin) / clock_khz : 0.f;
        const uint32_t msb = __float_as_uint(ms);
        out[0] = vote;
        out[2] = cv;
        out[3] = ce;
        out[4] = ci;
        out[5] = msb;
        gz_mirror_words(mirror, out);
    }
}

/* ---- host entry points ---------------------------------------------------- */

static float* gz_sel_dev_params = nullptr;
static float* gz_sel_dev_norm = nullptr;      /* mean[8], std[8] */
static gz_sel_dev_state* gz_sel_dev_st = nullptr;
static uint64_t* gz_sel_dev_partials = nullptr;
static uint64_t gz_sel_partials_blocks = 0;
static float gz_sel_clock_khz = 0.f;

extern "C" int gz_selector_device_upload_norm(const float* mean, const float* stdev)
{
    if (gz_sel_dev_norm == nullptr) gz_sel_dev_norm = (float*)gz_device_malloc(2 * GZ_SEL_IN * sizeof(float));
    gz_copy_to_device(gz_sel_dev_norm, mean, GZ_SEL_IN * sizeof(float));
    gz_copy_to_device(gz_sel_dev_norm + GZ_SEL_IN, stdev, GZ_SEL_IN * sizeof(float));
    return 0;
}

static uint64_t gz_sel_stats_blocks(uint64_t words)
{
    const uint64_t per_block = (uint64_t)GZ_WAVES_PER_BLOCK * GZ_SEL_WORDS_PER_WAVE;
    const uint64_t b = (words + per_block - 1) / per_block;
    return b ? b : 1;
}

__global__ void gz_k_sel_init(gz_sel_dev_state* st, uint64_t num_vertices, uint64_t num_edges, uint64_t root,
                              const double* __restrict__ outdeg, const uint32_t* __restrict__ indeg,
                              uint64_t* __restrict__ partials, int root_removed)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->unvisited = num_vertices;
        st->unvisited_in_edges = num_edges;
        st->t_begin = 0;
        if (root_removed) {
            st->unvisited = num_vertices - 1;
            st->unvisited_in_edges = num_edges - (indeg ? (uint64_t)indeg[root] : 0ull);
        } else if (root < num_vertices) {
            partials[0] = 1;
            partials[1] = (uint64_t)outdeg[root];
            partials[2] = indeg ? (uint64_t)indeg[root] : 0ull;
        }
    }
}

/* Start a run: unvisited = V / E (the first step removes the initial
 * frontier).  With root < V the initial frontier is that one vertex and its
 * statistics are staged for gz_sel_dev_step(..., GZ_SEL_FROM_ROOT); with
 * `root_removed` (root < V) the counts start past the root instead, for a
 * driver that takes the first decision on the host and runs no first step. */
extern "C" void gz_sel_dev_begin(uint64_t num_vertices, uint64_t num_edges, uint64_t root, const uint32_t* indeg,
                                 int root_removed)
{
    if (gz_sel_dev_st == nullptr) {
        gz_sel_dev_st = (gz_sel_dev_state*)gz_device_malloc(sizeof(gz_sel_dev_state));
        int rate = 0;
        int dev = 0;
        GZ_CHECK(hipGetDevice(&dev));
        GZ_CHECK(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev));
        gz_sel_clock_khz = (float)(rate > 0 ? rate : 100000);
    }
    if (gz_sel_dev_norm == nullptr) {
        const float zero[GZ_SEL_IN] = {0, 0, 0, 0, 0, 0, 0, 0};
        const float one[GZ_SEL_IN] = {1, 1, 1, 1, 1, 1, 1, 1};
        gz_selector_device_upload_norm(zero, one);
    }
    const uint64_t blocks = gz_sel_stats_blocks((num_vertices + 63) / 64);
    if (blocks > gz_sel_partials_blocks) {
        gz_device_free(gz_sel_dev_partials);
        gz_sel_dev_partials = (uint64_t*)gz_device_malloc((3 * blocks + 1) * sizeof(uint64_t));
        gz_sel_partials_blocks = blocks;
    }
    hipLaunchKernelGGL(gz_k_sel_init, dim3(1), dim3(64), 0, gz_stream, gz_sel_dev_st, num_vertices, num_edges, root,
                       gz_graph.vertex_outdegrees, indeg, gz_sel_dev_partials, root_removed && root < num_vertices);
    GZ_LAUNCH_CHECK();
}

/* Wall-clock stamp taken before an edge phase (learned engine).  With
 * `by_compaction` the edge phase's first kernel (the frontier compaction)
 * takes it, which saves a launch; otherwise a one-thread kernel does. */
extern "C" void gz_sel_dev_stamp(int by_compaction)
{
    if (by_compaction) {
        gz_scratch.stamp_target = &gz_sel_dev_st->t_begin;
        return;
    }
    hipLaunchKernelGGL(gz_k_sel_stamp, dim3(1), dim3(64), 0, gz_stream, gz_sel_dev_st);
    GZ_LAUNCH_CHECK();
}

/* Features and (with a model) the decision into reduce_buffer[1..6]; with
 * `with_vote` (edge-phase statistics only) also the converge vote into
 * reduce_buffer[0], summed from the edge phase's per-wave vote partials in
 * place of phase_op_write_global_accum_to_buf.  The
 * frontier statistics come from: GZ_SEL_FROM_EDGE_PHASE, the last BFS edge
 * phase's per-wave sums (gz_scratch.features, folded by the kernels);
 * GZ_SEL_FROM_ROOT, the root staged by gz_sel_dev_begin; GZ_SEL_FROM_BITMAP,
 * a pass over `frontier` (a bitmap over V vertices).  prev_ms < 0 takes the
 * previous edge phase's time from the device clock (stamp to the decision). */
extern "C" void gz_sel_dev_step(const uint64_t* frontier, const uint32_t* indeg, uint64_t* reduce_buffer,
                                int have_model, int bfs, uint64_t iteration, uint32_t prev_engine, float prev_ms,
                                int source, int with_vote)
{
    if (have_model && gz_sel_dev_params == nullptr) have_model = 0;
    const uint64_t V = gz_graph.num_vertices;
    const uint64_t* feat = gz_scratch.features;
    uint64_t blocks = gz_scratch.partials_count, stride = gz_scratch.partials_capacity;
    if (source == GZ_SEL_FROM_ROOT) {
        feat = gz_sel_dev_partials;
        blocks = 1;
        stride = 1;
    } else if (source == GZ_SEL_FROM_BITMAP) {
        const uint64_t words = (V + 63) / 64;
        blocks = gz_sel_stats_blocks(words);
        hipLaunchKernelGGL(gz_k_sel_stats, dim3((unsigned)blocks), dim3(GZ_BLOCK), 0, gz_stream, frontier,
                           gz_graph.vertex_outdegrees, indeg, V, words, gz_sel_dev_partials);
        GZ_LAUNCH_CHECK();
        feat = gz_sel_dev_partials;
        stride = blocks;
    }
    const uint64_t* votes = (with_vote && source == GZ_SEL_FROM_EDGE_PHASE && !gz_scratch.partials_is_double)
                                ? (const uint64_t*)gz_scratch.partials
                                : nullptr;
    /* the vote already sums the activated vertices' out-degrees in these
     * statistic modes, and the in-edge partials are 0 without in-degrees */
    int out_from_vote = 0;
    if (votes != nullptr && gz_phase_config.stat_mode == GZ_STAT_COUNT_AND_OUTDEG) out_from_vote = 1;
    if (votes != nullptr && gz_phase_config.stat_mode == GZ_STAT_OUTDEG_ONLY) out_from_vote = 2;
    const int with_in = indeg != nullptr || source != GZ_SEL_FROM_EDGE_PHASE;
    uint64_t* mirror = gz_reduce_mirror_target(reduce_buffer);
    /* timed as the vote reduction it replaces */
    gz_ktimer_begin(GZ_K_REDUCE);
    hipLaunchKernelGGL(gz_k_sel_decide, dim3(1), dim3(1024), 0, gz_stream, feat, blocks, stride, votes, gz_sel_dev_st,
                       gz_sel_dev_params, gz_sel_dev_norm, have_model, bfs, V, gz_graph.num_edges, (float)iteration,
                       (float)prev_engine, prev_ms, gz_sel_clock_khz, reduce_buffer, out_from_vote, with_in, mirror);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_REDUCE);
}

extern "C" void gz_sel_dev_sums(uint64_t* reduce_buffer, const uint32_t* indeg)
{
    int out_from_vote = 0;
    if (gz_phase_config.stat_mode == GZ_STAT_COUNT_AND_OUTDEG) out_from_vote = 1;
    if (gz_phase_config.stat_mode == GZ_STAT_OUTDEG_ONLY) out_from_vote = 2;
    uint64_t* mirror = gz_reduce_mirror_target(reduce_buffer);
    gz_ktimer_begin(GZ_K_REDUCE);
    hipLaunchKernelGGL(gz_k_sel_sums, dim3(1), dim3(1024), 0, gz_stream, gz_scratch.features, gz_scratch.partials_count,
                       gz_scratch.partials_capacity, (const uint64_t*)gz_scratch.partials, gz_sel_dev_st, gz_sel_clock_khz, out_from_vote,
                       indeg != nullptr ? 1 : 0, reduce_buffer, mirror);
    GZ_LAUNCH_CHECK();
    gz_ktimer_end(GZ_K_REDUCE);
}

extern "C" int gz_selector_device_upload(const float* params)
{
    if (gz_sel_dev_params == nullptr) gz_sel_dev_params = (float*)gz_device_malloc(GZ_SEL_PARAMS * sizeof(float));
    gz_copy_to_device(gz_sel_dev_params, params, GZ_SEL_PARAMS * sizeof(float));
    return 0;
}

/* Batched forward on the device: X[n][8] -> Y[n][2] (host arrays). */
extern "C" int gz_selector_forward_batch(const float* params, const float* X, float* Y, uint64_t n)
{
    float* dp = (float*)gz_device_malloc(GZ_SEL_PARAMS * sizeof(float));
    float* dx = (float*)gz_device_malloc((n ? n : 1) * GZ_SEL_IN * sizeof(float));
    float* dy = (float*)gz_device_malloc((n ? n : 1) * GZ_SEL_OUT * sizeof(float));
    gz_copy_to_device(dp, params, GZ_SEL_PARAMS * sizeof(float));
    gz_copy_to_device(dx, X, n * GZ_SEL_IN * sizeof(float));
    if (n) {
        hipLaunchKernelGGL(gz_k_sel_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, gz_stream, dp, dx, dy, n);
        GZ_LAUNCH_CHECK();
    }
    gz_copy_to_host(Y, dy, n * GZ_SEL_OUT * sizeof(float));
    gz_device_free(dp);
    gz_device_free(dx);
    gz_device_free(dy);
    return 0;
}

/* Full-batch gradient on the device (host arrays in and out); returns the
 * mean squared error. */
extern "C" float gz_selector_gradient(const float* params, const float* X, const float* T, uint64_t n, float* grad)
{
    const uint32_t blocks = (uint32_t)((n + GZ_BLOCK - 1) / GZ_BLOCK);
    float* dp = (float*)gz_device_malloc(GZ_SEL_PARAMS * sizeof(float));
    float* dx = (float*)gz_device_malloc((n ? n : 1) * GZ_SEL_IN * sizeof(float));
    float* dt = (float*)gz_device_malloc((n ? n : 1) * GZ_SEL_OUT * sizeof(float));
    float* bg = (float*)gz_device_malloc((blocks ? blocks : 1) * (uint64_t)GZ_SEL_PARAMS * sizeof(float));
    float* bl = (float*)gz_device_malloc((blocks ? blocks : 1) * sizeof(float));
    float* dg = (float*)gz_device_malloc(GZ_SEL_PARAMS * sizeof(float));
    gz_copy_to_device(dp, params, GZ_SEL_PARAMS * sizeof(float));
    gz_copy_to_device(dx, X, n * GZ_SEL_IN * sizeof(float));
    gz_copy_to_device(dt, T, n * GZ_SEL_OUT * sizeof(float));
    float loss = 0.f;
    if (blocks) {
        hipLaunchKernelGGL(gz_k_sel_grad, dim3(blocks), dim3(GZ_BLOCK), 0, gz_stream, dp, dx, dt, n, bg, bl);
        GZ_LAUNCH_CHECK();
        hipLaunchKernelGGL(gz_k_sel_sum_grads, dim3((GZ_SEL_PARAMS + 255) / 256), dim3(256), 0, gz_stream, bg, blocks, dg);
        GZ_LAUNCH_CHECK();
        float* hl = (float*)malloc(blocks * sizeof(float));
        gz_copy_to_host(hl, bl, blocks * sizeof(float));
        for (uint32_t b = 0; b < blocks; ++b) loss += hl[b];
        free(hl);
        loss /= (float)(n * GZ_SEL_OUT);
    }
    gz_copy_to_host(grad, dg, GZ_SEL_PARAMS * sizeof(float));
    gz_device_free(dp);
    gz_device_free(dx);
    gz_device_free(dt);
    gz_device_free(bg);
    gz_device_free(bl);
    gz_device_free(dg);
    return loss;
}

/* Train with full-batch Adam for `steps` steps, entirely on the device.
 * params is updated in place; loss_history (optional) receives the loss
 * before each step. */
extern "C" int gz_selector_train(float* params, const float* X, const float* T, uint64_t n, uint32_t steps, float lr,
                                 float* loss_history)
{
    const uint32_t blocks = (uint32_t)((n + GZ_BLOCK - 1) / GZ_BLOCK);
    if (blocks == 0) return -1;
    float* dp = (float*)gz_device_malloc(GZ_SEL_PARAMS * sizeof(float));
    float* dm = (float*)gz_device_malloc(GZ_SEL_PARAMS * sizeof(float));
    float* dv = (float*)gz_device_malloc(GZ_SEL_PARAMS * sizeof(float));
    float* dx = (float*)gz_device_malloc(n * GZ_SEL_IN * sizeof(float));
    float* dt = (float*)gz_device_malloc(n * GZ_SEL_OUT * sizeof(float));
    float* bg = (float*)gz_device_malloc(blocks * (uint64_t)GZ_SEL_PARAMS * sizeof(float));
    float* bl = (float*)gz_device_malloc(blocks * sizeof(float));
    float* dg = (float*)gz_device_malloc(GZ_SEL_PARAMS * sizeof(float));
    gz_copy_to_device(dp, params, GZ_SEL_PARAMS * sizeof(float));
    gz_copy_to_device(dx, X, n * GZ_SEL_IN * sizeof(float));
    gz_copy_to_device(dt, T, n * GZ_SEL_OUT * sizeof(float));
    float* hl = (float*)malloc(blocks * sizeof(float));
    const float b1 = 0.9f, b2 = 0.999f;
    float p1 = 1.f, p2 = 1.f;
    for (uint32_t s = 0; s < steps; ++s) {
        hipLaunchKernelGGL(gz_k_sel_grad, dim3(blocks), dim3(GZ_BLOCK), 0, gz_stream, dp, dx, dt, n, bg, bl);
        GZ_LAUNCH_CHECK();
        hipLaunchKernelGGL(gz_k_sel_sum_grads, dim3((GZ_SEL_PARAMS + 255) / 256), dim3(256), 0, gz_stream, bg, blocks, dg);
        GZ_LAUNCH_CHECK();
        p1 *= b1;
        p2 *= b2;
        hipLaunchKernelGGL(gz_k_sel_adam, dim3((GZ_SEL_PARAMS + 255) / 256), dim3(256), 0, gz_stream, dp, dm, dv, dg, lr, b1,
                           b2, 1e-8f, 1.f - p1, 1.f - p2);
        GZ_LAUNCH_CHECK();
        if (loss_history) {
            gz_copy_to_host(hl, bl, blocks * sizeof(float));
            float l = 0.f;
            for (uint32_t b = 0; b < blocks; ++b) l += hl[b];
            loss_history[s] = l / (float)(n * GZ_SEL_OUT);
        }
    }
    gz_copy_to_host(params, dp, GZ_SEL_PARAMS * sizeof(float));
    free(hl);
    gz_device_free(dp);
    gz_device_free(dm);
    gz_device_free(dv);
    gz_device_free(dx);
    gz_device_free(dt);
    gz_device_free(bg);
    gz_device_free(bl);
    gz_device_free(dg);
    return 0;
}
