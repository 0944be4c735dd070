// selector.hip -- matrix features and a small MLP (forward, backward, Adam)
// for the learned selectors.  Everything here runs on the device; the host
// only moves the (tiny) parameter, feature and score arrays.  The SpMV
// kernel selector's trained parameters are compiled into the code object.
//
// Features of a CSR (SPIRA_FEATURES entries, all O(1)-scaled):
//   0 log2(n_rows+1)/20   1 log2(nnz+1)/30   2 log2(mean_len+1)/10
//   3 min(cv of row lengths, 4)/4            4 log2(max_len+1)/16
//   5..12 fraction of rows with length 0, 1, 2-3, 4-7, 8-15, 16-31, 32-63, >=64
//   13 mean (last col - first col)/n_cols     14 mean bandwidth/n_cols
//   15 max bandwidth/n_cols                   16 log2(n_cols+1)/20
//   17 caller-defined (the batch selector puts log2(k)/5 there)
// MLP: n_in -> hidden (ReLU) -> hidden (ReLU) -> n_out, FP32, mean squared
// error over the outputs a mask marks valid, Adam (Kingma & Ba 2015).
#include "internal.hpp"

#include <cmath>
#include <random>

namespace spira_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kFeat = SPIRA_FEATURES;
constexpr int kStat = 16;  // accumulated statistics

constexpr unsigned kFeatBlocks = 512;  // partial-sum blocks of the feature pass (grid-stride over rows)

// stats: 0 sum len, 1 sum len^2, 2 max len, 3..10 bins, 11 sum spread, 12 sum band, 13 max band.
// Every statistic is a sum or max of integers below 2^53, so it is exact in
// FP64 whatever the order; each block writes its own partials (no atomics)
// and k_feature_fold adds them up.
__device__ __forceinline__ bool stat_is_max(int i) { return i == 2 || i == 13; }

__global__ __launch_bounds__(kThreads) void k_feature_stats(const uint64_t *__restrict__ off,
                                                             const uint32_t *__restrict__ col, uint64_t n_rows,
                                                             double *__restrict__ part)
{
    __shared__ double s[kThreads / 64][kStat];
    double loc[kStat] = {};
    for (uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x; r < n_rows;
         r += (uint64_t)gridDim.x * kThreads) {
        const uint64_t a = off[r], b = off[r + 1];
        const double len = (double)(b - a);
        loc[0] += len;
        loc[1] += len * len;
        loc[2] = fmax(loc[2], len);
        int bin = 0;
        if (b > a) {
            uint64_t L = b - a;
            bin = 1;
            while (L > 1 && bin < 7) {
                L >>= 1;
                ++bin;
            }
            const double first = (double)col[a], last = (double)col[b - 1], rr = (double)r;
            loc[11] += last - first;
            const double band = fmax(fabs(first - rr), fabs(last - rr));
            loc[12] += band;
            loc[13] = fmax(loc[13], band);
        }
        loc[3 + bin] += 1.0;
    }
    // wave then block reduction; max for entries 2 and 13
    for (int i = 0; i < kStat; ++i) {
        double v = loc[i];
        for (int o = 32; o > 0; o >>= 1) {
            const double t = __shfl_xor(v, o, 64);
            v = stat_is_max(i) ? fmax(v, t) : v + t;
        }
        if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < kStat) {
        const int i = threadIdx.x;
        double v = s[0][i];
        for (int w = 1; w < kThreads / 64; ++w) v = stat_is_max(i) ? fmax(v, s[w][i]) : v + s[w][i];
        part[(uint64_t)blockIdx.x * kStat + i] = v;
    }
}

// One block: the partials of nb blocks folded into the statistics, then the
// features (the list at the top of this file; slot 17 left 0) into feat, and
// the statistics into stats (kStat doubles) when it is not null.
constexpr int kFoldThreads = 1024;

__global__ __launch_bounds__(kFoldThreads) void k_feature_fold(const double *__restrict__ part, uint32_t nb,
                                                                uint64_t n_rows, uint64_t nnz, uint64_t n_cols,
                                                                float *__restrict__ feat, double *__restrict__ stats)
{
    __shared__ double s[kFoldThreads / 64][kStat];
    __shared__ double h[kStat];
    double loc[kStat] = {};
    for (uint32_t j = threadIdx.x; j < nb; j += kFoldThreads) {  // one block's partials: kStat contiguous words
        const double *pj = part + (uint64_t)j * kStat;
#pragma unroll
        for (int i = 0; i < kStat; ++i) loc[i] = stat_is_max(i) ? fmax(loc[i], pj[i]) : loc[i] + pj[i];
    }
#pragma unroll
    for (int i = 0; i < kStat; ++i) {
        double v = loc[i];
        for (int o = 32; o > 0; o >>= 1) {
            const double t = __shfl_xor(v, o, 64);
            v = stat_is_max(i) ? fmax(v, t) : v + t;
        }
        if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < kStat) {
        const int i = threadIdx.x;
        double v = s[0][i];
        for (int w = 1; w < kFoldThreads / 64; ++w) v = stat_is_max(i) ? fmax(v, s[w][i]) : v + s[w][i];
        h[i] = v;
        if (stats) stats[i] = v;
    }
    __syncthreads();
    // one thread per feature
    const int t = threadIdx.x;
    if (t >= kFeat) return;
    const double n = (double)(n_rows ? n_rows : 1), nc = (double)(n_cols ? n_cols : 1);
    const double mean = h[0] / n;
    double f = 0.0;
    if (t == 0) {
        f = log2((double)n_rows + 1) / 20.0;
    } else if (t == 1) {
        f = log2((double)nnz + 1) / 30.0;
    } else if (t == 2) {
        f = log2(mean + 1) / 10.0;
    } else if (t == 3) {
        const double var = h[1] / n - mean * mean;
        const double cv = mean > 0 ? sqrt(var > 0 ? var : 0.0) / mean : 0.0;
        f = fmin(cv, 4.0) / 4.0;
    } else if (t == 4) {
        f = log2(h[2] + 1) / 16.0;
    } else if (t <= 12) {
        f = h[3 + (t - 5)] / n;
    } else if (t <= 14) {
        const double nonempty = n - h[3] > 0 ? n - h[3] : 1.0;
        f = h[t == 13 ? 11 : 12] / nonempty / nc;
    } else if (t == 15) {
        f = h[13] / nc;
    } else if (t == 16) {
        f = log2((double)n_cols + 1) / 20.0;
    }
    feat[t] = (float)f;  // slot 17 stays 0
}

// ---- MLP kernels ----------------------------------------------------------
__global__ void k_dense_fwd(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ b,
                            float *__restrict__ Y, uint32_t B, uint32_t din, uint32_t dout, int relu)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * dout) return;
    const uint32_t s = t / dout, o = t % dout;
    float acc = b[o];
    for (uint32_t i = 0; i < din; ++i) acc = fmaf(W[o * din + i], X[s * din + i], acc);
    Y[t] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dY = 2 * M * (Y - T) / count, loss = sum M (Y - T)^2 / count
__global__ void k_mse_grad(const float *__restrict__ Y, const float *__restrict__ T, const float *__restrict__ M,
                           uint32_t n, float inv_count, float *__restrict__ dY, float *__restrict__ loss)
{
    __shared__ float s[kThreads];
    float l = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float m = M[i];
        const float d = m > 0.0f ? (Y[i] - T[i]) : 0.0f;
        dY[i] = 2.0f * d * inv_count;
        l += d * d;
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, s[0] * inv_count);
}

// dX[s,i] = sum_o dY[s,o] W[o,i], times 1[H[s,i] > 0] when the input came from a ReLU
__global__ void k_dense_bwd_x(const float *__restrict__ dY, const float *__restrict__ W, const float *__restrict__ H,
                              float *__restrict__ dX, uint32_t B, uint32_t din, uint32_t dout, int relu_in)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * din) return;
    const uint32_t s = t / din, i = t % din;
    float acc = 0.0f;
    for (uint32_t o = 0; o < dout; ++o) acc = fmaf(dY[s * dout + o], W[o * din + i], acc);
    dX[t] = (relu_in && H[t] <= 0.0f) ? 0.0f : acc;
}

// dW[o,i] = sum_s dY[s,o] X[s,i]; db[o] = sum_s dY[s,o]   (thread per (o, i), i == din -> bias)
__global__ void k_dense_bwd_w(const float *__restrict__ dY, const float *__restrict__ X, float *__restrict__ dW,
                              float *__restrict__ db, uint32_t B, uint32_t din, uint32_t dout)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dout * (din + 1)) return;
    const uint32_t o = t / (din + 1), i = t % (din + 1);
    float acc = 0.0f;
    if (i == din) {
        for (uint32_t s = 0; s < B; ++s) acc += dY[s * dout + o];
        db[o] = acc;
    } else {
        for (uint32_t s = 0; s < B; ++s) acc = fmaf(dY[s * dout + o], X[s * din + i], acc);
        dW[o * din + i] = acc;
    }
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint32_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

}  // namespace spira_amd

using namespace spira_amd;

struct spira_mlp {
    spira_dev *dev = nullptr;
    uint32_t n_in = 0, hidden = 0, n_out = 0;
    uint32_t dims[4] = {};
    uint64_t n_params = 0;
    size_t woff[3] = {}, boff[3] = {};  // offsets of W_l and b_l in the flat arrays
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    float *act = nullptr;  // activations of the last forward: X copy, H1, H2, Y
    uint32_t act_batch = 0;
    float *dact = nullptr;  // gradients of the activations
    float *loss = nullptr;
    uint64_t step = 0;
    std::vector<std::pair<void *, size_t>> owned;
    void *alloc(size_t bytes)
    {
        void *q = dev_alloc(dev, bytes);
        owned.push_back({q, bytes ? bytes : 8});
        return q;
    }
};

namespace {
thread_local std::string g_sel_error;

template <class F>
int sguard(spira_dev *dev, F &&f)
{
    std::string *sink = dev ? &dev->last_error : &g_sel_error;
    try {
        if (dev) SPIRA_HIP(hipSetDevice(dev->device));
        f();
        return SPIRA_OK;
    } catch (const status_error &e) {
        *sink = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        *sink = "host allocation failed";
        return SPIRA_ERR_NOMEM;
    } catch (const std::exception &e) {
        *sink = e.what();
        return SPIRA_ERR_DEVICE;
    }
}

// activation buffers for batch B: [X | H1 | H2 | Y] each B x dim
void ensure_act(spira_mlp *N, uint32_t B)
{
    if (N->act && N->act_batch >= B) return;
    if (N->act) {
        SPIRA_HIP(hipStreamSynchronize(N->dev->stream));
        for (auto it = N->owned.begin(); it != N->owned.end();) {
            if (it->first == N->act || it->first == N->dact) {
                dev_release(N->dev, it->first, it->second);
                it = N->owned.erase(it);
            } else {
                ++it;
            }
        }
    }
    const size_t per = (size_t)N->dims[0] + N->dims[1] + N->dims[2] + N->dims[3];
    N->act = (float *)N->alloc(per * B * sizeof(float));
    N->dact = (float *)N->alloc(per * B * sizeof(float));
    N->act_batch = B;
}

float *act_ptr(spira_mlp *N, float *base, int layer, uint32_t B)
{
    size_t o = 0;
    for (int l = 0; l < layer; ++l) o += (size_t)N->dims[l] * B;
    return base + o;
}

void forward(spira_mlp *N, const float *X, uint32_t B)
{
    hipStream_t st = N->dev->stream;
    ensure_act(N, B);
    SPIRA_HIP(hipMemcpyAsync(act_ptr(N, N->act, 0, B), X, (size_t)B * N->dims[0] * sizeof(float),
                             hipMemcpyDeviceToDevice, st));
    for (int l = 0; l < 3; ++l) {
        const uint32_t din = N->dims[l], dout = N->dims[l + 1];
        k_dense_fwd<<<grid_for((uint64_t)B * dout, kThreads), kThreads, 0, st>>>(
            act_ptr(N, N->act, l, B), N->p + N->woff[l], N->p + N->boff[l], act_ptr(N, N->act, l + 1, B), B, din,
            dout, l < 2 ? 1 : 0);
    }
    SPIRA_HIP(hipGetLastError());
}
}  // namespace

namespace spira_amd {

namespace {
#include "spmv_selector_weights.inc"

// The selector network's forward pass in one block: h1 = relu(W1 f + b1),
// h2 = relu(W2 h1 + b2), out = W3 h2 + b3, with the parameters read in place
// (kSpmvSelParams, compiled into the code object).  Each output is summed in
// the order k_dense_fwd uses (bias first, then fmaf over the inputs in order),
// so the scores equal the network's own forward pass bit for bit.
__global__ __launch_bounds__(kThreads) void k_sel_forward(const float *__restrict__ feat, float *__restrict__ out)
{
    __shared__ float s_a[kSpmvSelIn > kSpmvSelHidden ? kSpmvSelIn : kSpmvSelHidden];
    __shared__ float s_b[kSpmvSelHidden];
    const uint32_t t = threadIdx.x;
    if (t < kSpmvSelIn) s_a[t] = feat[t];
    __syncthreads();
    const float *P = kSpmvSelParams;
    if (t < kSpmvSelHidden) {  // layer 1
        const float *W = P, *b = P + kSpmvSelIn * kSpmvSelHidden;
        float acc = b[t];
        for (uint32_t i = 0; i < kSpmvSelIn; ++i) acc = fmaf(W[t * kSpmvSelIn + i], s_a[i], acc);
        s_b[t] = fmaxf(acc, 0.0f);
    }
    __syncthreads();
    P += kSpmvSelIn * kSpmvSelHidden + kSpmvSelHidden;
    if (t < kSpmvSelHidden) {  // layer 2
        const float *W = P, *b = P + kSpmvSelHidden * kSpmvSelHidden;
        float acc = b[t];
        for (uint32_t i = 0; i < kSpmvSelHidden; ++i) acc = fmaf(W[t * kSpmvSelHidden + i], s_b[i], acc);
        s_a[t] = fmaxf(acc, 0.0f);
    }
    __syncthreads();
    P += kSpmvSelHidden * kSpmvSelHidden + kSpmvSelHidden;
    if (t < kSpmvSelOut) {  // layer 3
        const float *W = P, *b = P + kSpmvSelHidden * kSpmvSelOut;
        float acc = b[t];
        for (uint32_t i = 0; i < kSpmvSelHidden; ++i) acc = fmaf(W[t * kSpmvSelHidden + i], s_a[i], acc);
        out[t] = acc;
    }
}

static_assert(kSpmvSelNParams == kSpmvSelIn * kSpmvSelHidden + kSpmvSelHidden + kSpmvSelHidden * kSpmvSelHidden +
                                     kSpmvSelHidden + kSpmvSelHidden * kSpmvSelOut + kSpmvSelOut,
              "selector parameter count");

// Queue the feature pass of A on its stream: kFeat floats at feat (device),
// and the kStat statistics at stats when it is not null.
void queue_features(spira_dcsr *A, float *feat, double *stats = nullptr)
{
    spira_dev *dev = A->dev;
    const uint32_t nb = A->n_rows ? (uint32_t)grid_for(A->n_rows, kThreads, kFeatBlocks) : 0;
    double *part = (double *)dev->ws.get(WS_MISC, ((size_t)nb + 1) * kStat * sizeof(double));
    if (nb) {
        k_feature_stats<<<nb, kThreads, 0, dev->stream>>>(A->off, A->col, A->n_rows, part);
        SPIRA_HIP(hipGetLastError());
    }
    k_feature_fold<<<1, kFoldThreads, 0, dev->stream>>>(part, nb, A->n_rows, A->nnz, A->n_cols, feat, stats);
    SPIRA_HIP(hipGetLastError());
}

}  // namespace

// The features of a locked CSR (the list at the top of this file), computed
// on the device; slot 17 is left 0.
void matrix_features_host(spira_dcsr *A, float *features)
{
    spira_dev *dev = A->dev;
    float *feat = (float *)dev->ws.get(WS_SEL, kFeat * sizeof(float));
    queue_features(A, feat);
    read_back(dev, features, feat, kFeat * sizeof(float));
}

// The trained network's outputs (predicted log2 slowdown of every variant
// against the best) for A's features, kept in A's plan; AUTO takes the lowest
// among the variants A fits.  Features and forward pass run on the device,
// back to back, and only the scores come back.  false when no trained model
// was compiled in.
// Row-length statistics (longest row, sum of squared lengths) and the
// selector's scores of A from one feature pass and one read-back: what the
// first AUTO call needs.  false (nothing done) when no trained model was
// compiled in.
bool spmv_stats_and_scores(spira_dcsr *A, uint64_t *max_len, double *sumsq)
{
    if (!kSpmvSelTrained || A->n_rows == 0) return false;
    spira_dev *dev = A->dev;
    struct words {
        double stats[kStat];
        float feat[kFeat];
        float score[kSpmvSelOut];
    };
    words *w = (words *)dev->ws.get(WS_SEL, sizeof(words));
    queue_features(A, w->feat, w->stats);
    k_sel_forward<<<1, kThreads, 0, dev->stream>>>(w->feat, w->score);
    SPIRA_HIP(hipGetLastError());
    words h;
    read_back(dev, &h, w, sizeof(h));
    *max_len = (uint64_t)h.stats[2];
    *sumsq = h.stats[1];
    std::copy(h.score, h.score + kSpmvSelOut, A->plan.sel_score);
    A->plan.scored = true;
    return true;
}

bool spmv_selector_scores(spira_dcsr *A)
{
    if (!kSpmvSelTrained || A->n_rows == 0) return false;
    if (A->plan.scored) return true;
    static_assert(kSpmvSelIn == kFeat && kSpmvSelOut == SPIRA_SPMV_N_VARIANTS, "selector model shape");
    spira_dev *dev = A->dev;
    float *feat = (float *)dev->ws.get(WS_SEL, (kFeat + kSpmvSelOut) * sizeof(float));
    queue_features(A, feat);
    k_sel_forward<<<1, kThreads, 0, dev->stream>>>(feat, feat + kFeat);
    SPIRA_HIP(hipGetLastError());
    read_back(dev, A->plan.sel_score, feat + kFeat, kSpmvSelOut * sizeof(float));
    A->plan.scored = true;
    return true;
}

}  // namespace spira_amd

extern "C" {

int spira_matrix_features(spira_dcsr *A, float *features)
{
    if (!A || !features) return SPIRA_ERR_INVALID_ARGUMENT;
    return sguard(A->dev, [&] {
        if (!A->built) fail(SPIRA_ERR_LOGIC, "features: matrix must be locked");
        matrix_features_host(A, features);
    });
}

int spira_mlp_create(spira_dev *dev, uint32_t n_in, uint32_t hidden, uint32_t n_out, uint64_t seed, spira_mlp **out)
{
    if (!dev || !out || !n_in || !hidden || !n_out) return SPIRA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    spira_mlp *N = new spira_mlp();
    N->dev = dev;
    int rc = sguard(dev, [&] {
        N->n_in = n_in;
        N->hidden = hidden;
        N->n_out = n_out;
        N->dims[0] = n_in;
        N->dims[1] = hidden;
        N->dims[2] = hidden;
        N->dims[3] = n_out;
        size_t o = 0;
        for (int l = 0; l < 3; ++l) {
            N->woff[l] = o;
            o += (size_t)N->dims[l + 1] * N->dims[l];
            N->boff[l] = o;
            o += N->dims[l + 1];
        }
        N->n_params = o;
        N->p = (float *)N->alloc(o * sizeof(float));
        N->g = (float *)N->alloc(o * sizeof(float));
        N->m = (float *)N->alloc(o * sizeof(float));
        N->v = (float *)N->alloc(o * sizeof(float));
        N->loss = (float *)N->alloc(64);
        // He-uniform weights, zero biases (seeded, host side)
        std::vector<float> init(o, 0.0f);
        std::mt19937_64 rng(seed);
        for (int l = 0; l < 3; ++l) {
            const float lim = std::sqrt(6.0f / (float)N->dims[l]);
            std::uniform_real_distribution<float> U(-lim, lim);
            for (size_t i = 0; i < (size_t)N->dims[l + 1] * N->dims[l]; ++i) init[N->woff[l] + i] = U(rng);
        }
        SPIRA_HIP(hipMemcpyAsync(N->p, init.data(), o * sizeof(float), hipMemcpyHostToDevice, dev->stream));
        SPIRA_HIP(hipMemsetAsync(N->m, 0, o * sizeof(float), dev->stream));
        SPIRA_HIP(hipMemsetAsync(N->v, 0, o * sizeof(float), dev->stream));
        SPIRA_HIP(hipMemsetAsync(N->g, 0, o * sizeof(float), dev->stream));
        dev_sync(dev);
    });
    if (rc != SPIRA_OK) {
        for (auto &q : N->owned) dev_release(dev, q.first, q.second);
        delete N;
        return rc;
    }
    *out = N;
    return SPIRA_OK;
}

void spira_mlp_destroy(spira_mlp *N)
{
    if (!N) return;
    (void)hipSetDevice(N->dev->device);
    (void)hipStreamSynchronize(N->dev->stream);
    for (auto &q : N->owned) dev_release(N->dev, q.first, q.second);
    delete N;
}

int spira_mlp_params(spira_mlp *N, float *get, const float *set, uint64_t *n_params)
{
    if (!N) return SPIRA_ERR_INVALID_ARGUMENT;
    return sguard(N->dev, [&] {
        if (n_params) *n_params = N->n_params;
        if (set) {
            SPIRA_HIP(hipMemcpyAsync(N->p, set, N->n_params * sizeof(float), hipMemcpyHostToDevice, N->dev->stream));
            N->step = 0;
            SPIRA_HIP(hipMemsetAsync(N->m, 0, N->n_params * sizeof(float), N->dev->stream));
            SPIRA_HIP(hipMemsetAsync(N->v, 0, N->n_params * sizeof(float), N->dev->stream));
        }
        if (get)
            SPIRA_HIP(hipMemcpyAsync(get, N->p, N->n_params * sizeof(float), hipMemcpyDeviceToHost, N->dev->stream));
        dev_sync(N->dev);
    });
}

int spira_mlp_forward(spira_mlp *N, const float *X, uint32_t batch, float *Y)
{
    if (!N || (batch && (!X || !Y))) return SPIRA_ERR_INVALID_ARGUMENT;
    return sguard(N->dev, [&] {
        if (!batch) return;
        forward(N, X, batch);
        SPIRA_HIP(hipMemcpyAsync(Y, act_ptr(N, N->act, 3, batch), (size_t)batch * N->n_out * sizeof(float),
                                 hipMemcpyDeviceToDevice, N->dev->stream));
        dev_sync(N->dev);
    });
}

int spira_mlp_loss_grad(spira_mlp *N, const float *X, const float *T, const float *M, uint32_t batch,
                        float mask_count, float *loss, float *grad)
{
    if (!N || !batch || !X || !T || !M) return SPIRA_ERR_INVALID_ARGUMENT;
    return sguard(N->dev, [&] {
        if (!(mask_count > 0.0f)) fail(SPIRA_ERR_INVALID_ARGUMENT, "mlp: mask_count must be > 0");
        hipStream_t st = N->dev->stream;
        forward(N, X, batch);
        const uint32_t B = batch;
        SPIRA_HIP(hipMemsetAsync(N->loss, 0, sizeof(float), st));
        const uint32_t no = B * N->n_out;
        k_mse_grad<<<grid_for(no, kThreads, 64), kThreads, 0, st>>>(act_ptr(N, N->act, 3, B), T, M, no,
                                                                     1.0f / mask_count, act_ptr(N, N->dact, 3, B),
                                                                     N->loss);
        for (int l = 2; l >= 0; --l) {
            const uint32_t din = N->dims[l], dout = N->dims[l + 1];
            k_dense_bwd_w<<<grid_for((uint64_t)dout * (din + 1), kThreads), kThreads, 0, st>>>(
                act_ptr(N, N->dact, l + 1, B), act_ptr(N, N->act, l, B), N->g + N->woff[l], N->g + N->boff[l], B, din,
                dout);
            if (l > 0)
                k_dense_bwd_x<<<grid_for((uint64_t)B * din, kThreads), kThreads, 0, st>>>(
                    act_ptr(N, N->dact, l + 1, B), N->p + N->woff[l], act_ptr(N, N->act, l, B),
                    act_ptr(N, N->dact, l, B), B, din, dout, 1);
        }
        SPIRA_HIP(hipGetLastError());
        if (loss) SPIRA_HIP(hipMemcpyAsync(loss, N->loss, sizeof(float), hipMemcpyDeviceToHost, st));
        if (grad) SPIRA_HIP(hipMemcpyAsync(grad, N->g, N->n_params * sizeof(float), hipMemcpyDeviceToHost, st));
        dev_sync(N->dev);
    });
}

int spira_mlp_adam_step(spira_mlp *N, float lr, float beta1, float beta2, float eps)
{
    if (!N) return SPIRA_ERR_INVALID_ARGUMENT;
    return sguard(N->dev, [&] {
        N->step += 1;
        const float c1 = 1.0f - std::pow(beta1, (float)N->step), c2 = 1.0f - std::pow(beta2, (float)N->step);
        k_adam<<<grid_for(N->n_params, kThreads), kThreads, 0, N->dev->stream>>>(N->p, N->g, N->m, N->v,
                                                                                 (uint32_t)N->n_params, lr, beta1,
                                                                                 beta2, eps, c1, c2);
        SPIRA_HIP(hipGetLastError());
    });
}

}  // extern "C"
