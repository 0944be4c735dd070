// batch.hip -- many matrices of uneven size, k right-hand sides (k = 1, 4, 16
// or 32), one persistent launch fed by a device work queue; the repeated call
// can be captured in a hipGraph.
//
// A batch is stored as one block-diagonal CSR: rows and columns of matrix m
// are shifted by row_base[m] / col_base[m], so X (sum of n_cols x k, row
// major) and Y (sum of n_rows x k, row major) are single arrays and the kernel
// needs no per-matrix metadata.  The rows are cut into tiles of bounded work
// (nnz * k); every tile carries the lanes-per-row of its matrix's schedule.
// Tiles are queued heaviest first; each block of the persistent grid takes
// the next tile with one atomic add until the queue is empty (every block
// leaves once the counter passes the tile count).
//
// Values are FP32 / BF16 / split BF16, X and Y are FP32, and each output is
// accumulated with the compensated Dot2 scheme (Ogita, Rump, Oishi 2005).
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace spira_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ float bf16f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

struct tile_t {
    uint32_t r0, r1;   // rows [r0, r1) of the block-diagonal CSR
    uint32_t elog2;    // lanes per row = 1 << elog2 (per right-hand side)
    uint32_t pad;
};

struct bvals {
    int storage;
    const float *f32;
    const uint16_t *hi, *lo;
    __device__ __forceinline__ float operator()(uint64_t i) const
    {
        if (storage == SPIRA_VAL_F32) return __builtin_nontemporal_load(f32 + i);
        if (storage == SPIRA_VAL_BF16) return bf16f(__builtin_nontemporal_load(hi + i));
        return bf16f(__builtin_nontemporal_load(hi + i)) + bf16f(__builtin_nontemporal_load(lo + i));
    }
};

struct acc2 {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float a, float x)
    {
        const float p = a * x;
        const float e = fmaf(a, x, -p);
        const float t = s + p;
        const float bp = t - s;
        c += ((s - (t - bp)) + (p - bp)) + e;
        s = t;
    }
    __device__ __forceinline__ void merge(float s2, float c2)
    {
        const float t = s + s2;
        const float bp = t - s;
        c += c2 + ((s - (t - bp)) + (s2 - bp));
        s = t;
    }
};

__global__ void k_reset_counter(uint32_t *counter) { *counter = 0u; }

template <int K>
__global__ __launch_bounds__(kThreads) void k_batch_spmm(const tile_t *__restrict__ tiles, uint32_t n_tiles,
                                                          uint32_t *__restrict__ counter,
                                                          const uint32_t *__restrict__ off,
                                                          const uint32_t *__restrict__ col, bvals vals,
                                                          const float *__restrict__ X, float *__restrict__ Y)
{
    __shared__ uint32_t s_t;
    __shared__ float s_part[2][kWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = lane % K;  // right-hand side of this lane
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t t = s_t;
        __syncthreads();  // every lane has read s_t before it is overwritten
        if (t >= n_tiles) break;  // block-uniform: the whole block leaves
        const tile_t tl = tiles[t];
        if (tl.r1 - tl.r0 == 1 && off[tl.r0 + 1] - off[tl.r0] > 256u) {
            // one long row: all four waves stride over it, then combine
            const uint32_t r = tl.r0;
            constexpr int slots = 64 / K;
            const int e = lane / K;
            acc2 a;
            for (uint32_t i = off[r] + w * slots + e; i < off[r + 1]; i += kWaves * slots)
                a.add(vals(i), X[(uint64_t)col[i] * K + j]);
            for (int o = K; o < 64; o <<= 1) a.merge(__shfl_xor(a.s, o, 64), __shfl_xor(a.c, o, 64));
            if (e == 0) {
                s_part[0][w][j] = a.s;
                s_part[1][w][j] = a.c;
            }
            __syncthreads();
            if (w == 0 && e == 0) {
                acc2 b;
                for (int ww = 0; ww < kWaves; ++ww) b.merge(s_part[0][ww][j], s_part[1][ww][j]);
                Y[(uint64_t)r * K + j] = b.s + b.c;
            }
            __syncthreads();
            continue;
        }
        // lanes = (row slot, entry slot e, rhs j): E entry lanes per row,
        // 64 / (E * K) rows per wavefront step
        const int E = 1 << tl.elog2;
        const int e = (lane / K) & (E - 1);
        const int rs = lane / (K * E);
        const int rpw = 64 / (K * E);
        // the lanes of one row (same rs) share the row and so the trip count:
        // the cross-lane merge only pairs lanes of the same row
        for (uint32_t r = tl.r0 + w * rpw + rs; r < tl.r1; r += kWaves * rpw) {
            acc2 a;
            for (uint32_t i = off[r] + e; i < off[r + 1]; i += E) a.add(vals(i), X[(uint64_t)col[i] * K + j]);
            for (int o = K; o < K * E; o <<= 1) a.merge(__shfl_xor(a.s, o, 64), __shfl_xor(a.c, o, 64));
            if (e == 0) Y[(uint64_t)r * K + j] = a.s + a.c;
        }
    }
}

// Concatenation of one matrix into the batch (grid-stride).
__global__ void k_cat_off(const uint64_t *__restrict__ off, uint64_t n, uint64_t nnz_base, uint64_t row_base,
                          uint64_t *__restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x)
        out[row_base + i] = off[i] + nnz_base;
}

__global__ void k_cat_entries(const uint32_t *__restrict__ col, const double *__restrict__ val, uint64_t nnz,
                              uint64_t nnz_base, uint32_t col_base, int storage, uint32_t *__restrict__ ocol,
                              float *__restrict__ of32, uint16_t *__restrict__ ohi, uint16_t *__restrict__ olo)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
         i += (uint64_t)gridDim.x * blockDim.x) {
        ocol[nnz_base + i] = col[i] + col_base;
        const float f = (float)val[i];
        if (storage == SPIRA_VAL_F32) {
            of32[nnz_base + i] = f;
        } else {
            uint32_t u = __float_as_uint(f);
            uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
            const uint16_t h = (u & 0x7fffffffu) > 0x7f800000u ? (uint16_t)((u >> 16) | 0x40u) : (uint16_t)(r >> 16);
            ohi[nnz_base + i] = h;
            if (storage == SPIRA_VAL_SPLIT_BF16) {
                const float rest = (float)(val[i] - (double)bf16f(h));
                uint32_t v = __float_as_uint(rest);
                uint32_t rv = v + 0x7fffu + ((v >> 16) & 1u);
                olo[nnz_base + i] = (v & 0x7fffffffu) > 0x7f800000u ? (uint16_t)((v >> 16) | 0x40u) : (uint16_t)(rv >> 16);
            }
        }
    }
}

__global__ void k_u64_to_u32_b(const uint64_t *__restrict__ a, uint32_t *__restrict__ b, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (uint32_t)a[i];
}

}  // namespace

}  // namespace spira_amd

using namespace spira_amd;

struct spira_batch {
    spira_dev *dev = nullptr;
    uint32_t n_mats = 0;
    int storage = SPIRA_VAL_F32;
    uint64_t n_rows = 0, n_cols = 0, nnz = 0;
    std::vector<uint64_t> row_base, col_base, nnz_base;  // n_mats+1 each
    std::vector<uint64_t> h_off;                          // host copy of the offsets (tile building)
    uint32_t *off32 = nullptr;
    uint32_t *col = nullptr;
    float *f32 = nullptr;
    uint16_t *hi = nullptr, *lo = nullptr;
    // schedule
    uint32_t k = 0;
    std::vector<uint8_t> elog2;      // per matrix
    std::vector<uint32_t> tile_nnz;  // per matrix
    tile_t *tiles = nullptr;
    size_t tiles_bytes = 0;
    uint32_t n_tiles = 0;
    uint32_t *counter = nullptr;
    // captured graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<void *, size_t>> owned;
    void *alloc(size_t bytes)
    {
        void *p = dev_alloc(dev, bytes);
        owned.push_back({p, bytes ? bytes : 8});
        return p;
    }
};

namespace spira_amd {

static void batch_build_tiles(spira_batch *b)
{
    const uint32_t K = b->k;
    std::vector<tile_t> tl;
    std::vector<uint64_t> work;
    for (uint32_t m = 0; m < b->n_mats; ++m) {
        const uint64_t r0 = b->row_base[m], r1 = b->row_base[m + 1];
        const uint64_t budget = std::max<uint64_t>(b->tile_nnz[m], 1);
        const uint32_t el = b->elog2[m];
        const uint64_t max_rows = 1024;
        uint64_t r = r0;
        while (r < r1) {
            uint64_t e = r + 1, acc = b->h_off[r + 1] - b->h_off[r];
            while (e < r1 && e - r < max_rows && acc + (b->h_off[e + 1] - b->h_off[e]) <= budget) {
                acc += b->h_off[e + 1] - b->h_off[e];
                ++e;
            }
            tl.push_back(tile_t{(uint32_t)r, (uint32_t)e, el, 0});
            work.push_back(acc * K + (e - r) * K);
            r = e;
        }
    }
    // heaviest tiles first (longest-processing-time order for the queue)
    std::vector<uint32_t> order(tl.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return work[a] > work[c]; });
    std::vector<tile_t> sorted(tl.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = tl[order[i]];
    if (b->tiles) {
        SPIRA_HIP(hipStreamSynchronize(b->dev->stream));
        dev_release(b->dev, b->tiles, b->tiles_bytes);
        b->tiles = nullptr;
    }
    b->tiles_bytes = std::max<size_t>(sorted.size(), 1) * sizeof(tile_t);
    b->tiles = (tile_t *)dev_alloc(b->dev, b->tiles_bytes);
    SPIRA_HIP(hipMemcpyAsync(b->tiles, sorted.data(), sorted.size() * sizeof(tile_t), hipMemcpyHostToDevice,
                             b->dev->stream));
    dev_sync(b->dev);
    b->n_tiles = (uint32_t)sorted.size();
}

static void batch_launch(spira_batch *b, const float *X, float *Y)
{
    hipStream_t st = b->dev->stream;
    bvals v{b->storage, b->f32, b->hi, b->lo};
    const unsigned grid = (unsigned)std::max<uint32_t>(1u, std::min<uint32_t>(b->n_tiles, (uint32_t)b->dev->n_cu * 8u));
    k_reset_counter<<<1, 1, 0, st>>>(b->counter);
    switch (b->k) {
    case 1: k_batch_spmm<1><<<grid, kThreads, 0, st>>>(b->tiles, b->n_tiles, b->counter, b->off32, b->col, v, X, Y); break;
    case 4: k_batch_spmm<4><<<grid, kThreads, 0, st>>>(b->tiles, b->n_tiles, b->counter, b->off32, b->col, v, X, Y); break;
    case 16: k_batch_spmm<16><<<grid, kThreads, 0, st>>>(b->tiles, b->n_tiles, b->counter, b->off32, b->col, v, X, Y); break;
    default: k_batch_spmm<32><<<grid, kThreads, 0, st>>>(b->tiles, b->n_tiles, b->counter, b->off32, b->col, v, X, Y); break;
    }
    SPIRA_HIP(hipGetLastError());
}

static void batch_drop_graph(spira_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static uint32_t default_elog2(double mean_len, uint32_t K)
{
    uint32_t maxe = 0;
    while ((1u << (maxe + 1)) * K <= 64u) ++maxe;
    uint32_t el = 0;
    while (el < maxe && (double)(1u << el) * 2.0 < mean_len) ++el;
    return el;
}

}  // namespace spira_amd

namespace {
thread_local std::string g_batch_error;

template <class F>
int bguard(spira_dev *dev, F &&f)
{
    std::string *sink = dev ? &dev->last_error : &g_batch_error;
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
}  // namespace

extern "C" {

int spira_batch_create(spira_dev *dev, uint32_t n_mats, spira_dcsr *const *mats, int storage, spira_batch **out)
{
    if (!dev || !out || (n_mats && !mats)) return SPIRA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    spira_batch *b = new spira_batch();
    b->dev = dev;
    int rc = bguard(dev, [&] {
        if (storage != SPIRA_VAL_F32 && storage != SPIRA_VAL_BF16 && storage != SPIRA_VAL_SPLIT_BF16)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: unknown value storage");
        b->n_mats = n_mats;
        b->storage = storage;
        b->row_base.assign(n_mats + 1, 0);
        b->col_base.assign(n_mats + 1, 0);
        b->nnz_base.assign(n_mats + 1, 0);
        for (uint32_t m = 0; m < n_mats; ++m) {
            const spira_dcsr *A = mats[m];
            if (!A || !A->built) fail(SPIRA_ERR_LOGIC, "batch: every matrix must be locked");
            if (A->dev != dev) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: matrices must belong to this device handle");
            b->row_base[m + 1] = b->row_base[m] + A->n_rows;
            b->col_base[m + 1] = b->col_base[m] + A->n_cols;
            b->nnz_base[m + 1] = b->nnz_base[m] + A->nnz;
        }
        b->n_rows = b->row_base[n_mats];
        b->n_cols = b->col_base[n_mats];
        b->nnz = b->nnz_base[n_mats];
        if (b->nnz >= (1ull << 32) || b->n_rows >= (1ull << 32) || b->n_cols >= (1ull << 32))
            fail(SPIRA_ERR_UNSUPPORTED, "batch: totals must stay below 2^32 rows, columns and nonzeros");
        uint64_t *off = (uint64_t *)dev->ws.get(WS_MISC2, (b->n_rows + 1) * sizeof(uint64_t));
        b->col = (uint32_t *)b->alloc((b->nnz ? b->nnz : 1) * sizeof(uint32_t));
        if (storage == SPIRA_VAL_F32) b->f32 = (float *)b->alloc((b->nnz ? b->nnz : 1) * sizeof(float));
        else b->hi = (uint16_t *)b->alloc((b->nnz ? b->nnz : 1) * sizeof(uint16_t));
        if (storage == SPIRA_VAL_SPLIT_BF16) b->lo = (uint16_t *)b->alloc((b->nnz ? b->nnz : 1) * sizeof(uint16_t));
        for (uint32_t m = 0; m < n_mats; ++m) {
            const spira_dcsr *A = mats[m];
            k_cat_off<<<grid_for(A->n_rows + 1, 256, 1024), 256, 0, dev->stream>>>(A->off, A->n_rows, b->nnz_base[m],
                                                                                 b->row_base[m], off);
            if (A->nnz)
                k_cat_entries<<<grid_for(A->nnz, 256, 2048), 256, 0, dev->stream>>>(
                    A->col, A->val, A->nnz, b->nnz_base[m], (uint32_t)b->col_base[m], storage, b->col, b->f32, b->hi,
                    b->lo);
        }
        SPIRA_HIP(hipGetLastError());
        if (n_mats == 0) SPIRA_HIP(hipMemsetAsync(off, 0, sizeof(uint64_t), dev->stream));
        b->off32 = (uint32_t *)b->alloc((b->n_rows + 1) * sizeof(uint32_t));
        k_u64_to_u32_b<<<grid_for(b->n_rows + 1, 256), 256, 0, dev->stream>>>(off, b->off32, b->n_rows + 1);
        b->h_off.resize(b->n_rows + 1);
        SPIRA_HIP(hipMemcpyAsync(b->h_off.data(), off, (b->n_rows + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                 dev->stream));
        b->counter = (uint32_t *)b->alloc(64);
        dev_sync(dev);
    });
    if (rc != SPIRA_OK) {
        for (auto &o : b->owned) dev_release(dev, o.first, o.second);
        delete b;
        return rc;
    }
    *out = b;
    return SPIRA_OK;
}

void spira_batch_destroy(spira_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->dev->device);
    (void)hipStreamSynchronize(b->dev->stream);
    batch_drop_graph(b);
    if (b->tiles) dev_release(b->dev, b->tiles, b->tiles_bytes);
    for (auto &o : b->owned) dev_release(b->dev, o.first, o.second);
    delete b;
}

int spira_batch_info(const spira_batch *b, uint64_t *n_rows, uint64_t *n_cols, uint64_t *nnz, uint32_t *n_tiles)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    if (n_rows) *n_rows = b->n_rows;
    if (n_cols) *n_cols = b->n_cols;
    if (nnz) *nnz = b->nnz;
    if (n_tiles) *n_tiles = b->n_tiles;
    return SPIRA_OK;
}

int spira_batch_bases(const spira_batch *b, uint64_t *row_base, uint64_t *col_base)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    if (row_base) std::memcpy(row_base, b->row_base.data(), (b->n_mats + 1) * sizeof(uint64_t));
    if (col_base) std::memcpy(col_base, b->col_base.data(), (b->n_mats + 1) * sizeof(uint64_t));
    return SPIRA_OK;
}

int spira_batch_set_schedule(spira_batch *b, uint32_t k, const uint8_t *elog2, const uint32_t *tile_nnz)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (k != 1 && k != 4 && k != 16 && k != 32)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: k must be 1, 4, 16 or 32");
        uint32_t maxe = 0;
        while ((1u << (maxe + 1)) * k <= 64u) ++maxe;
        b->k = k;
        b->elog2.resize(b->n_mats);
        b->tile_nnz.resize(b->n_mats);
        for (uint32_t m = 0; m < b->n_mats; ++m) {
            const uint64_t rows = b->row_base[m + 1] - b->row_base[m];
            const uint64_t nz = b->nnz_base[m + 1] - b->nnz_base[m];
            const double mean = rows ? (double)nz / (double)rows : 0.0;
            uint32_t el = elog2 ? elog2[m] : default_elog2(mean, k);
            if (el > maxe) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: lanes per row exceed 64 / k");
            b->elog2[m] = (uint8_t)el;
            b->tile_nnz[m] = tile_nnz ? tile_nnz[m] : std::max<uint32_t>(256u, 16384u / k);
        }
        batch_drop_graph(b);
        batch_build_tiles(b);
    });
}

int spira_batch_spmm(spira_batch *b, const float *X, float *Y, uint32_t flags)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (!b->k) fail(SPIRA_ERR_LOGIC, "batch: set a schedule first");
        if ((b->n_cols && !X) || (b->n_rows && !Y)) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: X or Y is NULL");
        batch_launch(b, X, Y);
        if (!(flags & SPIRA_NO_SYNC)) dev_sync(b->dev);
    });
}

int spira_batch_graph_capture(spira_batch *b, const float *X, float *Y)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (!b->k) fail(SPIRA_ERR_LOGIC, "batch: set a schedule first");
        batch_drop_graph(b);
        hipStream_t st = b->dev->stream;
        SPIRA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            batch_launch(b, X, Y);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SPIRA_HIP(hipStreamEndCapture(st, &b->graph));
        SPIRA_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    });
}

int spira_batch_graph_launch(spira_batch *b, uint32_t flags)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (!b->exec) fail(SPIRA_ERR_LOGIC, "batch: no captured graph");
        SPIRA_HIP(hipGraphLaunch(b->exec, b->dev->stream));
        if (!(flags & SPIRA_NO_SYNC)) dev_sync(b->dev);
    });
}

}  // extern "C"
