// stage.hip -- open-mode insert staging on the device.
//
// The reference keeps open-mode inserts in per-row host buffers
// (row.hpp:162-169; parallel_matrix::insert routes them to the owning
// partition, parallel_matrix.hpp:581-602, optionally through per-partition
// staging arrays that are burst-flushed in order, insert_staging.hpp:58-96).
// Here a spira_stage is one device COO append buffer (row, col, value in
// insertion order).  spira_lock_staged_f64 turns it into the per-row
// segments lock() consumes with a stable radix sort by row (rocPRIM), which
// keeps each row's inserts in insertion order -- the order last-write-wins
// deduplication depends on -- then runs the device lock pipeline.
#include "internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <cstring>

struct spira_stage {
    spira_dev *dev = nullptr;
    uint64_t n_rows = 0, n_cols = 0;
    uint64_t n = 0, cap = 0;
    uint32_t *row = nullptr, *col = nullptr;
    double *val = nullptr;
};

namespace spira_amd {

namespace {

constexpr int kThreads = 256;
enum : uint32_t { BAD_ROW = 1u, BAD_COL = 2u };

__global__ void k_stage_check(const uint32_t *__restrict__ row, const uint32_t *__restrict__ col, uint64_t n,
                              uint64_t n_rows, uint64_t n_cols, uint32_t *__restrict__ err)
{
    uint32_t bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (row[i] >= n_rows) bad |= BAD_ROW;
        if (col[i] >= n_cols) bad |= BAD_COL;
    }
    if (bad) atomicOr(err, bad);
}

__global__ void k_iota_u32(uint32_t *__restrict__ v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}

// stage_col / stage_val in sorted order
__global__ void k_stage_gather(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ col,
                               const double *__restrict__ val, uint64_t n, uint32_t *__restrict__ scol,
                               double *__restrict__ sval)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = perm[i];
        scol[i] = col[p];
        sval[i] = val[p];
    }
}

// stage_ptr[r] = first sorted position with row >= r, r = 0..n_rows
__global__ void k_stage_ptr(const uint32_t *__restrict__ keys, uint64_t n, uint64_t n_rows, uint64_t *__restrict__ ptr)
{
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= n_rows;
         r += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if ((uint64_t)keys[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        ptr[r] = lo;
    }
}

unsigned bits_for(uint64_t n_rows)
{
    unsigned b = 1;
    while (b < 32 && (1ull << b) < n_rows) ++b;
    return b;
}

void stage_release(spira_stage *s)
{
    const size_t c = s->cap ? s->cap : 1;
    if (s->row) dev_release(s->dev, s->row, c * sizeof(uint32_t));
    if (s->col) dev_release(s->dev, s->col, c * sizeof(uint32_t));
    if (s->val) dev_release(s->dev, s->val, c * sizeof(double));
    s->row = s->col = nullptr;
    s->val = nullptr;
}

void stage_grow(spira_stage *s, uint64_t need)
{
    if (need <= s->cap) return;
    const uint64_t cap = std::max<uint64_t>(need, std::max<uint64_t>(2 * s->cap, 1u << 16));
    spira_dev *dev = s->dev;
    uint32_t *r = (uint32_t *)dev_alloc(dev, cap * sizeof(uint32_t));
    uint32_t *c = (uint32_t *)dev_alloc(dev, cap * sizeof(uint32_t));
    double *v = (double *)dev_alloc(dev, cap * sizeof(double));
    if (s->n) {
        SPIRA_HIP(hipMemcpyAsync(r, s->row, s->n * sizeof(uint32_t), hipMemcpyDeviceToDevice, dev->stream));
        SPIRA_HIP(hipMemcpyAsync(c, s->col, s->n * sizeof(uint32_t), hipMemcpyDeviceToDevice, dev->stream));
        SPIRA_HIP(hipMemcpyAsync(v, s->val, s->n * sizeof(double), hipMemcpyDeviceToDevice, dev->stream));
    }
    SPIRA_HIP(hipStreamSynchronize(dev->stream));  // the old arrays go back to the cache
    stage_release(s);
    s->row = r;
    s->col = c;
    s->val = v;
    s->cap = cap;
}

thread_local std::string g_stage_error;

template <class F>
int sguard(spira_dev *dev, F &&f)
{
    std::string *sink = dev ? &dev->last_error : &g_stage_error;
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

}  // namespace spira_amd

using namespace spira_amd;

extern "C" {

int spira_stage_create(spira_dev *dev, uint64_t n_rows, uint64_t n_cols, spira_stage **out)
{
    if (!dev || !out) return SPIRA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (n_rows >= (1ull << 32) || n_cols > (1ull << 32)) {
        dev->last_error = "stage: n_rows must be < 2^32 and n_cols <= 2^32";
        return SPIRA_ERR_UNSUPPORTED;
    }
    spira_stage *s = new spira_stage();
    s->dev = dev;
    s->n_rows = n_rows;
    s->n_cols = n_cols;
    *out = s;
    return SPIRA_OK;
}

void spira_stage_destroy(spira_stage *s)
{
    if (!s) return;
    (void)hipSetDevice(s->dev->device);
    (void)hipStreamSynchronize(s->dev->stream);
    stage_release(s);
    delete s;
}

int spira_stage_size(const spira_stage *s, uint64_t *n)
{
    if (!s || !n) return SPIRA_ERR_INVALID_ARGUMENT;
    *n = s->n;
    return SPIRA_OK;
}

int spira_stage_clear(spira_stage *s)
{
    if (!s) return SPIRA_ERR_INVALID_ARGUMENT;
    s->n = 0;
    return SPIRA_OK;
}

int spira_stage_append(spira_stage *s, const uint32_t *rows, const uint32_t *cols, const double *vals, uint64_t n,
                       uint32_t flags)
{
    if (!s) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = s->dev;
    return sguard(dev, [&] {
        if (n == 0) return;
        if (!rows || !cols || !vals) fail(SPIRA_ERR_INVALID_ARGUMENT, "stage: NULL input array");
        if (s->n + n >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "stage: fewer than 2^32 staged inserts supported");
        stage_grow(s, s->n + n);
        const hipMemcpyKind kind = (flags & SPIRA_IN_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        hipStream_t st = dev->stream;
        SPIRA_HIP(hipMemcpyAsync(s->row + s->n, rows, n * sizeof(uint32_t), kind, st));
        SPIRA_HIP(hipMemcpyAsync(s->col + s->n, cols, n * sizeof(uint32_t), kind, st));
        SPIRA_HIP(hipMemcpyAsync(s->val + s->n, vals, n * sizeof(double), kind, st));
        // the whole append is rejected if any index is out of range (matrix.hpp:161-171)
        uint32_t *err = (uint32_t *)dev->ws.get(WS_MISC, 64);
        SPIRA_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), st));
        k_stage_check<<<grid_for(n, kThreads, 4096), kThreads, 0, st>>>(s->row + s->n, s->col + s->n, n, s->n_rows,
                                                                       s->n_cols, err);
        SPIRA_HIP(hipGetLastError());
        const uint32_t bad = read_scalar(dev, err);
        if (bad & BAD_ROW) fail(SPIRA_ERR_OUT_OF_RANGE, "spira::matrix: row_index out of range");
        if (bad & BAD_COL) fail(SPIRA_ERR_OUT_OF_RANGE, "spira::matrix: col_index out of range");
        s->n += n;
    });
}

int spira_lock_staged_f64(spira_dcsr *A, spira_stage *s, const uint8_t *dirty, uint32_t flags, uint64_t *out_nnz)
{
    if (!A || !s) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return sguard(dev, [&] {
        if (s->dev != dev) fail(SPIRA_ERR_INVALID_ARGUMENT, "stage: belongs to another device handle");
        if (s->n_rows != A->n_rows || s->n_cols != A->n_cols)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "stage: shape does not match the matrix");
        hipStream_t st = dev->stream;
        const uint64_t n = A->n_rows, m = s->n;
        const size_t mb = (m ? m : 1);
        // sorted arrays (released at the end of the call)
        uint32_t *keys = (uint32_t *)dev_alloc(dev, mb * sizeof(uint32_t));
        uint32_t *idx = (uint32_t *)dev_alloc(dev, mb * sizeof(uint32_t));
        uint32_t *perm = (uint32_t *)dev_alloc(dev, mb * sizeof(uint32_t));
        uint32_t *scol = (uint32_t *)dev_alloc(dev, mb * sizeof(uint32_t));
        double *sval = (double *)dev_alloc(dev, mb * sizeof(double));
        uint64_t *sptr = (uint64_t *)dev_alloc(dev, (n + 1) * sizeof(uint64_t));
        auto release = [&] {
            dev_release(dev, keys, mb * sizeof(uint32_t));
            dev_release(dev, idx, mb * sizeof(uint32_t));
            dev_release(dev, perm, mb * sizeof(uint32_t));
            dev_release(dev, scol, mb * sizeof(uint32_t));
            dev_release(dev, sval, mb * sizeof(double));
            dev_release(dev, sptr, (n + 1) * sizeof(uint64_t));
        };
        try {
            if (m) {
                k_iota_u32<<<grid_for(m, kThreads, 4096), kThreads, 0, st>>>(idx, m);
                size_t tmp_bytes = 0;
                SPIRA_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, s->row, keys, idx, perm, (unsigned)m, 0,
                                                    bits_for(n), st));
                void *tmp = dev->ws.get(WS_SORT, tmp_bytes ? tmp_bytes : 8);
                // LSD radix sort: stable, so equal rows keep insertion order
                SPIRA_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, s->row, keys, idx, perm, (unsigned)m, 0,
                                                    bits_for(n), st));
                k_stage_gather<<<grid_for(m, kThreads, 4096), kThreads, 0, st>>>(perm, s->col, s->val, m, scol, sval);
            }
            k_stage_ptr<<<grid_for(n + 1, kThreads, 4096), kThreads, 0, st>>>(keys, m, n, sptr);
            SPIRA_HIP(hipGetLastError());
            const uint8_t *d_dirty = dirty;
            if (dirty && !(flags & SPIRA_IN_DEVICE)) {
                uint8_t *d = (uint8_t *)dev->ws.get(WS_X, n ? n : 1);
                SPIRA_HIP(hipMemcpyAsync(d, dirty, n, hipMemcpyHostToDevice, st));
                d_dirty = d;
            }
            const uint64_t nnz = lock_device_f64(A, sptr, scol, sval, m, d_dirty);
            if (out_nnz) *out_nnz = nnz;
        } catch (...) {
            (void)hipStreamSynchronize(st);
            release();
            throw;
        }
        release();  // stream order keeps them intact for the lock's queued kernels
        s->n = 0;
    });
}

}  // extern "C"
