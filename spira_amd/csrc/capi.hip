// capi.hip -- the extern "C" boundary of libspira_amd (include/spira_amd.h).
// Converts status_error / bad_alloc into status codes and keeps the message on
// the handle.
#include "internal.hpp"

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>

using namespace spira_amd;

namespace {

thread_local std::string g_last_error;  // for calls without a usable handle

// Fixed scratch of spira_spmv_f64_noalloc: the XPHASE_PACED pacing counters
// or the XSLICE/XSTREAM partial sums.  Every GPU holds its own copy until the
// process ends.  On the host `scratch` only names the symbol: kernels get the
// device address from hipGetSymbolAddress, queried once per device with that
// device current (the first spira_dev_create on it) and kept below.
constexpr size_t kNoAllocScratchBytes = SPIRA_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (the caller has made `device` current).
void *scratch_of_device(int device)
{
    if (device < 0 || device >= kMaxDevices) fail(SPIRA_ERR_UNSUPPORTED, "device ordinal beyond this build's table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        SPIRA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

template <class F>
int guarded(spira_dev *dev, F &&f)
{
    std::string *sink = dev ? &dev->last_error : &g_last_error;
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

// compute_partition_boundaries (partition.hpp:78-126) on a locked CSR: the
// per-row nnz prefix is the offsets array itself.  Boundary t is the first
// row whose prefix reaches max(1, t * total / n_threads); the targets grow
// with t, so searching the whole prefix gives the reference's search from
// boundary t-1.  total == 0 gives the uniform split t * n_rows / n_threads.
__global__ void k_partition_bounds(const uint64_t *__restrict__ off, uint64_t n_rows, uint64_t n_threads,
                                   uint64_t *__restrict__ b)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > n_threads) return;
    if (t == 0) {
        b[0] = 0;
        return;
    }
    if (t == n_threads || n_threads <= 1 || n_rows == 0) {
        b[t] = n_rows;
        return;
    }
    const uint64_t total = off[n_rows];
    if (total == 0) {
        b[t] = t * n_rows / n_threads;
        return;
    }
    // t * total / n_threads without overflowing 64 bits
    const uint64_t q = total / n_threads, r = total % n_threads;
    uint64_t target = t * q + (t * r) / n_threads;
    if (target < 1) target = 1;
    uint64_t lo = 0, hi = n_rows + 1;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (off[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    b[t] = lo < n_rows ? lo : n_rows;
}

void copy_in(spira_dev *dev, void *dst, const void *src, size_t bytes, bool src_device)
{
    if (bytes == 0) return;
    SPIRA_HIP(hipMemcpyAsync(dst, src, bytes, src_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                             dev->stream));
}

}  // namespace

namespace spira_amd {

void free_dcsr_arrays(spira_dcsr *A)
{
    spira_dev *dev = A->dev;
    if (A->off) dev_release(dev, A->off, off_bytes(A));
    if (A->col) dev_release(dev, A->col, col_bytes(A));
    if (A->val) dev_release(dev, A->val, val_bytes(A));
    const size_t cap = A->capacity ? A->capacity : 1;
    if (A->val_f32) dev_release(dev, A->val_f32, cap * sizeof(float));
    if (A->val_bf16) dev_release(dev, A->val_bf16, cap * sizeof(uint16_t));
    if (A->val_split_hi) dev_release(dev, A->val_split_hi, cap * sizeof(uint16_t));
    if (A->val_bf16_lo) dev_release(dev, A->val_bf16_lo, cap * sizeof(uint16_t));
    A->off = nullptr;
    A->col = nullptr;
    A->val = nullptr;
    A->val_f32 = nullptr;
    A->val_bf16 = nullptr;
    A->val_split_hi = nullptr;
    A->val_bf16_lo = nullptr;
    spmv_free_plan(A);
}

}  // namespace spira_amd

extern "C" {

int spira_dev_create(spira_dev **out, int device_ordinal, void *stream)
{
    if (!out) {
        g_last_error = "spira_dev_create: out is NULL";
        return SPIRA_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    spira_dev *dev = nullptr;
    int rc = guarded(nullptr, [&] {
        int count = 0;
        SPIRA_HIP(hipGetDeviceCount(&count));
        if (device_ordinal < 0 || device_ordinal >= count)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "spira_dev_create: no such device ordinal");
        SPIRA_HIP(hipSetDevice(device_ordinal));
        dev = new spira_dev();
        dev->device = device_ordinal;
        hipDeviceProp_t prop;
        SPIRA_HIP(hipGetDeviceProperties(&prop, device_ordinal));
        dev->arch_major = prop.major;
        dev->arch_minor = prop.minor;
        dev->n_cu = prop.multiProcessorCount;
        dev->noalloc_scratch = scratch_of_device(device_ordinal);
        SPIRA_HIP(hipHostMalloc(&dev->host_words, kHostWordsBytes, hipHostMallocDefault));
        // the XPHASE_PACED wait bound (a test forces give-ups with a small one)
        if (const char *e = std::getenv("SPIRA_PHASE_SPIN_MAX")) dev->phase_spin_max = (uint32_t)std::strtoul(e, nullptr, 10);
        if (stream) {
            dev->stream = (hipStream_t)stream;
        } else {
            SPIRA_HIP(hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking));
            dev->own_stream = true;
        }
        dev->ws.stream = dev->stream;
    });
    if (rc != SPIRA_OK) {
        if (dev && dev->host_words) (void)hipHostFree(dev->host_words);
        delete dev;
        return rc;
    }
    *out = dev;
    return SPIRA_OK;
}

void spira_dev_destroy(spira_dev *dev)
{
    if (!dev) return;
    (void)hipSetDevice(dev->device);
    (void)hipStreamSynchronize(dev->stream);
    dev->ws.release();
    dev->cache.release_all();
    if (dev->own_stream) (void)hipStreamDestroy(dev->stream);
    if (dev->host_words) (void)hipHostFree(dev->host_words);
    delete dev;
}

const char *spira_dev_last_error(const spira_dev *dev)
{
    return dev ? dev->last_error.c_str() : g_last_error.c_str();
}

void *spira_dev_stream(const spira_dev *dev) { return dev ? (void *)dev->stream : nullptr; }

int spira_dev_synchronize(spira_dev *dev)
{
    if (!dev) return SPIRA_ERR_INVALID_ARGUMENT;
    return guarded(dev, [&] { dev_sync(dev); });
}

int spira_dcsr_create(spira_dev *dev, uint64_t n_rows, uint64_t n_cols, spira_dcsr **out)
{
    if (!dev || !out) {
        g_last_error = "spira_dcsr_create: NULL argument";
        return SPIRA_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    return guarded(dev, [&] {
        spira_dcsr *A = new spira_dcsr();
        A->dev = dev;
        A->n_rows = n_rows;
        A->n_cols = n_cols;
        *out = A;
    });
}

void spira_dcsr_destroy(spira_dcsr *A)
{
    if (!A) return;
    (void)hipSetDevice(A->dev->device);
    (void)hipStreamSynchronize(A->dev->stream);
    free_dcsr_arrays(A);
    delete A;
}

int spira_dcsr_info(const spira_dcsr *A, uint64_t *n_rows, uint64_t *n_cols, uint64_t *nnz, int *built)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    if (n_rows) *n_rows = A->n_rows;
    if (n_cols) *n_cols = A->n_cols;
    if (nnz) *nnz = A->nnz;
    if (built) *built = A->built ? 1 : 0;
    return SPIRA_OK;
}

int spira_dcsr_upload_f64(spira_dcsr *A, uint64_t nnz, const uint64_t *offsets, const uint32_t *cols,
                          const double *vals, uint32_t flags)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        if (!offsets) fail(SPIRA_ERR_INVALID_ARGUMENT, "spira::matrix::load_csr: CSR offsets array is null");
        if (nnz > 0 && (!cols || !vals)) fail(SPIRA_ERR_INVALID_ARGUMENT, "load_csr: cols/vals are null");
        const bool in_dev = (flags & SPIRA_IN_DEVICE) != 0;
        uint64_t *off = nullptr;
        uint32_t *col = nullptr;
        double *val = nullptr;
        const size_t ob = (A->n_rows + 1) * sizeof(uint64_t), cb = (nnz ? nnz : 1) * sizeof(uint32_t),
                     vb = (nnz ? nnz : 1) * sizeof(double);
        try {
            off = (uint64_t *)dev_alloc(dev, ob);
            col = (uint32_t *)dev_alloc(dev, cb);
            val = (double *)dev_alloc(dev, vb);
            copy_in(dev, off, offsets, (A->n_rows + 1) * sizeof(uint64_t), in_dev);
            copy_in(dev, col, cols, nnz * sizeof(uint32_t), in_dev);
            copy_in(dev, val, vals, nnz * sizeof(double), in_dev);
            bool zeros = false;
            validate_csr_device(A, nnz, off, col, val, &zeros);
            free_dcsr_arrays(A);
            A->off = off;
            A->col = col;
            A->val = val;
            A->nnz = nnz;
            A->capacity = nnz;
            A->built = true;
            A->has_zeros = zeros;
        } catch (...) {
            (void)hipStreamSynchronize(dev->stream);
            dev_release(dev, off, ob);
            dev_release(dev, col, cb);
            dev_release(dev, val, vb);
            throw;
        }
    });
}

int spira_dcsr_download_f64(const spira_dcsr *A, uint64_t *offsets, uint32_t *cols, double *vals, uint32_t flags)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        if (!A->built) fail(SPIRA_ERR_LOGIC, "csr(): no CSR has been built (lock() first)");
        const hipMemcpyKind kind = (flags & SPIRA_OUT_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        if (offsets)
            SPIRA_HIP(hipMemcpyAsync(offsets, A->off, (A->n_rows + 1) * sizeof(uint64_t), kind, dev->stream));
        if (cols && A->nnz)
            SPIRA_HIP(hipMemcpyAsync(cols, A->col, A->nnz * sizeof(uint32_t), kind, dev->stream));
        if (vals && A->nnz)
            SPIRA_HIP(hipMemcpyAsync(vals, A->val, A->nnz * sizeof(double), kind, dev->stream));
        dev_sync(dev);
    });
}

int spira_lock_f64(spira_dcsr *A, const uint64_t *stage_ptr, const uint32_t *stage_col, const double *stage_val,
                   const uint8_t *dirty, uint32_t flags, uint64_t *out_nnz)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        if (!stage_ptr) fail(SPIRA_ERR_INVALID_ARGUMENT, "lock: stage_ptr is NULL");
        const bool in_dev = (flags & SPIRA_IN_DEVICE) != 0;
        const uint64_t n = A->n_rows;
        uint64_t m = 0;
        if (in_dev && (flags & SPIRA_COUNT_GIVEN)) {
            if (!out_nnz) fail(SPIRA_ERR_INVALID_ARGUMENT, "lock: SPIRA_COUNT_GIVEN needs out_nnz to hold the count");
            m = *out_nnz;  // checked against stage_ptr[n] on the device before any column is read
        } else if (in_dev) {
            m = read_scalar(dev, stage_ptr + n);
        } else {
            m = stage_ptr[n];
        }
        if (m > 0 && (!stage_col || !stage_val)) fail(SPIRA_ERR_INVALID_ARGUMENT, "lock: staged arrays are NULL");
        const uint64_t *d_ptr = stage_ptr;
        const uint32_t *d_col = stage_col;
        const double *d_val = stage_val;
        const uint8_t *d_dirty = dirty;
        if (!in_dev) {
            uint64_t *p = (uint64_t *)dev->ws.get(WS_STAGE_PTR, (n + 1) * sizeof(uint64_t));
            uint32_t *c = (uint32_t *)dev->ws.get(WS_STAGE_COL, m * sizeof(uint32_t));
            double *v = (double *)dev->ws.get(WS_STAGE_VAL, m * sizeof(double));
            copy_in(dev, p, stage_ptr, (n + 1) * sizeof(uint64_t), false);
            copy_in(dev, c, stage_col, m * sizeof(uint32_t), false);
            copy_in(dev, v, stage_val, m * sizeof(double), false);
            d_ptr = p;
            d_col = c;
            d_val = v;
            if (dirty) {
                uint8_t *d = (uint8_t *)dev->ws.get(WS_X, n);
                copy_in(dev, d, dirty, n, false);
                d_dirty = d;
            }
        }
        const uint64_t nnz = lock_device_f64(A, d_ptr, d_col, d_val, m, d_dirty);
        if (out_nnz) *out_nnz = nnz;
    });
}

int spira_spmv_f64(spira_dcsr *A, const double *x, uint64_t x_len, double *y, uint64_t y_len, int variant,
                   uint32_t flags)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        // serial/spmv.hpp:150-158: sizes first, then the mode.
        if (x_len != A->n_cols)
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "The size of the input vector x does not match the number of columns of the matrix");
        if (y_len != A->n_rows)
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "The size of the output vector y does not match the number of rows of the matrix");
        if (!A->built) fail(SPIRA_ERR_LOGIC, "spmv: matrix must be locked");
        if (variant < SPIRA_SPMV_AUTO || variant >= SPIRA_SPMV_N_VARIANTS)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: unknown kernel variant");
        if ((x_len && !x) || (y_len && !y)) fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: x or y is NULL");
        const double *dx = x;
        double *dy = y;
        const bool xd = flags & SPIRA_X_DEVICE, yd = flags & SPIRA_Y_DEVICE;
        if (!xd) {
            double *t = (double *)dev->ws.get(WS_X, x_len * sizeof(double));
            copy_in(dev, t, x, x_len * sizeof(double), false);
            dx = t;
        }
        if (!yd) dy = (double *)dev->ws.get(WS_Y, y_len * sizeof(double));
        spmv_launch_f64(A, dx, dy, variant, (flags & SPIRA_REPRODUCIBLE) != 0);
        if (!yd && y_len)
            SPIRA_HIP(hipMemcpyAsync(y, dy, y_len * sizeof(double), hipMemcpyDeviceToHost, dev->stream));
        if (!(flags & SPIRA_NO_SYNC) || !xd || !yd) dev_sync(dev);
    });
}

int spira_spmv_f64_noalloc(spira_dcsr *A, const double *x, uint64_t x_len, double *y, uint64_t y_len, int variant,
                           uint32_t flags)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        // the checks of spira_spmv_f64, in the same order
        if (x_len != A->n_cols)
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "The size of the input vector x does not match the number of columns of the matrix");
        if (y_len != A->n_rows)
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "The size of the output vector y does not match the number of rows of the matrix");
        if (!A->built) fail(SPIRA_ERR_LOGIC, "spmv: matrix must be locked");
        if (variant < SPIRA_SPMV_AUTO || variant >= SPIRA_SPMV_N_VARIANTS)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: unknown kernel variant");
        if ((x_len && !x) || (y_len && !y)) fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: x or y is NULL");
        if ((flags & (SPIRA_X_DEVICE | SPIRA_Y_DEVICE)) != (SPIRA_X_DEVICE | SPIRA_Y_DEVICE))
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "spmv_noalloc: x and y must be device pointers (SPIRA_X_DEVICE | SPIRA_Y_DEVICE)");
        if (A->n_rows == 0) return;
        if (A->n_rows >= (1ull << 32) - 1) fail(SPIRA_ERR_UNSUPPORTED, "spmv: n_rows must be < 2^32 - 1");
        if (!dev->noalloc_scratch) fail(SPIRA_ERR_DEVICE, "spmv_noalloc: no scratch address for this device");
        const char *unprepared =
            "spmv_noalloc: the matrix has no SpMV plan for this variant (one spira_spmv_f64 call with the "
            "variant after the last lock or upload builds it)";
        const bool repro = (flags & SPIRA_REPRODUCIBLE) != 0;
        if (repro && variant >= 0 && !spmv_order_fixed(variant))
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "spmv: XPHASE variants add in a varying order; drop SPIRA_REPRODUCIBLE or choose another variant");
        int v = variant < 0 ? spmv_auto_choice(A, repro) : variant;
        // AUTO whose choice has no plan yet: the variant the last AUTO call of
        // spira_spmv_f64 ran (a plan-free one on the first call after a lock)
        // if it fits the call's flags
        const int ran = A->plan.auto_ran;
        if (variant < 0 && v >= 0 && ran >= 0 && !spmv_plan_ready(A, v) && spmv_plan_ready(A, ran) &&
            (!repro || spmv_order_fixed(ran)))
            v = ran;
        if (v < 0 || !A->plan.stats_valid) fail(SPIRA_ERR_LOGIC, unprepared);
        const uint64_t need = spmv_scratch_bytes(A, v);
        if (need > kNoAllocScratchBytes)
            fail(SPIRA_ERR_UNSUPPORTED, "spmv_noalloc: the variant needs " + std::to_string(need) +
                                            " bytes of scratch, more than SPIRA_NOALLOC_SCRATCH_BYTES (" +
                                            std::to_string(kNoAllocScratchBytes) + ")");
        if (!spmv_plan_ready(A, v)) fail(SPIRA_ERR_LOGIC, unprepared);
        spmv_launch_planned(A, x, y, v, dev->noalloc_scratch);
        if (!(flags & SPIRA_NO_SYNC)) dev_sync(dev);
    });
}

int spira_spmv_f32(spira_dcsr *A, const float *x, uint64_t x_len, float *y, uint64_t y_len, int variant,
                   uint32_t flags)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        // serial/spmv.hpp:98-106: sizes first, then the mode
        if (x_len != A->n_cols)
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "The size of the input vector x does not match the number of columns of the matrix");
        if (y_len != A->n_rows)
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "The size of the output vector y does not match the number of rows of the matrix");
        if (!A->built) fail(SPIRA_ERR_LOGIC, "spmv: matrix must be locked");
        if (variant != SPIRA_SPMV_AUTO && !mixed_variant_supported(variant))
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "spmv (float): variant must be VEC4..VEC32, STREAM1K, STREAM2K, XPHASE or XPHASE_PACED");
        if ((x_len && !x) || (y_len && !y)) fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: x or y is NULL");
        const float *dx = x;
        float *dy = y;
        const bool xd = flags & SPIRA_X_DEVICE, yd = flags & SPIRA_Y_DEVICE;
        if (!xd) {
            float *t = (float *)dev->ws.get(WS_X, x_len * sizeof(float));
            copy_in(dev, t, x, x_len * sizeof(float), false);
            dx = t;
        }
        if (!yd) dy = (float *)dev->ws.get(WS_Y, y_len * sizeof(float));
        spmv_f32_launch(A, dx, dy, variant);
        if (!yd && y_len)
            SPIRA_HIP(hipMemcpyAsync(y, dy, y_len * sizeof(float), hipMemcpyDeviceToHost, dev->stream));
        if (!(flags & SPIRA_NO_SYNC) || !xd || !yd) dev_sync(dev);
    });
}

int spira_spmv_mixed_f64(spira_dcsr *A, int storage, const double *x, uint64_t x_len, double *y, uint64_t y_len,
                         int variant, float recompute_ratio, uint32_t flags, uint64_t *n_recomputed)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        if (x_len != A->n_cols)
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "The size of the input vector x does not match the number of columns of the matrix");
        if (y_len != A->n_rows)
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "The size of the output vector y does not match the number of rows of the matrix");
        if (!A->built) fail(SPIRA_ERR_LOGIC, "spmv: matrix must be locked");
        if (storage != SPIRA_VAL_F32 && storage != SPIRA_VAL_BF16 && storage != SPIRA_VAL_SPLIT_BF16)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "mixed spmv: unknown value storage");
        if (variant != SPIRA_SPMV_AUTO && !mixed_variant_supported(variant))
            fail(SPIRA_ERR_INVALID_ARGUMENT,
                 "mixed spmv: variant must be VEC4..VEC32, STREAM1K, STREAM2K, XPHASE or XPHASE_PACED");
        if ((x_len && !x) || (y_len && !y)) fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: x or y is NULL");
        const double *dx = x;
        double *dy = y;
        const bool xd = flags & SPIRA_X_DEVICE, yd = flags & SPIRA_Y_DEVICE;
        if (!xd) {
            double *t = (double *)dev->ws.get(WS_X, x_len * sizeof(double));
            copy_in(dev, t, x, x_len * sizeof(double), false);
            dx = t;
        }
        if (!yd) dy = (double *)dev->ws.get(WS_Y, y_len * sizeof(double));
        spmv_mixed_launch(A, storage, dx, dy, variant, recompute_ratio);
        if (!yd && y_len)
            SPIRA_HIP(hipMemcpyAsync(y, dy, y_len * sizeof(double), hipMemcpyDeviceToHost, dev->stream));
        if (n_recomputed) {
            uint32_t c = 0;
            if (A->n_rows && recompute_ratio > 0.0f)
                SPIRA_HIP(hipMemcpyAsync(&c, mixed_queue_count(A), sizeof(uint32_t), hipMemcpyDeviceToHost,
                                         dev->stream));
            dev_sync(dev);
            *n_recomputed = c;
        }
        if (!(flags & SPIRA_NO_SYNC) || !xd || !yd) dev_sync(dev);
    });
}

int spira_spmv_choose(spira_dcsr *A, int *variant)
{
    if (!A || !variant) return SPIRA_ERR_INVALID_ARGUMENT;
    return guarded(A->dev, [&] {
        if (!A->built) fail(SPIRA_ERR_LOGIC, "spmv: matrix must be locked");
        *variant = A->n_rows ? spmv_resolve_auto(A) : 0;
    });
}

int spira_scan_u64(spira_dev *dev, const uint64_t *in_dev, uint64_t *out_dev, uint64_t n)
{
    if (!dev || !out_dev || (!in_dev && n)) return SPIRA_ERR_INVALID_ARGUMENT;  // in is not read when n == 0
    return guarded(dev, [&] {
        scan_excl_u64(dev, in_dev, out_dev, n);
        dev_sync(dev);
    });
}

int spira_probe_ceilings(spira_dev *dev, double *gather_req_per_s, double *stream_bytes_per_s)
{
    if (!dev) return SPIRA_ERR_INVALID_ARGUMENT;
    return guarded(dev, [&] { probe_ceilings(dev, gather_req_per_s, stream_bytes_per_s); });
}

int spira_spmv_autotune_f64(spira_dcsr *A, const double *x_dev, double *y_dev, int reps, float *ms, int *best)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        if (!A->built) fail(SPIRA_ERR_LOGIC, "spmv: matrix must be locked");
        if (reps < 1) fail(SPIRA_ERR_INVALID_ARGUMENT, "autotune: reps must be >= 1");
        if ((A->n_cols && !x_dev) || (A->n_rows && !y_dev)) fail(SPIRA_ERR_INVALID_ARGUMENT, "autotune: NULL vector");
        spmv_build_plan(A);
        hipEvent_t e0, e1;
        SPIRA_HIP(hipEventCreate(&e0));
        SPIRA_HIP(hipEventCreate(&e1));
        int best_v = -1;
        float best_ms = 0.0f;
        float t_ms[SPIRA_SPMV_N_VARIANTS];
        try {
            for (int v = 0; v < SPIRA_SPMV_N_VARIANTS; ++v) {
                t_ms[v] = -1.0f;
                try {
                    spmv_launch_f64(A, x_dev, y_dev, v);  // warm-up (builds the plan parts)
                } catch (const status_error &e) {
                    if (e.code != SPIRA_ERR_UNSUPPORTED) throw;
                    if (ms) ms[v] = -1.0f;  // this variant cannot run this matrix (size limits)
                    continue;
                }
                SPIRA_HIP(hipEventRecord(e0, dev->stream));
                for (int r = 0; r < reps; ++r) spmv_launch_f64(A, x_dev, y_dev, v);
                SPIRA_HIP(hipEventRecord(e1, dev->stream));
                SPIRA_HIP(hipEventSynchronize(e1));
                float t = 0.0f;
                SPIRA_HIP(hipEventElapsedTime(&t, e0, e1));
                t /= (float)reps;
                if (ms) ms[v] = t;
                t_ms[v] = t;
                if (best_v < 0 || t < best_ms) {
                    best_ms = t;
                    best_v = v;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        A->plan.variant = best_v;
        A->plan.tuned = true;
        for (int v = 0; v < SPIRA_SPMV_N_VARIANTS; ++v) A->plan.tuned_ms[v] = t_ms[v];
        A->plan.auto_variant = -1;  // AUTO's own choices now follow the measured times
        A->plan.auto_repro = -1;
        if (best) *best = best_v;
    });
}

int spira_spmv_set_variant(spira_dcsr *A, int variant)
{
    if (!A) return SPIRA_ERR_INVALID_ARGUMENT;
    return guarded(A->dev, [&] {
        if (variant < SPIRA_SPMV_AUTO || variant >= SPIRA_SPMV_N_VARIANTS)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: unknown kernel variant");
        A->plan.variant = variant;
    });
}

int spira_partition_boundaries(spira_dcsr *A, uint64_t n_threads, uint64_t *boundaries)
{
    if (!A || !boundaries) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return guarded(dev, [&] {
        if (n_threads == 0) fail(SPIRA_ERR_INVALID_ARGUMENT, "parallel_matrix requires at least one thread");
        if (!A->built) fail(SPIRA_ERR_LOGIC, "rebalance: matrix must be locked");
        const size_t bytes = (n_threads + 1) * sizeof(uint64_t);
        uint64_t *b = (uint64_t *)dev_alloc(dev, bytes);
        try {
            k_partition_bounds<<<grid_for(n_threads + 1, 256), 256, 0, dev->stream>>>(A->off, A->n_rows, n_threads, b);
            SPIRA_HIP(hipGetLastError());
            SPIRA_HIP(hipMemcpyAsync(boundaries, b, bytes, hipMemcpyDeviceToHost, dev->stream));
            dev_sync(dev);
        } catch (...) {
            (void)hipStreamSynchronize(dev->stream);
            dev_release(dev, b, bytes);
            throw;
        }
        dev_release(dev, b, bytes);
    });
}

}  // extern "C"
