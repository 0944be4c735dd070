// internal.hpp -- internal interfaces between the translation units of
// libspira_amd (not part of the C-ABI).
#pragma once

#include "common.hpp"

#include <cstring>

namespace spira_amd {

// scan.hip: out[0..n) = exclusive prefix of in, out[n] = total.  in may equal
// out.  total_copy (optional, device): a second copy of the total.
void scan_excl_u64(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_copy = nullptr);
// The same scan when the kernel that wrote `in` also wrote the sum of every
// chunk of `chunk` words (chunk divides 4,096) into scan_chunk_sums(dev, n,
// chunk)'s array; the scan then skips its own summing pass.
uint64_t *scan_chunk_sums(spira_dev *dev, uint64_t n, uint32_t chunk);
void scan_excl_u64_chunks(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *sums,
                          uint32_t chunk, uint64_t *total_copy = nullptr);

// spmv.hip
void spmv_build_plan(spira_dcsr *A);
void spmv_free_plan(spira_dcsr *A);
// AUTO's choice: the pinned variant, else the resolved learned/rule choice
// (-1 while unresolved; no device work).  spmv_resolve_auto resolves it
// (features + the selector network on the device, once per locked matrix).
// With `repro`, only variants whose summation order is fixed (not XPHASE*).
int spmv_auto_choice(const spira_dcsr *A, bool repro = false);
int spmv_resolve_auto(spira_dcsr *A, bool repro = false);
inline bool spmv_order_fixed(int variant)
{
    return variant != SPIRA_SPMV_XPHASE && variant != SPIRA_SPMV_XPHASE_PACED;
}
// The rule used when no trained selector is compiled in (row-length stats).
int spmv_rule_variant(const spira_dcsr *A);
// Whether A is within `variant`'s size limits (no SPIRA_ERR_UNSUPPORTED).
bool spmv_variant_fits(const spira_dcsr *A, int variant);

// selector.hip: the features of a locked CSR; the selector network's scores
// for A into A->plan.sel_score (false when no trained model is compiled in).
void matrix_features_host(spira_dcsr *A, float *features);
bool spmv_selector_scores(spira_dcsr *A);
bool spmv_stats_and_scores(spira_dcsr *A, uint64_t *max_len, double *sumsq);
void spmv_launch_f64(spira_dcsr *A, const double *x, double *y, int variant, bool repro = false);
// XPHASE (paced: XPHASE_PACED) with values stored as SPIRA_VAL_* in the plan's
// order and x in FP32 (half the x lines to gather), products and row sums in
// FP64: FP64 y (mixed precision), or FP32 y (the reference's float
// instantiation, FP32 values).
// x already rounded to FP32, or FP64 x rounded where the kernel gathers it
void spmv_xphase_reduced(spira_dcsr *A, int storage, const float *x, double *y, bool paced);
void spmv_xphase_reduced(spira_dcsr *A, int storage, const double *x, double *y, bool paced);
void spmv_xphase_f32(spira_dcsr *A, const float *x, float *y, bool paced);
// The allocation-free launch: true when A's plan already holds what `variant`
// (AUTO resolved to the pinned or heuristic choice) needs; the bytes of
// scratch it takes (XPHASE_PACED counters, XSLICE/XSTREAM partial sums); and
// the launch with that scratch passed in.  None of them allocates.
bool spmv_plan_ready(const spira_dcsr *A, int variant);
uint64_t spmv_scratch_bytes(const spira_dcsr *A, int variant);
void spmv_launch_planned(spira_dcsr *A, const double *x, double *y, int variant, void *scratch);
// Row-block list of the STREAM kernels (which: 0 = chunk 1024, 1 = 2048).
void spmv_stream_blocks(spira_dcsr *A, int which, const uint32_t **blk, uint32_t *nblk);

// spmv_mixed.hip
void spmv_mixed_launch(spira_dcsr *A, int storage, const double *x, double *y, int variant, float ratio);
uint32_t *mixed_queue_count(spira_dcsr *A);
bool mixed_variant_supported(int variant);
// The reference's float instantiation: FP32 values (A's values rounded), FP32
// x and y; each row summed in FP32 with compensation (VEC/STREAM) or in FP64
// (XPHASE), so y is the row sum rounded once to FP32 up to the compensation's
// error.  variant: AUTO, VEC4..VEC32, STREAM1K/2K, XPHASE, XPHASE_PACED.
void spmv_f32_launch(spira_dcsr *A, const float *x, float *y, int variant);

// probe.hip: measured ceilings of the part (L2-resident random 8-byte
// gathers, one request each, in requests/s; a 1 GiB streaming read in B/s).
void probe_ceilings(spira_dev *dev, double *gather_req_per_s, double *stream_bytes_per_s);

// lock.hip: runs the device lock pipeline on already-device-resident staged
// arrays; installs the new CSR in A.  Returns the new nnz.
uint64_t lock_device_f64(spira_dcsr *A, const uint64_t *d_stage_ptr, const uint32_t *d_stage_col,
                         const double *d_stage_val, uint64_t m, const uint8_t *d_dirty);

// lock.hip: load_csr checks on device arrays; throws status_error.
void validate_csr_device(spira_dcsr *A, uint64_t nnz, const uint64_t *d_off, const uint32_t *d_col,
                         const double *d_val, bool *has_zeros);

// Synchronise the handle's stream (throws on failure).
inline void dev_sync(spira_dev *dev) { SPIRA_HIP(hipStreamSynchronize(dev->stream)); }

// Copy `bytes` of device memory to host memory after the stream's queued
// work and wait.  Up to kHostWordsBytes go through the handle's pinned
// buffer: an asynchronous copy into pinned memory, where pageable memory
// would take a staged, blocking copy.
constexpr size_t kHostWordsBytes = 65536;
inline void read_back(spira_dev *dev, void *dst, const void *dsrc, size_t bytes)
{
    if (bytes == 0) return;
    if (dev->host_words && bytes <= kHostWordsBytes) {
        SPIRA_HIP(hipMemcpyAsync(dev->host_words, dsrc, bytes, hipMemcpyDeviceToHost, dev->stream));
        dev_sync(dev);
        std::memcpy(dst, dev->host_words, bytes);
    } else {
        SPIRA_HIP(hipMemcpyAsync(dst, dsrc, bytes, hipMemcpyDeviceToHost, dev->stream));
        dev_sync(dev);
    }
}

// Read one device word.
template <class T>
inline T read_scalar(spira_dev *dev, const T *dptr)
{
    T v{};
    read_back(dev, &v, dptr, sizeof(T));
    return v;
}

void free_dcsr_arrays(spira_dcsr *A);

// Device memory through the handle's block cache.
inline void *dev_alloc(spira_dev *dev, size_t bytes) { return dev->cache.alloc(bytes); }
inline void dev_release(spira_dev *dev, void *p, size_t bytes) { dev->cache.release(p, bytes); }

// Plan memory: recorded on the plan and released with it.
inline void *plan_alloc(spira_dcsr *A, size_t bytes)
{
    void *p = dev_alloc(A->dev, bytes);
    A->plan.owned.push_back({p, bytes ? bytes : 8});
    return p;
}

// Bytes of the CSR arrays a handle holds.
inline size_t off_bytes(const spira_dcsr *A) { return (A->n_rows + 1) * sizeof(uint64_t); }
inline size_t col_bytes(const spira_dcsr *A) { return (A->capacity ? A->capacity : 1) * sizeof(uint32_t); }
inline size_t val_bytes(const spira_dcsr *A) { return (A->capacity ? A->capacity : 1) * sizeof(double); }

}  // namespace spira_amd
