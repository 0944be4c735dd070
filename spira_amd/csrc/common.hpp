// common.hpp -- shared device helpers and the internal host-side objects of
// libspira_amd (device handle, device CSR, workspace).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spira_amd.h"

namespace spira_amd {

// ---------------------------------------------------------------------------
// Errors.  Every C-ABI entry point converts these into status codes; the codes
// map 1:1 onto the exceptions the reference throws (invalid_argument,
// logic_error, out_of_range) plus device and allocation failures.
// ---------------------------------------------------------------------------
struct status_error : std::runtime_error {
    int code;
    status_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw status_error(code, msg); }

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            fail(SPIRA_ERR_NOMEM, std::string(what) + ": " + hipGetErrorString(e));
        fail(SPIRA_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define SPIRA_HIP(call) ::spira_amd::hip_check((call), #call)

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

// Native vector types (usable with the non-temporal load builtin).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ T ld_stream(const T *p)
{
    // Streaming (non-temporal) load: A's values and columns are touched once
    // per SpMV, so they should not evict x from L2 / MALL.
    return __builtin_nontemporal_load(p);
}

// Unsigned 32-bit division by a divisor fixed for a launch: one multiply-high,
// one add and one shift (Granlund & Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994, Figure 4.1).  For 1 <= d < 2^32,
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1 (< 2^32):
// n / d = (mulhi(m, n) + n) >> l, the sum taken in 33 bits.
struct udiv32 {
    uint32_t mul = 1, shift = 0;  // d = 1
};
inline udiv32 make_udiv32(uint32_t d)  // d >= 1 (d <= 1 gives the identity)
{
    udiv32 f;
    if (d <= 1) return f;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.mul = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    f.shift = l;
    return f;
}
__device__ __forceinline__ uint32_t udiv(uint32_t n, udiv32 f)
{
    return (uint32_t)(((uint64_t)__umulhi(n, f.mul) + n) >> f.shift);
}

// Sum over aligned groups of G lanes (G a power of two <= 64).
template <int G>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, G);
    return v;
}

__device__ __forceinline__ double group_sum_rt(double v, int g)
{
    for (int s = g / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    return v;
}

template <int G>
__device__ __forceinline__ float group_sum_f(float v)
{
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, G);
    return v;
}

// BF16 <-> FP32.  Round-to-nearest-even FP32 -> BF16 (NaN kept quiet).
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// is_zero(v) <=> |v| <= 0 (traits.hpp:73-88): +0.0 and -0.0 only; NaN and
// subnormals are not zero.  Bit test, so no dependence on denormal flushing.
__device__ __forceinline__ bool is_zero_f64(double v)
{
    return (__double_as_longlong(v) & 0x7fffffffffffffffll) == 0;
}

// Largest r in [lo, hi) with ptr[r] <= k (the row that owns element k when the
// rows are non-empty).  ptr is non-decreasing.
template <class OffT>
__device__ __forceinline__ uint64_t row_of(const OffT *ptr, uint64_t lo, uint64_t hi, uint64_t k)
{
    while (hi - lo > 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)ptr[mid] <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}

// row_of with the 64 lanes of a wavefront probing together: each step narrows
// [lo, hi) 64-fold, so a million rows cost 4 dependent loads instead of 20.
// Every lane of the wavefront must be active and pass the same arguments, and
// ptr[lo] <= k must hold.
template <class OffT>
__device__ __forceinline__ uint64_t wave_row_of(const OffT *ptr, uint64_t lo, uint64_t hi, uint64_t k)
{
    const uint64_t lane = threadIdx.x & (kWave - 1);
    while (hi - lo > 1) {
        const uint64_t step = (hi - lo + kWave - 1) / kWave;
        const uint64_t p = lo + lane * step;
        const bool ok = p < hi && (uint64_t)ptr[p] <= k;  // true on a prefix of the lanes, lane 0 included
        const int cnt = __popcll(__ballot(ok));
        lo += (uint64_t)(cnt - 1) * step;
        hi = min(hi, lo + step);
    }
    return lo;
}

// First index in [lo, hi) with a[i] >= key (a strictly increasing).
__device__ __forceinline__ uint64_t lower_bound_u32(const uint32_t *a, uint64_t lo, uint64_t hi, uint32_t key)
{
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

inline unsigned grid_for(uint64_t items, unsigned per_block, unsigned cap = 0x7fffffffu)
{
    uint64_t g = (items + per_block - 1) / per_block;
    if (g == 0) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// ---------------------------------------------------------------------------
// Host objects
// ---------------------------------------------------------------------------

// Growable device scratch owned by a handle.  Slot i keeps its own buffer so
// one operation can hold several at once.
struct workspace {
    static constexpr int kSlots = 26;
    void *ptr[kSlots] = {};
    size_t cap[kSlots] = {};
    hipStream_t stream = nullptr;  // work that may still read a slot runs here
    void *get(int slot, size_t bytes)
    {
        if (bytes == 0) bytes = 8;
        if (cap[slot] < bytes) {
            // Growing frees the old buffer: let queued kernels finish first.
            if (ptr[slot]) {
                SPIRA_HIP(hipStreamSynchronize(stream));
                SPIRA_HIP(hipFree(ptr[slot]));
            }
            ptr[slot] = nullptr;
            cap[slot] = 0;
            size_t want = bytes + bytes / 4;
            SPIRA_HIP(hipMalloc(&ptr[slot], want));
            cap[slot] = want;
        }
        return ptr[slot];
    }
    void release()
    {
        for (int i = 0; i < kSlots; ++i) {
            if (ptr[i]) (void)hipFree(ptr[i]);
            ptr[i] = nullptr;
            cap[i] = 0;
        }
    }
};

// Caching allocator for CSR arrays and plans.  Every lock() allocates a new
// CSR and releases the old one; on a handle all work runs on one stream, so a
// released block can be handed out again at once (stream order protects any
// kernel still reading it).  Blocks are reused when they are at least the
// requested size and at most twice it; the cache is capped and is released
// with the handle.
struct block_cache {
    struct blk {
        void *p;
        size_t bytes;
    };
    std::vector<blk> free_list;
    size_t cached = 0;
    size_t cap = size_t(24) << 30;  // bytes kept for reuse at most
    void *alloc(size_t bytes)
    {
        if (bytes == 0) bytes = 8;
        int best = -1;
        for (size_t i = 0; i < free_list.size(); ++i) {
            const size_t b = free_list[i].bytes;
            if (b >= bytes && b <= 2 * bytes && (best < 0 || b < free_list[best].bytes)) best = (int)i;
        }
        if (best >= 0) {
            void *p = free_list[best].p;
            cached -= free_list[best].bytes;
            free_list.erase(free_list.begin() + best);
            return p;
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            release_all();  // give cached memory back and retry once
            SPIRA_HIP(hipMalloc(&p, bytes));
        }
        return p;
    }
    void release(void *p, size_t bytes)
    {
        if (!p) return;
        if (bytes == 0) bytes = 8;
        if (cached + bytes > cap) {
            (void)hipFree(p);
            return;
        }
        free_list.push_back({p, bytes});
        cached += bytes;
    }
    void release_all()
    {
        for (auto &b : free_list) (void)hipFree(b.p);
        free_list.clear();
        cached = 0;
    }
};

// Workspace slot assignment (one operation at a time per handle).
enum ws_slot : int {
    WS_X = 0,        // host-x staging
    WS_Y = 1,        // host-y staging
    WS_SCAN = 2,     // scan block sums
    WS_STAGE_PTR = 3,
    WS_STAGE_COL = 4,
    WS_STAGE_VAL = 5,
    WS_DED_COL = 6,
    WS_DED_VAL = 7,
    WS_FLAGS = 8,    // lock: packed flags of staged slots, then their prefix
    WS_ROWS = 9,     // lock: deduped count per row
    WS_MISC = 10,    // small control words
    WS_MISC2 = 11,   // spmv plan: block-start prefix
    WS_QPOS = 12,    // lock: insertion point of each staged slot in its old row
    WS_BIG = 13,     // lock: rows with many staged inserts
    WS_CNT = 14,     // lock: new row lengths
    WS_PZ = 15,      // lock: prefix of dropped old zeros
    WS_MIXQ = 16,    // mixed spmv: flag byte per row for the FP64 recompute + its count
    WS_X32 = 17,     // mixed spmv: x rounded to FP32
    WS_SORT = 18,    // device staging: radix sort temporary storage
    WS_BLK = 19,     // lock: row tables of the element-parallel passes
    WS_LONG_ROWS = 20,  // lock: rows with more than 8192 staged inserts
    WS_LONG_OFF = 21,   // lock: their inserts' offsets, back to back
    WS_LONG_KEY = 22,   // lock: (list index, column) keys, unsorted then sorted
    WS_LONG_IDX = 23,   // lock: insertion index of each key, unsorted then sorted
    WS_LONG_P = 24,     // lock: prefix of the kept keys
    WS_SEL = 25,        // matrix features and the SpMV selector's scores
};

}  // namespace spira_amd

// SpMV plan cached on a device CSR (dropped by every lock/upload).
struct spira_plan {
    bool stats_valid = false;
    uint64_t max_len = 0;        // longest row
    double mean_len = 0.0;       // nnz / n_rows
    double cv_len = 0.0;         // coefficient of variation of row lengths
    uint32_t *off32 = nullptr;   // 32-bit copy of the offsets (nnz < 2^32)
    uint32_t *blk[2] = {nullptr, nullptr}; // stream row blocks, chunk 1024 / 2048
    uint32_t nblk[2] = {0, 0};
    // XSLICE: a slice-major copy of A, one column slice per XCD
    uint32_t *xs_off = nullptr;  // kSlices*n_rows+1 offsets into xs_col/xs_val
    uint32_t *xs_col = nullptr;
    double *xs_val = nullptr;
    double *xs_part = nullptr;   // kSlices*n_rows partial sums
    uint32_t *xs_blk = nullptr;  // XSTREAM row-block lists of all slices
    uint32_t *xs_blk_start = nullptr;  // kSlices+1 list starts
    uint32_t xs_nblk_max = 0;
    // XPHASE: entries grouped by (row block, column slice), rows in order
    uint64_t *ph_goff = nullptr;  // ph_nb*ph_S+1 group starts
    uint32_t *ph_rc = nullptr;    // row inside its block << 18 | column inside its slice
    double *ph_val = nullptr;
    uint32_t *ph_sync = nullptr;  // phase-pacing counters, one per XCD
    bool ph_sync_init = false;    // ph_sync zeroed (before this plan's first paced call)
    uint32_t ph_epoch = 0;        // paced calls on ph_sync since it was zeroed
    uint32_t *ph_order = nullptr; // row blocks by decreasing entry count
    float *ph_v32 = nullptr;      // the XPHASE values as FP32, BF16, split BF16 (built on demand)
    uint16_t *ph_b16 = nullptr;
    uint32_t *ph_s32 = nullptr;   // split BF16: hi in the low half, lo in the high half
    uint32_t ph_S = 0, ph_RB = 0, ph_nb = 0, ph_width = 0;
    int32_t variant = -1;        // pinned AUTO choice (-1 = none)
    int32_t auto_variant = -1;   // AUTO's own choice once resolved (learned selector, else the rule)
    int32_t auto_repro = -1;     // ... among the variants with a fixed summation order
    int32_t auto_ran = -1;       // the variant the last AUTO call ran (-1: no AUTO call since the lock)
    bool scored = false;         // sel_score holds the selector's outputs for this matrix
    float sel_score[SPIRA_SPMV_N_VARIANTS] = {};
    bool tuned = false;          // tuned_ms holds the autotuner's times (-1: did not run)
    float tuned_ms[SPIRA_SPMV_N_VARIANTS] = {};
    std::vector<std::pair<void *, size_t>> owned;  // every plan allocation
};

struct spira_dcsr {
    spira_dev *dev = nullptr;
    uint64_t n_rows = 0, n_cols = 0, nnz = 0, capacity = 0;
    bool built = false;          // false until the first lock()/upload
    bool has_zeros = false;      // CSR may hold explicit zeros (only via upload)
    uint64_t *off = nullptr;     // n_rows+1
    uint32_t *col = nullptr;     // capacity
    double *val = nullptr;       // capacity
    spira_plan plan;
    // Reduced-precision value copies (built on demand, dropped on lock).
    float *val_f32 = nullptr;
    uint16_t *val_bf16 = nullptr;     // BF16 storage
    uint16_t *val_split_hi = nullptr; // split BF16: hi = bf16(v)
    uint16_t *val_bf16_lo = nullptr;  // split BF16: lo = bf16(v - hi)
};

struct spira_dev {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    spira_amd::workspace ws;
    spira_amd::block_cache cache;
    std::string last_error;
    int arch_major = 0, arch_minor = 0;
    int n_cu = 0;
    void *noalloc_scratch = nullptr;  // this GPU's copy of the fixed scratch array (capi.hip)
    void *host_words = nullptr;       // pinned host memory for small read-backs (kHostWordsBytes)
    uint32_t phase_spin_max = 1u << 14;  // loads before an XPHASE_PACED wait gives up (SPIRA_PHASE_SPIN_MAX)
};
