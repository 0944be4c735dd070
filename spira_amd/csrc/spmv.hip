// spmv.hip -- FP64 CSR SpMV kernels for gfx950 and their per-matrix plan.
//
// Replaces the per-row SIMD dot pointer of the reference
// (kernels.h:5, dispatch.cpp:29-67, dot_avx512.cpp:7-56) called from
// serial/spmv.hpp:167-173 and parallel/algorithms/spmv.hpp:147-151.  Instead of
// one call per row, one launch covers the whole matrix:
//
//   VEC<G>    G lanes (G = 1..64) per row, strided over the row, reduced with
//             cross-lane shuffles; for rows of roughly uniform length.
//   STREAM<C> row blocks holding <= C rows+nnz (merge-path style split of the
//             rows+nnz diagonal); the block streams its values/columns
//             coalesced, stages products in LDS, then reduces each row with a
//             power-of-two team of lanes sized to the rows in the block.  Rows
//             longer than C get a block of their own reduced by all 256 lanes.
//   XSLICE<G> the columns cut into 8 slices, one per XCD (slice-major copy of
//             A), G lanes per row; partial y per slice, summed in slice order.
//   XSTREAM   the same slices, each run with the STREAM row-block kernel.
//   XPHASE    row blocks that keep their y in LDS; the blocks of one XCD walk
//             L2-sized column slices in the same order (entries grouped by
//             block and slice), so x gathers hit L2 and y is written once.
//             XPHASE_PACED also paces the blocks of an XCD to within three
//             slices.  Each (block, slice) group is sorted by x line.
//
// Values and columns are read with non-temporal loads so the x gathers keep
// L2 and the MALL.  Blocks are renumbered so that each XCD (8 on MI355X, which
// receive workgroups round-robin) processes one contiguous slice of rows,
// keeping a row neighbourhood's x lines in one L2.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace spira_amd {

namespace {

constexpr int kThreads = 256;

// Map hardware block b (dispatched to XCD b % 8) to a logical block so that
// each XCD owns one contiguous range of logical blocks.  A bijection on [0,nb).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    const uint32_t per = nb >> 3, rem = nb & 7u;
    const uint32_t x = b & 7u, idx = b >> 3;
    const uint32_t start = x * per + (x < rem ? x : rem);
    return start + idx;
}

// ---------------------------------------------------------------------------
// VEC<G>: G lanes per row
// ---------------------------------------------------------------------------
template <int G, class OffT>
__global__ __launch_bounds__(kThreads) void k_spmv_vec(uint64_t n_rows, const OffT *__restrict__ off,
                                                        const uint32_t *__restrict__ col,
                                                        const double *__restrict__ val,
                                                        const double *__restrict__ x, double *__restrict__ y)
{
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const uint64_t gid = (uint64_t)lb * kThreads + threadIdx.x;
    const uint64_t row = gid / G;
    const uint32_t lane = (uint32_t)(gid % G);
    if (row >= n_rows) return;  // whole G-groups leave together
    const uint64_t b = off[row], e = off[row + 1];
    double acc0 = 0.0, acc1 = 0.0;
    uint64_t k = b + lane;
    for (; k + G < e; k += 2 * G) {
        const uint32_t c0 = ld_stream(col + k), c1 = ld_stream(col + k + G);
        const double v0 = ld_stream(val + k), v1 = ld_stream(val + k + G);
        acc0 += v0 * x[c0];
        acc1 += v1 * x[c1];
    }
    if (k < e) acc0 += ld_stream(val + k) * x[ld_stream(col + k)];
    double acc = group_sum<G>(acc0 + acc1);
    if (lane == 0) y[row] = acc;
}

// ---------------------------------------------------------------------------
// STREAM<C>: LDS-staged row blocks
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double s_part[kThreads / kWave];
    v = group_sum<kWave>(v);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_part[wid] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / kWave; ++w) t += s_part[w];
    return t;
}

// One row block [r0, r1) of a CSR (off/col/val), y[r] written for its rows.
template <int C, class OffT>
__device__ __forceinline__ void stream_block(uint32_t r0, uint32_t r1, const OffT *__restrict__ off,
                                             const uint32_t *__restrict__ col, const double *__restrict__ val,
                                             const double *__restrict__ x, double *__restrict__ y)
{
    __shared__ double s_prod[2 * C];
    __shared__ uint32_t s_off[C + 1];
    const uint32_t nr = r1 - r0;
    const uint64_t k0 = off[r0], k1 = off[r1];

    if (nr == 1) {
        // One row (possibly far longer than C): all 256 lanes stride over it.
        double a0 = 0.0, a1 = 0.0;
        uint64_t k = k0 + threadIdx.x;
        for (; k + kThreads < k1; k += 2 * kThreads) {
            const uint32_t c0 = ld_stream(col + k), c1 = ld_stream(col + k + kThreads);
            const double v0 = ld_stream(val + k), v1 = ld_stream(val + k + kThreads);
            a0 += v0 * x[c0];
            a1 += v1 * x[c1];
        }
        if (k < k1) a0 += ld_stream(val + k) * x[ld_stream(col + k)];
        const double t = block_sum(a0 + a1);
        if (threadIdx.x == 0) y[r0] = t;
        return;
    }

    // Phase 1: coalesced stream of this block's values/columns; products to LDS.
    const uint32_t nk = (uint32_t)(k1 - k0);  // < 2C by construction of the plan
    constexpr int IT = 2 * C / kThreads;
    uint32_t cc[IT];
    double vv[IT];
#pragma unroll
    for (int t = 0; t < IT; ++t) {
        const uint32_t i = threadIdx.x + t * kThreads;
        const bool ok = i < nk;
        cc[t] = ok ? ld_stream(col + k0 + i) : 0u;
        vv[t] = ok ? ld_stream(val + k0 + i) : 0.0;
    }
    for (uint32_t j = threadIdx.x; j <= nr; j += kThreads) s_off[j] = (uint32_t)(off[r0 + j] - k0);
#pragma unroll
    for (int t = 0; t < IT; ++t) {
        const uint32_t i = threadIdx.x + t * kThreads;
        if (i < nk) s_prod[i] = vv[t] * x[cc[t]];
    }
    __syncthreads();

    // Phase 2: a team of T lanes per row, T = largest power of two with
    // T * nr <= 256 (at most one wavefront).
    int T = 1;
    while (T < kWave && (uint32_t)(2 * T) * nr <= (uint32_t)kThreads) T <<= 1;
    const uint32_t teams = kThreads / T;
    const uint32_t lane = threadIdx.x & (T - 1);
    for (uint32_t r = threadIdx.x / T; r < nr; r += teams) {
        const uint32_t a = s_off[r], e = s_off[r + 1];
        double acc = 0.0;
        for (uint32_t k = a + lane; k < e; k += T) acc += s_prod[k];
        acc = group_sum_rt(acc, T);
        if (lane == 0) y[r0 + r] = acc;
    }
}

template <int C, class OffT>
__global__ __launch_bounds__(kThreads) void k_spmv_stream(const uint32_t *__restrict__ blk_row,
                                                           const OffT *__restrict__ off,
                                                           const uint32_t *__restrict__ col,
                                                           const double *__restrict__ val,
                                                           const double *__restrict__ x, double *__restrict__ y)
{
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    stream_block<C, OffT>(blk_row[b], blk_row[b + 1], off, col, val, x, y);
}

// ---------------------------------------------------------------------------
// XSLICE<G>: the columns are cut into kSlices slices, one per XCD, and the
// plan keeps a slice-major copy of A (slice s holds every row's entries whose
// column falls in slice s, rows in order).  Hardware block b runs on XCD
// b % 8, so it takes slice b % 8: each XCD gathers only from its 1/8 of x,
// which stays in that XCD's L2 instead of missing to the MALL on almost every
// gather.  Each slice writes a partial y; k_xslice_sum adds the 8 partials in
// slice order.
// ---------------------------------------------------------------------------
constexpr uint32_t kSlices = 8;

template <int G>
__global__ __launch_bounds__(kThreads) void k_spmv_xslice(uint64_t n_rows, const uint32_t *__restrict__ xoff,
                                                           const uint32_t *__restrict__ xcol,
                                                           const double *__restrict__ xval,
                                                           const double *__restrict__ x, double *__restrict__ part)
{
    const uint32_t s = blockIdx.x % kSlices;
    const uint64_t rb = blockIdx.x / kSlices;
    const uint64_t row = rb * (kThreads / G) + threadIdx.x / G;
    const uint32_t lane = threadIdx.x % G;
    if (row >= n_rows) return;
    const uint64_t base = (uint64_t)s * n_rows + row;
    const uint32_t b = xoff[base], e = xoff[base + 1];
    double acc0 = 0.0, acc1 = 0.0;
    uint32_t k = b + lane;
    for (; k + G < e; k += 2 * G) {
        const uint32_t c0 = ld_stream(xcol + k), c1 = ld_stream(xcol + k + G);
        const double v0 = ld_stream(xval + k), v1 = ld_stream(xval + k + G);
        acc0 += v0 * x[c0];
        acc1 += v1 * x[c1];
    }
    if (k < e) acc0 += ld_stream(xval + k) * x[ld_stream(xcol + k)];
    const double acc = group_sum<G>(acc0 + acc1);
    if (lane == 0) part[base] = acc;
}

// XSTREAM<C>: the same slices, each processed with the LDS row-block kernel.
// blk holds every slice's row-block list back to back (each list ends with
// n_rows); blk_start[s] is where slice s's list begins.
template <int C>
__global__ __launch_bounds__(kThreads) void k_spmv_xstream(const uint32_t *__restrict__ blk,
                                                            const uint32_t *__restrict__ blk_start, uint64_t n_rows,
                                                            const uint32_t *__restrict__ xoff,
                                                            const uint32_t *__restrict__ xcol,
                                                            const double *__restrict__ xval,
                                                            const double *__restrict__ x, double *__restrict__ part)
{
    const uint32_t s = blockIdx.x % kSlices;
    const uint32_t j = blockIdx.x / kSlices;
    const uint32_t a = blk_start[s];
    const uint32_t nb = blk_start[s + 1] - a - 1;
    if (j >= nb) return;
    stream_block<C, uint32_t>(blk[a + j], blk[a + j + 1], xoff + (uint64_t)s * n_rows, xcol, xval, x,
                              part + (uint64_t)s * n_rows);
}

__global__ __launch_bounds__(kThreads) void k_xslice_sum(uint64_t n_rows, const double *__restrict__ part,
                                                          double *__restrict__ y)
{
    const uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= n_rows) return;
    double acc = ld_stream(part + r);
#pragma unroll
    for (uint32_t s = 1; s < kSlices; ++s) acc += ld_stream(part + (uint64_t)s * n_rows + r);
    y[r] = acc;
}

// Plan: entries of row r per slice (columns are sorted, so one pass).
__global__ __launch_bounds__(kThreads) void k_xslice_count(uint64_t n_rows, uint64_t width,
                                                            const uint64_t *__restrict__ off,
                                                            const uint32_t *__restrict__ col,
                                                            uint64_t *__restrict__ cnt)
{
    const uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= n_rows) return;
    uint64_t c[kSlices] = {};
    for (uint64_t k = off[r]; k < off[r + 1]; ++k) {
        const uint32_t s = (uint32_t)(col[k] / width);
#pragma unroll
        for (uint32_t t = 0; t < kSlices; ++t) c[t] += (t == s) ? 1u : 0u;
    }
#pragma unroll
    for (uint32_t t = 0; t < kSlices; ++t) cnt[(uint64_t)t * n_rows + r] = c[t];
}

// Plan: copy row r's entries into their slices at the scanned positions.
__global__ __launch_bounds__(kThreads) void k_xslice_scatter(uint64_t n_rows, uint64_t width,
                                                              const uint64_t *__restrict__ off,
                                                              const uint32_t *__restrict__ col,
                                                              const double *__restrict__ val,
                                                              const uint64_t *__restrict__ pos,
                                                              uint32_t *__restrict__ xcol, double *__restrict__ xval)
{
    const uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= n_rows) return;
    uint64_t w[kSlices];
#pragma unroll
    for (uint32_t t = 0; t < kSlices; ++t) w[t] = pos[(uint64_t)t * n_rows + r];
    for (uint64_t k = off[r]; k < off[r + 1]; ++k) {
        const uint32_t c = col[k];
        const uint32_t s = (uint32_t)(c / width);
        uint64_t d = 0;
#pragma unroll
        for (uint32_t t = 0; t < kSlices; ++t) {
            if (t == s) {
                d = w[t];
                w[t] = d + 1;
            }
        }
        xcol[d] = c;
        xval[d] = val[k];
    }
}

__global__ void k_u64_to_u32(const uint64_t *__restrict__ a, uint32_t *__restrict__ b, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (uint32_t)a[i];
}

// ---------------------------------------------------------------------------
// XPHASE: a block owns a run of RB <= kPhaseRows rows and keeps their y in
// LDS for the whole call.  The columns are cut into S slices of at most
// kPhaseSliceBytes of x; the plan groups the entries by (block, slice), rows
// in order, each entry carrying its row inside the block and its column inside
// the slice in one u32 (14 + 18 bits): 12 bytes per entry, as in CSR.  A block
// walks the S slices in an order that depends only on its XCD (block b runs
// on XCD b % 8), so the blocks of one XCD gather from the same slice at the
// same time and that slice stays in the XCD's L2 while A streams past it.
// Products are added into LDS with native FP64 atomics (the order of the
// additions within a row can differ from call to call; the rounding of the
// row sum is that of any summation order), and y is written once at the end:
// no partial-y vectors and no second kernel.
// ---------------------------------------------------------------------------
constexpr uint32_t kPhaseRows = 16384;                 // rows per block: 128 KiB of FP64 y in LDS
constexpr uint64_t kPhaseSliceBytes = 2ull << 20;       // x bytes per slice: half of an XCD's L2
constexpr uint32_t kPhaseMaxSlices = 256;
// An entry's row inside its block and column inside its slice share one u32.
constexpr uint32_t kPhaseColBits = 18;
static_assert(kPhaseRows <= (1u << (32 - kPhaseColBits)), "local rows must fit the high bits");
constexpr int kPhaseThreads = 1024;
// XPHASE: 6 entries per thread per step, phases not paced (banded matrices:
// a block meets few slices).  XPHASE_PACED: 20 entries per thread per step
// (a 2 MiB slice of a scattered matrix is one step for most blocks) and paced
// phases (scattered columns: every block meets every slice).

// Phase pacing.  A block may start phase t only after every block of its XCD
// has finished phase t - 3, so an XCD's blocks gather from at most three
// slices at a time (they drift apart otherwise, and the XCD's L2 then holds
// many slices' worth of x).  sync[c * kPhaseSyncStride] counts finished (block,
// phase) pairs of XCD c, mod 2^32, and runs on across calls: a call adds
// nbx * passes * S to it, so the kernel offsets its targets by epoch (the
// calls since the counters were zeroed) times that.  The wait only paces
// and never decides a result: it gives up after a bounded number of sleeps
// (spin_max loads), and the host passes sync only when every block of the
// grid can be resident at once (one block per CU).  A block whose wait gives
// up stops pacing for the rest of the call (other work on the GPU may hold
// CUs, so some blocks of its XCD are not running): one timeout per block at
// most, never one per phase.
constexpr uint32_t kPhaseSyncStride = 16;  // counters 64 bytes apart

// true when the wait gave up (the same value in every thread of the block)
__device__ __forceinline__ bool phase_wait(const uint32_t *ctr, uint32_t target, uint32_t spin_max)
{
    __shared__ uint32_t s_gave_up;
    if (threadIdx.x == 0) {
        uint32_t i = 0;
        for (; i < spin_max; ++i) {
            // counts and targets are taken mod 2^32 (the counters run on across calls)
            if ((int32_t)(__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) >= 0) break;
            __builtin_amdgcn_s_sleep(4);
        }
        s_gave_up = i == spin_max ? 1u : 0u;
    }
    __syncthreads();
    return s_gave_up != 0;
}

__global__ void k_phase_sync_zero(uint32_t *sync)
{
    if (threadIdx.x < 8 * kPhaseSyncStride) sync[threadIdx.x] = 0u;
}

// Values of the XPHASE copy: stored FP64, FP32, BF16 or split BF16 (hi + lo,
// summed exactly in FP64).  load() reads the stored form and wide() turns it
// into FP64 where the product is formed, after the step's gathers are issued.
struct ph_f64 {
    const double *v;
    using raw = double;
    __device__ __forceinline__ raw load(uint64_t q) const { return ld_stream(v + q); }
    __device__ __forceinline__ static double wide(raw r) { return r; }
};
struct ph_f32 {
    const float *v;
    using raw = float;
    __device__ __forceinline__ raw load(uint64_t q) const { return ld_stream(v + q); }
    __device__ __forceinline__ static double wide(raw r) { return (double)r; }
};
struct ph_bf16 {
    const uint16_t *v;
    using raw = uint16_t;
    __device__ __forceinline__ raw load(uint64_t q) const { return ld_stream(v + q); }
    __device__ __forceinline__ static double wide(raw r) { return (double)bf16_to_f32(r); }
};
struct ph_split {
    const uint32_t *v;     // one word per entry: one load instead of two
    using raw = uint32_t;  // hi in the low half, lo in the high half
    __device__ __forceinline__ raw load(uint64_t q) const { return ld_stream(v + q); }
    __device__ __forceinline__ static double wide(raw r)
    {
        return (double)bf16_to_f32((uint16_t)r) + (double)bf16_to_f32((uint16_t)(r >> 16));
    }
};

// x as used by the products: RX rounds an FP64 x to FP32 first (the mixed
// path's x, rounded where it is gathered instead of in a pass of its own).
template <bool RX, class XT>
__device__ __forceinline__ double x_used(XT v)
{
    if constexpr (RX) return (double)(float)v;
    else return (double)v;
}

// One step of U entries per thread from entry k on, each slot guarded
// against the group's end e1.
template <int U, bool RX, class Vals, class XT>
__device__ __forceinline__ void xphase_guarded_step(uint64_t k, uint64_t e1, uint32_t cbase,
                                                    const uint32_t *__restrict__ prc, Vals pval,
                                                    const XT *__restrict__ x, double *s_y)
{
    uint32_t rr[U], cc[U];
    typename Vals::raw vr[U];
    double xx[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t q = k + (uint64_t)u * kPhaseThreads;
        ok[u] = q < e1;
        const uint32_t rc = ok[u] ? ld_stream(prc + q) : 0u;  // a masked lane reads x[cbase], adds nothing
        rr[u] = rc >> kPhaseColBits;
        cc[u] = cbase + (rc & ((1u << kPhaseColBits) - 1u));
        vr[u] = ok[u] ? pval.load(q) : typename Vals::raw{};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) xx[u] = x_used<RX>(x[cc[u]]);
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (ok[u])
            __hip_atomic_fetch_add(&s_y[rr[u]], Vals::wide(vr[u]) * xx[u], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Products are formed and summed in FP64 whatever the stored precision (a
// FP32 value times a FP32 x is exact in FP64), and y is written as YT.
// Block b of the grid takes row blocks b, b + grid, b + 2 grid, ... (one
// "pass" each; with pacing the grid is one block per CU, so matrices with more
// row blocks than CUs are paced too).  Every block runs every pass, so the
// XCD's pacing count is the same in each; a block past the last row block
// runs its phases empty and still counts them.
template <int U, class Vals, class XT, class YT, bool RX = false>
__global__ __launch_bounds__(kPhaseThreads) void k_spmv_xphase(uint64_t n_rows, uint32_t RB, uint32_t S,
                                                                uint32_t nb, const uint32_t *__restrict__ order,
                                                                const uint64_t *__restrict__ goff,
                                                                uint32_t width, const uint32_t *__restrict__ prc,
                                                                Vals pval, const XT *__restrict__ x,
                                                                YT *__restrict__ y, uint32_t *__restrict__ sync,
                                                                uint32_t spin_max, uint32_t full_steps,
                                                                uint32_t epoch)
{
    __shared__ double s_y[kPhaseRows];
    __shared__ uint64_t s_goff[kPhaseMaxSlices + 1];  // this block's group starts
    const uint32_t xcd = blockIdx.x % 8;
    const uint32_t per = S / 8;  // S is a multiple of 8; XCD c starts at slice c * S / 8
    const uint32_t first = xcd * per;
    const uint32_t nbx = (gridDim.x - xcd + 7) / 8;  // blocks of this XCD
    uint32_t *ctr = sync ? sync + xcd * kPhaseSyncStride : nullptr;
    bool pace = ctr != nullptr;  // counting continues after a give-up; waiting does not
    const uint32_t passes = (nb + gridDim.x - 1) / gridDim.x;
    // every call on these counters adds nbx * passes * S to this XCD's count:
    // the count before this call (mod 2^32)
    const uint32_t base = epoch * (nbx * passes * S);
    if (pace) {  // a count behind base means the counters missed a call: pace nothing
        __shared__ uint32_t s_behind;
        if (threadIdx.x == 0)
            s_behind = (int32_t)(__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - base) < 0;
        __syncthreads();
        if (s_behind) pace = false;
    }
    for (uint32_t pass = 0; pass < passes; ++pass) {
    const uint32_t slot = blockIdx.x + pass * gridDim.x;
    const bool live = slot < nb;
    const uint32_t b = live ? order[slot] : 0u;  // row blocks heaviest first
    const uint64_t r0 = (uint64_t)b * RB;
    const uint32_t nr = live ? (uint32_t)min((uint64_t)RB, n_rows - r0) : 0u;
    for (uint32_t i = threadIdx.x; i < nr; i += kPhaseThreads) s_y[i] = 0.0;
    for (uint32_t i = threadIdx.x; i <= S; i += kPhaseThreads) s_goff[i] = live ? goff[(uint64_t)b * S + i] : 0ull;
    __syncthreads();
    for (uint32_t t = 0; t < S; ++t) {
        const uint32_t T = pass * S + t;  // phase index across passes
        if (pace && T >= 3 && phase_wait(ctr, base + nbx * (T - 2), spin_max)) pace = false;
        const uint32_t s = first + t < S ? first + t : first + t - S;
        const uint64_t e0 = s_goff[s], e1 = s_goff[s + 1];
        const uint32_t cbase = s * width;
        // Reduced value types with full_steps: full steps (every slot of every
        // thread inside the group) load without branches, so every load of
        // the step is in flight before the first wait, and one guarded step
        // takes the group's remainder.  Otherwise (FP64 values, or
        // full_steps clear) every step guards its slots.  Each form is kept
        // where it measured fastest; the compiler schedules the guarded step
        // differently inside a loop and as a single step, hence both.
        constexpr uint64_t kStep = (uint64_t)U * kPhaseThreads;
        if constexpr (std::is_same_v<Vals, ph_f64>) {
            for (uint64_t k = e0 + threadIdx.x; k < e1; k += kStep) xphase_guarded_step<U, RX>(k, e1, cbase, prc, pval, x, s_y);
        } else if (full_steps) {
            uint64_t kb = e0;  // the block's step base
            for (; kb + kStep <= e1; kb += kStep) {
                const uint64_t k = kb + threadIdx.x;
                uint32_t rc[U];
                typename Vals::raw vr[U];
                XT xv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    rc[u] = ld_stream(prc + k + (uint64_t)u * kPhaseThreads);
                    vr[u] = pval.load(k + (uint64_t)u * kPhaseThreads);
                }
                __builtin_amdgcn_sched_barrier(0);  // all stream loads issued before the first wait
#pragma unroll
                for (int u = 0; u < U; ++u) xv[u] = x[cbase + (rc[u] & ((1u << kPhaseColBits) - 1u))];
                __builtin_amdgcn_sched_barrier(0);  // all gathers issued before the first add
#pragma unroll
                for (int u = 0; u < U; ++u)
                    __hip_atomic_fetch_add(&s_y[rc[u] >> kPhaseColBits], Vals::wide(vr[u]) * x_used<RX>(xv[u]),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (kb < e1) xphase_guarded_step<U, RX>(kb + threadIdx.x, e1, cbase, prc, pval, x, s_y);
        } else {
            for (uint64_t k = e0 + threadIdx.x; k < e1; k += kStep) xphase_guarded_step<U, RX>(k, e1, cbase, prc, pval, x, s_y);
        }
        if (ctr) {
            __syncthreads();  // the whole block is done with phase t
            if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nr; i += kPhaseThreads) y[r0 + i] = (YT)s_y[i];
    __syncthreads();  // s_y and s_goff are reused by the next pass
    }
}

// XPHASE plan, built element-parallel in three passes (no per-row state, so
// its cost and memory grow with nnz alone).  Entries are binned into
// sub-groups (row block, column slice, sub-slice of 2^16 columns); a slice's
// sub-groups are adjacent, so a (block, slice) group is their union.
//   bin<count>    a workgroup takes 8192 consecutive entries, finds each
//                 one's row from an LDS copy of its rows' offsets, counts its
//                 sub-groups in LDS and adds them to the sub-group counters;
//                 an exclusive scan gives the sub-group starts.
//   bin<scatter>  the same chunks reserve their share of each sub-group with
//                 one atomic per sub-group and write (row in block << 18 |
//                 column in slice, value) there.
//   lines         one workgroup per sub-group orders its entries by x line
//                 (16 columns, 128 bytes): line counts and their prefix in
//                 LDS, then windows of consecutive lines holding at most
//                 kLineStage entries are staged in LDS in line order and
//                 written out whole, so every store is coalesced.  The 64
//                 lanes of a wavefront then sweep x in order and lanes on one
//                 line share its request.  (The order inside a line is not
//                 fixed; the FP64 LDS adds are not order-fixed either.)
constexpr int kPlanThreads = 1024;
constexpr int kPlanItems = 8;
constexpr uint32_t kPlanChunk = kPlanThreads * kPlanItems;
constexpr uint32_t kPlanRowCache = 2048;
constexpr uint32_t kPlanLocalGroups = 4096;
constexpr uint32_t kLineCols = 16;
constexpr uint32_t kSubBits = 16;                              // columns per sub-slice: 2^16
constexpr uint32_t kSubLines = (1u << kSubBits) / kLineCols;  // 4096 lines
constexpr uint32_t kWinLines = kSubLines;
constexpr uint32_t kLineStage = 4096;                          // entries staged in LDS per window
constexpr int kLineItems = 8;                                  // entries per thread held in registers

// Row table of the bin passes: tab[b] = row owning entry b * kPlanChunk, and
// tab[chunks] = row owning the last entry, so chunk b's rows are
// [tab[b], tab[b + 1]] without a search.  One thread per row.
__global__ void k_xphase_chunk_rows(const uint64_t *__restrict__ off, uint64_t n_rows, uint64_t nnz,
                                    uint64_t *__restrict__ tab)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    const uint64_t e0 = off[r], e1 = off[r + 1];
    if (e0 == e1) return;
    for (uint64_t b = (e0 + kPlanChunk - 1) / kPlanChunk; b * kPlanChunk < e1; ++b) tab[b] = r;
    if (e1 == nnz) tab[(nnz + kPlanChunk - 1) / kPlanChunk] = r;
}

// In-place running maximum of s[0..n) (n <= kPlanChunk) by a kPlanThreads
// block, kPlanItems consecutive values per thread.
__device__ __forceinline__ void block_max_scan_u16(uint16_t *s, uint32_t n)
{
    __shared__ uint32_t s_w[kPlanThreads / kWave];
    const uint32_t a = threadIdx.x * kPlanItems;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    uint32_t v[kPlanItems];
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < kPlanItems; ++i) {
        v[i] = a + i < n ? s[a + i] : 0u;
        m = max(m, v[i]);
        v[i] = m;
    }
    uint32_t inc = m;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(inc, d, kWave);
        if (lane >= d) inc = max(inc, t);
    }
    if (lane == kWave - 1) s_w[wid] = inc;
    __syncthreads();
    uint32_t ex = __shfl_up(inc, 1, kWave);
    if (lane == 0) ex = 0;
    for (int w = 0; w < wid; ++w) ex = max(ex, s_w[w]);
#pragma unroll
    for (int i = 0; i < kPlanItems; ++i)
        if (a + i < n) s[a + i] = (uint16_t)max(v[i], ex);
    __syncthreads();
}

// Entries of one chunk (kPlanChunk of them) binned by (row block, slice,
// sub-slice).  The row of each entry: every non-empty row of the chunk marks
// its first entry with its index in the chunk's row range and a running
// maximum spreads the marks over the row's entries (up to kPlanRowCache rows;
// a chunk touching more searches the offsets in global memory).  Divisions by
// the row-block size and the slice width are multiply-shifts (udiv32).
template <bool SCATTER>
__global__ __launch_bounds__(kPlanThreads) void k_xphase_bin(uint64_t nnz, uint64_t n_rows, uint32_t RB, udiv32 rb_div,
                                                              uint32_t S, uint32_t SUB, uint32_t width,
                                                              udiv32 width_div, const uint64_t *__restrict__ off,
                                                              const uint32_t *__restrict__ col,
                                                              const double *__restrict__ val,
                                                              const uint64_t *__restrict__ tab,
                                                              unsigned long long *__restrict__ gctr,
                                                              uint32_t *__restrict__ trc, double *__restrict__ tval)
{
    static_assert(kPlanRowCache <= 65536, "16-bit local rows");
    __shared__ uint16_t s_rid[kPlanChunk];   // local row of each entry of the chunk
    __shared__ uint32_t s_cnt[kPlanLocalGroups];
    __shared__ uint64_t s_base[SCATTER ? kPlanLocalGroups : 1];
    const uint64_t first = (uint64_t)blockIdx.x * kPlanChunk;
    if (first >= nnz) return;
    const uint64_t last = min(first + kPlanChunk, nnz) - 1;
    const uint32_t nq = (uint32_t)(last - first + 1);
    // the entries first: their loads are in flight while the row marks fill
    uint32_t cc[kPlanItems];
    double v[kPlanItems];
#pragma unroll
    for (int it = 0; it < kPlanItems; ++it) {
        const uint64_t k = first + (uint64_t)it * kPlanThreads + threadIdx.x;
        cc[it] = k <= last ? ld_stream(col + k) : 0u;
        if (SCATTER) v[it] = k <= last ? ld_stream(val + k) : 0.0;
    }
    const uint64_t rlo = tab[blockIdx.x], rhi = tab[blockIdx.x + 1];
    const uint64_t nr = rhi - rlo + 1;
    const bool cached = nr <= kPlanRowCache;
    const uint64_t gps = (uint64_t)S * SUB;  // sub-groups per block
    const uint32_t b0 = udiv((uint32_t)rlo, rb_div);
    const uint64_t nlg = (uint64_t)(udiv((uint32_t)rhi, rb_div) - b0 + 1) * gps;  // sub-groups this chunk can touch
    const bool local = nlg <= kPlanLocalGroups;
    if (cached)
        for (uint32_t j = threadIdx.x; j < nq; j += kPlanThreads) s_rid[j] = 0;  // row rlo owns entry `first`
    if (local)
        for (uint32_t j = threadIdx.x; j < nlg; j += kPlanThreads) s_cnt[j] = 0;
    __syncthreads();
    if (cached) {
        // rows after rlo start after `first`; a non-empty one marks its first entry
        for (uint32_t j = threadIdx.x + 1; j < nr; j += kPlanThreads) {
            const uint64_t a = off[rlo + j];
            if (a <= last && off[rlo + j + 1] > a) s_rid[a - first] = (uint16_t)j;
        }
        __syncthreads();
        block_max_scan_u16(s_rid, nq);
    }

    uint32_t lg[kPlanItems], rc[kPlanItems];
#pragma unroll
    for (int it = 0; it < kPlanItems; ++it) {
        const uint32_t q = (uint32_t)it * kPlanThreads + threadIdx.x;
        lg[it] = 0xffffffffu;
        if (q >= nq) continue;
        const uint32_t c = cc[it];
        const uint32_t r = cached ? (uint32_t)rlo + s_rid[q] : (uint32_t)row_of(off, rlo, rhi + 1, first + q);
        const uint32_t b = udiv(r, rb_div);
        const uint32_t sl = udiv(c, width_div);
        const uint32_t cl = c - sl * width;
        rc[it] = ((r - b * RB) << kPhaseColBits) | cl;
        lg[it] = (uint32_t)((uint64_t)(b - b0) * gps + (uint64_t)sl * SUB + (cl >> kSubBits));
        if (!local) {
            const uint64_t g = (uint64_t)b0 * gps + lg[it];
            if (SCATTER) {
                const uint64_t p = atomicAdd(gctr + g, 1ull);
                trc[p] = rc[it];
                tval[p] = v[it];
            } else {
                atomicAdd(gctr + g, 1ull);
            }
        } else {
            atomicAdd(&s_cnt[lg[it]], 1u);
        }
    }
    if (!local) return;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nlg; j += kPlanThreads) {
        const uint32_t n = s_cnt[j];
        if (n == 0) continue;
        if (SCATTER) {
            s_base[j] = atomicAdd(gctr + (uint64_t)b0 * gps + j, (unsigned long long)n);
            s_cnt[j] = 0;  // now the running rank inside the reserved run
        } else {
            atomicAdd(gctr + (uint64_t)b0 * gps + j, (unsigned long long)n);
        }
    }
    if (!SCATTER) return;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kPlanItems; ++it) {
        if (lg[it] == 0xffffffffu) continue;
        const uint64_t p = s_base[lg[it]] + atomicAdd(&s_cnt[lg[it]], 1u);
        trc[p] = rc[it];
        tval[p] = v[it];
    }
}

// In-place exclusive scan of c[0..n) (n <= kWinLines) by a 1024-thread block
// (each thread owns a run of 4); c[n] receives the total.
__device__ __forceinline__ void block_scan_lines(uint32_t *c, uint32_t n)
{
    __shared__ uint32_t s_w[kPlanThreads / kWave];
    constexpr uint32_t per = kWinLines / kPlanThreads;
    const uint32_t a = threadIdx.x * per;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < per; ++i) sum += (a + i < n) ? c[a + i] : 0u;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    uint32_t inc = sum;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(inc, d, kWave);
        if (lane >= d) inc += t;
    }
    if (lane == kWave - 1) s_w[wid] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int w = 0; w < kPlanThreads / kWave; ++w) {
        if (w < wid) base += s_w[w];
        tot += s_w[w];
    }
    uint32_t run = base + inc - sum;
#pragma unroll
    for (uint32_t i = 0; i < per; ++i) {
        if (a + i < n) {
            const uint32_t t = c[a + i];
            c[a + i] = run;
            run += t;
        }
    }
    if (threadIdx.x == 0) c[n] = tot;
    __syncthreads();
}

// One workgroup per sub-group at a time.  Sub-groups of up to kLineItems
// entries per thread are held in registers: counted per line, then placed in
// windows of consecutive lines of at most kLineStage entries into LDS in line
// order and written out whole.  Larger ones read their entries again for each
// window; a single line holding more than a window is written in place (its
// own range; the order inside a line is free).
__global__ __launch_bounds__(kPlanThreads) void k_xphase_lines(const uint64_t *__restrict__ gsub, uint64_t n_sub,
                                                                uint32_t SUB, uint32_t width,
                                                                const uint32_t *__restrict__ trc,
                                                                const double *__restrict__ tval,
                                                                uint32_t *__restrict__ prc, double *__restrict__ pval)
{
    __shared__ uint32_t s_pre[kSubLines + 1];  // line counts, then their exclusive prefix
    __shared__ uint32_t s_cur[kSubLines];      // cursors of the current window's lines
    __shared__ uint32_t s_rc[kLineStage];
    __shared__ double s_val[kLineStage];
    constexpr uint32_t kColMask = (1u << kPhaseColBits) - 1u;
    for (uint64_t g = blockIdx.x; g < n_sub; g += gridDim.x) {
        const uint64_t e0 = gsub[g], e1 = gsub[g + 1];
        if (e0 == e1) continue;  // the same for every thread of the block
        const uint32_t n = (uint32_t)(e1 - e0);  // < 2^32 (checked by the host)
        const uint32_t c0 = (uint32_t)(g % SUB) << kSubBits;  // first column of the sub-slice in its slice
        const uint32_t nl = (min(width, c0 + (1u << kSubBits)) - c0 + kLineCols - 1) / kLineCols;
        const bool reg = n <= kLineItems * kPlanThreads;
        uint32_t rw[kLineItems];
        double rv[kLineItems];
        if (reg) {
#pragma unroll
            for (int it = 0; it < kLineItems; ++it) {
                const uint32_t i = threadIdx.x + it * kPlanThreads;
                rw[it] = i < n ? ld_stream(trc + e0 + i) : 0u;
                rv[it] = i < n ? ld_stream(tval + e0 + i) : 0.0;
            }
        }
        for (uint32_t j = threadIdx.x; j < nl; j += kPlanThreads) s_pre[j] = 0;
        __syncthreads();
        if (reg) {
#pragma unroll
            for (int it = 0; it < kLineItems; ++it)
                if (threadIdx.x + it * kPlanThreads < n) atomicAdd(&s_pre[((rw[it] & kColMask) - c0) / kLineCols], 1u);
        } else {
            for (uint32_t i = threadIdx.x; i < n; i += kPlanThreads)
                atomicAdd(&s_pre[((trc[e0 + i] & kColMask) - c0) / kLineCols], 1u);
        }
        __syncthreads();
        block_scan_lines(s_pre, nl);
        uint32_t ws = 0;  // windows of consecutive lines; identical in every thread
        while (ws < nl) {
            uint32_t lo = ws + 1, hi = nl;  // last we in (ws, nl] with s_pre[we] - s_pre[ws] <= kLineStage
            const uint32_t lim = s_pre[ws] + kLineStage;
            if (s_pre[lo] > lim) {
                hi = lo;  // one line holds more than a window
            } else {
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= lim) lo = mid;
                    else hi = mid - 1;
                }
                hi = lo;
            }
            const uint32_t we = hi;
            const uint32_t wbase = s_pre[ws], wn = s_pre[we] - wbase;
            const bool staged = wn <= kLineStage;
            for (uint32_t l = ws + threadIdx.x; l < we; l += kPlanThreads) s_cur[l - ws] = s_pre[l] - wbase;
            __syncthreads();
            auto place = [&](uint32_t w, double v) {
                const uint32_t l = ((w & kColMask) - c0) / kLineCols;
                if (l < ws || l >= we) return;
                const uint32_t d = atomicAdd(&s_cur[l - ws], 1u);
                if (staged) {
                    s_rc[d] = w;
                    s_val[d] = v;
                } else {
                    prc[e0 + wbase + d] = w;
                    pval[e0 + wbase + d] = v;
                }
            };
            if (reg) {
#pragma unroll
                for (int it = 0; it < kLineItems; ++it)
                    if (threadIdx.x + it * kPlanThreads < n) place(rw[it], rv[it]);
            } else {
                for (uint32_t i = threadIdx.x; i < n; i += kPlanThreads) {
                    const uint32_t w = trc[e0 + i];
                    const uint32_t l = ((w & kColMask) - c0) / kLineCols;
                    if (l >= ws && l < we) place(w, tval[e0 + i]);
                }
            }
            __syncthreads();
            if (staged) {
                for (uint32_t d = threadIdx.x; d < wn; d += kPlanThreads) {
                    prc[e0 + wbase + d] = s_rc[d];
                    pval[e0 + wbase + d] = s_val[d];
                }
                __syncthreads();
            }
            ws = we;
        }
    }
}

// Slice-group starts of the kernel: every SUB-th sub-group start.
__global__ void k_xphase_goff(const uint64_t *__restrict__ gsub, uint64_t groups, uint32_t SUB,
                              uint64_t *__restrict__ goff)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g <= groups) goff[g] = gsub[g * SUB];
}

// ---------------------------------------------------------------------------
// Plan kernels
// ---------------------------------------------------------------------------
__global__ void k_off32(const uint64_t *__restrict__ off, uint32_t *__restrict__ o32, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o32[i] = (uint32_t)off[i];
}

// Per-block max and sum of squares of the row lengths (one pass), then one
// block folds the per-block results.
__global__ __launch_bounds__(kThreads) void k_row_stats(const uint64_t *__restrict__ off, uint64_t n_rows,
                                                         uint64_t *__restrict__ bmax, double *__restrict__ bsq)
{
    __shared__ uint64_t s_mx[kThreads / kWave];
    __shared__ double s_sq[kThreads / kWave];
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    uint64_t mx = 0;
    if (i < n_rows) mx = off[i + 1] - off[i];
    double sq = (double)mx * (double)mx;
    for (int s = 32; s > 0; s >>= 1) {
        const uint64_t o = __shfl_xor(mx, s, 64);
        mx = o > mx ? o : mx;
        sq += __shfl_xor(sq, s, 64);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_mx[wid] = mx;
        s_sq[wid] = sq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t m = 0;
        double q = 0.0;
        for (int w = 0; w < kThreads / kWave; ++w) {
            m = s_mx[w] > m ? s_mx[w] : m;
            q += s_sq[w];
        }
        bmax[blockIdx.x] = m;
        bsq[blockIdx.x] = q;
    }
}

__global__ __launch_bounds__(kThreads) void k_fold_stats(const uint64_t *__restrict__ bmax,
                                                          const double *__restrict__ bsq, uint64_t nb,
                                                          uint64_t *__restrict__ out_max, double *__restrict__ out_sq)
{
    __shared__ uint64_t s_mx[kThreads];
    __shared__ double s_sq[kThreads];
    uint64_t m = 0;
    double q = 0.0;
    for (uint64_t i = threadIdx.x; i < nb; i += kThreads) {
        m = bmax[i] > m ? bmax[i] : m;
        q += bsq[i];
    }
    s_mx[threadIdx.x] = m;
    s_sq[threadIdx.x] = q;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_mx[threadIdx.x] = s_mx[threadIdx.x + s] > s_mx[threadIdx.x] ? s_mx[threadIdx.x + s] : s_mx[threadIdx.x];
            s_sq[threadIdx.x] += s_sq[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out_max = s_mx[0];
        *out_sq = s_sq[0];
    }
}

__device__ __forceinline__ bool stream_block_start(const uint64_t *off, uint64_t r, uint32_t C)
{
    if (r == 0) return true;
    const uint64_t len = off[r + 1] - off[r];
    const uint64_t lenp = off[r] - off[r - 1];
    if (len > C || lenp > C) return true;
    return (off[r] + r) / C != (off[r - 1] + r - 1) / C;
}

__global__ void k_plan_mark(const uint64_t *__restrict__ off, uint64_t n_rows, uint32_t C,
                            uint64_t *__restrict__ mark)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n_rows) mark[r] = stream_block_start(off, r, C) ? 1u : 0u;
}

__global__ void k_plan_scatter(const uint64_t *__restrict__ off, uint64_t n_rows, uint32_t C,
                               const uint64_t *__restrict__ pos, uint32_t *__restrict__ blk)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n_rows && stream_block_start(off, r, C)) blk[pos[r]] = (uint32_t)r;
    if (r == 0) blk[pos[n_rows]] = (uint32_t)n_rows;
}

// XPHASE's shape for A: rows per block (all CUs busy on mid-sized matrices,
// at most kPhaseRows), blocks, column slices (a multiple of 8, each at most
// kPhaseSliceBytes of x, at most kPhaseMaxSlices) and columns per slice.
void xphase_geometry(const spira_dcsr *A, uint64_t *rb, uint64_t *nb, uint64_t *S, uint64_t *width)
{
    const uint64_t n = A->n_rows;
    *rb = std::min<uint64_t>(kPhaseRows, std::max<uint64_t>(1, (n + 255) / 256));
    *nb = (n + *rb - 1) / *rb;
    const uint64_t xbytes = A->n_cols * sizeof(double);
    *S = std::min<uint64_t>(kPhaseMaxSlices,
                            8 * std::max<uint64_t>(1, (xbytes + 8 * kPhaseSliceBytes - 1) / (8 * kPhaseSliceBytes)));
    *width = std::max<uint64_t>(1, (A->n_cols + *S - 1) / *S);
}

template <int G, class OffT>
void launch_vec(spira_dcsr *A, const OffT *off, const double *x, double *y)
{
    const uint64_t threads = A->n_rows * (uint64_t)G;
    const unsigned grid = grid_for(threads, kThreads);
    k_spmv_vec<G, OffT><<<grid, kThreads, 0, A->dev->stream>>>(A->n_rows, off, A->col, A->val, x, y);
}

template <class OffT>
void launch_vec_g(int g, spira_dcsr *A, const OffT *off, const double *x, double *y)
{
    switch (g) {
    case 1: launch_vec<1>(A, off, x, y); break;
    case 2: launch_vec<2>(A, off, x, y); break;
    case 4: launch_vec<4>(A, off, x, y); break;
    case 8: launch_vec<8>(A, off, x, y); break;
    case 16: launch_vec<16>(A, off, x, y); break;
    case 32: launch_vec<32>(A, off, x, y); break;
    default: launch_vec<64>(A, off, x, y); break;
    }
}

template <int C, class OffT>
void launch_stream(spira_dcsr *A, const uint32_t *blk, uint32_t nblk, const OffT *off, const double *x,
                   double *y)
{
    k_spmv_stream<C, OffT><<<nblk, kThreads, 0, A->dev->stream>>>(blk, off, A->col, A->val, x, y);
}

// Row-length statistics and the 32-bit offsets.  for_auto: the call resolves
// AUTO, so the selector's feature pass (which yields the same statistics)
// runs instead of the plain one, with its scores, under one read-back.
void ensure_stats(spira_dcsr *A, bool for_auto = false)
{
    spira_plan &P = A->plan;
    if (P.stats_valid) return;
    spira_dev *dev = A->dev;
    if (A->nnz < (1ull << 32) && !P.off32) {  // queued first: it runs while the statistics come back
        P.off32 = (uint32_t *)plan_alloc(A, (A->n_rows + 1) * sizeof(uint32_t));
        k_off32<<<grid_for(A->n_rows + 1, kThreads), kThreads, 0, dev->stream>>>(A->off, P.off32, A->n_rows + 1);
        SPIRA_HIP(hipGetLastError());
    }
    if (A->n_rows > 0) {
        uint64_t max_len = 0;
        double sumsq = 0.0;
        if (!(for_auto && spmv_stats_and_scores(A, &max_len, &sumsq))) {
            const uint64_t nb = grid_for(A->n_rows, kThreads);
            char *w = (char *)dev->ws.get(WS_MISC, nb * 16 + 64);
            uint64_t *bmax = (uint64_t *)w;
            double *bsq = (double *)(w + nb * 8);
            uint64_t *omax = (uint64_t *)(w + nb * 16);
            double *osq = (double *)(w + nb * 16 + 8);
            k_row_stats<<<(unsigned)nb, kThreads, 0, dev->stream>>>(A->off, A->n_rows, bmax, bsq);
            SPIRA_HIP(hipGetLastError());
            k_fold_stats<<<1, kThreads, 0, dev->stream>>>(bmax, bsq, nb, omax, osq);
            SPIRA_HIP(hipGetLastError());
            uint64_t h[2];
            read_back(dev, h, omax, sizeof(h));
            max_len = h[0];
            std::memcpy(&sumsq, &h[1], sizeof(double));
        }
        P.max_len = max_len;
        P.mean_len = (double)A->nnz / (double)A->n_rows;
        double var = sumsq / (double)A->n_rows - P.mean_len * P.mean_len;
        P.cv_len = P.mean_len > 0 ? std::sqrt(var > 0 ? var : 0.0) / P.mean_len : 0.0;
    }
    P.stats_valid = true;
}

void ensure_blocks(spira_dcsr *A, int which)
{
    spira_plan &P = A->plan;
    if (P.blk[which]) return;
    spira_dev *dev = A->dev;
    const uint32_t C = which == 0 ? 1024u : 2048u;
    const uint64_t n = A->n_rows;
    uint64_t *pos = (uint64_t *)dev->ws.get(WS_MISC2, (n + 1) * sizeof(uint64_t));
    k_plan_mark<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(A->off, n, C, pos);
    SPIRA_HIP(hipGetLastError());
    scan_excl_u64(dev, pos, pos, n);
    const uint64_t nb = read_scalar(dev, pos + n);
    P.blk[which] = (uint32_t *)plan_alloc(A, (nb + 1) * sizeof(uint32_t));
    k_plan_scatter<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(A->off, n, C, pos, P.blk[which]);
    SPIRA_HIP(hipGetLastError());
    P.nblk[which] = (uint32_t)nb;
}

// Slice-major copy of A for the XSLICE kernels (one-time per locked matrix).
void ensure_xslice(spira_dcsr *A)
{
    spira_plan &P = A->plan;
    if (P.xs_off) return;
    if (A->nnz >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "spmv: XSLICE needs nnz < 2^32");
    spira_dev *dev = A->dev;
    const uint64_t n = A->n_rows;
    const uint64_t width = (A->n_cols + kSlices - 1) / kSlices > 0 ? (A->n_cols + kSlices - 1) / kSlices : 1;
    const uint64_t N = kSlices * n;
    uint64_t *pos = (uint64_t *)dev->ws.get(WS_MISC2, (N + 1) * sizeof(uint64_t));
    k_xslice_count<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(n, width, A->off, A->col, pos);
    SPIRA_HIP(hipGetLastError());
    scan_excl_u64(dev, pos, pos, N);
    // (plan memory is recorded on the plan: a failure part-way leaves it to
    // spmv_free_plan, and xs_off is set last so the plan is never half-used)
    uint32_t *xs_off = nullptr;
    P.xs_col = (uint32_t *)plan_alloc(A, (A->nnz ? A->nnz : 1) * sizeof(uint32_t));
    P.xs_val = (double *)plan_alloc(A, (A->nnz ? A->nnz : 1) * sizeof(double));
    P.xs_part = (double *)plan_alloc(A, (N ? N : 1) * sizeof(double));
    xs_off = (uint32_t *)plan_alloc(A, (N + 1) * sizeof(uint32_t));
    k_xslice_scatter<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(n, width, A->off, A->col, A->val, pos,
                                                                          P.xs_col, P.xs_val);
    SPIRA_HIP(hipGetLastError());
    k_u64_to_u32<<<grid_for(N + 1, kThreads), kThreads, 0, dev->stream>>>(pos, xs_off, N + 1);
    SPIRA_HIP(hipGetLastError());

    // Row-block lists of every slice for XSTREAM (chunk 1024), back to back:
    // one pass to size them, one to fill them.
    constexpr uint32_t C = 1024;
    uint64_t *mark = (uint64_t *)dev->ws.get(WS_CNT, (n + 1) * sizeof(uint64_t));
    std::vector<uint32_t> start(kSlices + 1, 0);
    uint64_t mx = 0;
    for (uint32_t s = 0; s < kSlices; ++s) {
        k_plan_mark<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(pos + (uint64_t)s * n, n, C, mark);
        SPIRA_HIP(hipGetLastError());
        scan_excl_u64(dev, mark, mark, n);
        const uint64_t nb = read_scalar(dev, mark + n);
        start[s + 1] = start[s] + (uint32_t)nb + 1;
        mx = nb > mx ? nb : mx;
    }
    P.xs_blk = (uint32_t *)plan_alloc(A, (uint64_t)start[kSlices] * sizeof(uint32_t));
    P.xs_blk_start = (uint32_t *)plan_alloc(A, (kSlices + 1) * sizeof(uint32_t));
    for (uint32_t s = 0; s < kSlices; ++s) {
        const uint64_t *so = pos + (uint64_t)s * n;
        k_plan_mark<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(so, n, C, mark);
        SPIRA_HIP(hipGetLastError());
        scan_excl_u64(dev, mark, mark, n);
        k_plan_scatter<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(so, n, C, mark, P.xs_blk + start[s]);
        SPIRA_HIP(hipGetLastError());
    }
    SPIRA_HIP(hipMemcpyAsync(P.xs_blk_start, start.data(), (kSlices + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                             dev->stream));
    dev_sync(dev);  // pos and mark live in the workspace: finish before they can be reused
    P.xs_nblk_max = (uint32_t)mx;
    P.xs_off = xs_off;  // the slice plan is complete
}

// (block, slice)-grouped copy of A for XPHASE (one-time per locked matrix).
void ensure_xphase(spira_dcsr *A)
{
    spira_plan &P = A->plan;
    if (P.ph_goff) return;
    spira_dev *dev = A->dev;
    const uint64_t n = A->n_rows;
    uint64_t rb, nb, S, width;
    xphase_geometry(A, &rb, &nb, &S, &width);
    if (width > (1ull << kPhaseColBits)) fail(SPIRA_ERR_UNSUPPORTED, "spmv: too many columns for XPHASE");
    if (nb >= (1ull << 31)) fail(SPIRA_ERR_UNSUPPORTED, "spmv: matrix too large for XPHASE");
    const uint64_t G = nb * S;
    const uint64_t SUB = (width + (1ull << kSubBits) - 1) >> kSubBits;
    const uint64_t GS = G * SUB;  // sub-groups
    const uint64_t nz = A->nnz ? A->nnz : 1;
    const uint64_t chunks = (A->nnz + kPlanChunk - 1) / kPlanChunk;
    if (chunks >= (1ull << 31)) fail(SPIRA_ERR_UNSUPPORTED, "spmv: matrix too large for XPHASE");
    P.ph_rc = (uint32_t *)plan_alloc(A, nz * sizeof(uint32_t));
    P.ph_val = (double *)plan_alloc(A, nz * sizeof(double));
    uint64_t *goff = (uint64_t *)plan_alloc(A, (G + 1) * sizeof(uint64_t));
    P.ph_sync = (uint32_t *)plan_alloc(A, 8 * kPhaseSyncStride * sizeof(uint32_t));
    // sub-group counters, then starts (GS + 1), then cursors (GS)
    uint64_t *gsub = (uint64_t *)dev->ws.get(WS_MISC2, (2 * GS + 1) * sizeof(uint64_t));
    unsigned long long *cur = (unsigned long long *)(gsub + GS + 1);
    struct temp {  // released on every exit (stream order protects queued work)
        spira_dev *d;
        void *p;
        size_t bytes;
        ~temp() { dev_release(d, p, bytes); }
    } t_rc{dev, dev_alloc(dev, nz * sizeof(uint32_t)), nz * sizeof(uint32_t)},
        t_val{dev, dev_alloc(dev, nz * sizeof(double)), nz * sizeof(double)};
    uint32_t *trc = (uint32_t *)t_rc.p;
    double *tval = (double *)t_val.p;
    hipStream_t st = dev->stream;
    SPIRA_HIP(hipMemsetAsync(gsub, 0, (GS + 1) * sizeof(uint64_t), st));
    uint64_t *tab = (uint64_t *)dev->ws.get(WS_CNT, (chunks + 1) * sizeof(uint64_t));
    if (chunks > 0) {
        k_xphase_chunk_rows<<<grid_for(n, kThreads), kThreads, 0, st>>>(A->off, n, A->nnz, tab);
        SPIRA_HIP(hipGetLastError());
        k_xphase_bin<false><<<(unsigned)chunks, kPlanThreads, 0, st>>>(
            A->nnz, n, (uint32_t)rb, make_udiv32((uint32_t)rb), (uint32_t)S, (uint32_t)SUB, (uint32_t)width,
            make_udiv32((uint32_t)width), A->off, A->col, A->val, tab, (unsigned long long *)gsub, nullptr, nullptr);
        SPIRA_HIP(hipGetLastError());
    }
    scan_excl_u64(dev, gsub, gsub, GS);  // sub-group starts; gsub[GS] = nnz
    k_xphase_goff<<<grid_for(G + 1, kThreads), kThreads, 0, st>>>(gsub, G, (uint32_t)SUB, goff);
    SPIRA_HIP(hipGetLastError());
    if (chunks > 0) {
        SPIRA_HIP(hipMemcpyAsync(cur, gsub, GS * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        k_xphase_bin<true><<<(unsigned)chunks, kPlanThreads, 0, st>>>(
            A->nnz, n, (uint32_t)rb, make_udiv32((uint32_t)rb), (uint32_t)S, (uint32_t)SUB, (uint32_t)width,
            make_udiv32((uint32_t)width), A->off, A->col, A->val, tab, cur, trc, tval);
        SPIRA_HIP(hipGetLastError());
        if (A->nnz >= (1ull << 32)) {
            // a sub-group's entry count is held in 32 bits by the line pass
            uint64_t mx = 0;
            std::vector<uint64_t> h(GS + 1);
            read_back(dev, h.data(), gsub, (GS + 1) * sizeof(uint64_t));
            for (uint64_t g = 0; g < GS; ++g) mx = std::max(mx, h[g + 1] - h[g]);
            if (mx >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "spmv: an XPHASE group holds 2^32 entries or more");
        }
        const uint64_t grid = std::min<uint64_t>(GS, 2u * (uint64_t)std::max(dev->n_cu, 1));
        k_xphase_lines<<<(unsigned)grid, kPlanThreads, 0, st>>>(gsub, GS, (uint32_t)SUB, (uint32_t)width, trc, tval,
                                                                P.ph_rc, P.ph_val);
        SPIRA_HIP(hipGetLastError());
    }
    // row blocks in decreasing order of entries (the longest start first, so
    // the light ones fill the tail)
    {
        std::vector<uint64_t> h(G + 1);
        read_back(dev, h.data(), goff, (G + 1) * sizeof(uint64_t));
        std::vector<uint32_t> ord(nb);
        for (uint32_t i = 0; i < nb; ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
            return h[(uint64_t)(a + 1) * S] - h[(uint64_t)a * S] > h[(uint64_t)(b + 1) * S] - h[(uint64_t)b * S];
        });
        P.ph_order = (uint32_t *)plan_alloc(A, nb * sizeof(uint32_t));
        SPIRA_HIP(hipMemcpyAsync(P.ph_order, ord.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
    dev_sync(dev);  // the cursors live in the workspace: finish before they can be reused
    P.ph_S = (uint32_t)S;
    P.ph_width = (uint32_t)width;
    P.ph_RB = (uint32_t)rb;
    P.ph_nb = (uint32_t)nb;
    P.ph_goff = goff;  // the plan is complete
}

// PACE: pace the phases (worth it when every block has entries in every
// slice; a banded matrix gives each block one or two slices, and pacing would
// then make the blocks of an XCD take turns).  Only when the whole grid is
// resident at once (one block per CU).
// ctr: the 8 * kPhaseSyncStride pacing counters (zeroed here by a kernel).
// Branch-free full steps for the reduced value types (the FP64 kernels
// ignore the flag).  The paced kernel (20 entries per thread per step) always
// takes them.  The unpaced one (6 per step) does unless the longest row holds
// 1/32 or more of an average row block's entries: that row's entries then
// meet in most wavefronts of the column-sorted groups (two or more lanes of a
// wavefront adding into one LDS word), and such matrices ran faster with
// every step guarded (FP32 values, rows of up to 50,000 entries: 868 against
// 911 us, while rows clipped at 1,024 ran 406 against 369 us and the
// Laplacian 64.4 against 39.3 us).  Without row statistics: guarded.
bool xphase_full_steps(const spira_dcsr *A, bool paced)
{
    const spira_plan &P = A->plan;
    if (paced) return true;
    if (!P.stats_valid || P.ph_nb == 0) return false;
    return P.max_len * 32 < A->nnz / P.ph_nb + 1;
}

// The pacing counters of a plan (ctr == the plan's ph_sync) are zeroed before
// its first paced call only: each call adds a fixed count per XCD, so the
// kernel offsets its targets by the plan's call number (epoch).  Counters
// shared with other kernels (the allocation-free scratch) are zeroed before
// every call (epoch 0).
template <int U, bool PACE, class Vals = ph_f64, class XT = double, class YT = double, bool RX = false>
void launch_xphase(spira_dcsr *A, const XT *x, YT *y, uint32_t *ctr, Vals vals = Vals{})
{
    spira_plan &P = A->plan;
    if constexpr (std::is_same_v<Vals, ph_f64>) vals.v = P.ph_val;
    uint32_t *sync = nullptr;
    uint32_t grid = P.ph_nb, epoch = 0;
    if (PACE && A->dev->n_cu > 0) {
        // one block per CU, all resident: more row blocks are taken in passes
        grid = std::min<uint32_t>(P.ph_nb, (uint32_t)A->dev->n_cu);
        sync = ctr;
        const uint64_t per_call = (uint64_t)((grid + 7) / 8) * ((P.ph_nb + grid - 1) / grid) * P.ph_S;
        const bool own = ctr == P.ph_sync && per_call < (1ull << 30);  // targets compared mod 2^32
        if (!own || !P.ph_sync_init) {
            k_phase_sync_zero<<<1, 8 * kPhaseSyncStride, 0, A->dev->stream>>>(sync);
            P.ph_epoch = 0;
            P.ph_sync_init = own;
        }
        if (own) epoch = P.ph_epoch++;
    }
    k_spmv_xphase<U, Vals, XT, YT, RX><<<grid, kPhaseThreads, 0, A->dev->stream>>>(
        A->n_rows, P.ph_RB, P.ph_S, P.ph_nb, P.ph_order, P.ph_goff, P.ph_width, P.ph_rc, vals, x, y, sync,
        A->dev->phase_spin_max, xphase_full_steps(A, PACE) ? 1u : 0u, epoch);
}

__global__ void k_ph_to_f32(const double *__restrict__ v, float *__restrict__ o, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (float)v[i];
}

__global__ void k_ph_to_bf16(const double *__restrict__ v, uint16_t *__restrict__ o, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = f32_to_bf16((float)v[i]);
}

__global__ void k_ph_to_split(const double *__restrict__ v, uint32_t *__restrict__ o, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t h = f32_to_bf16((float)v[i]);
    const uint16_t l = f32_to_bf16((float)(v[i] - (double)bf16_to_f32(h)));
    o[i] = (uint32_t)h | ((uint32_t)l << 16);
}

// The XPHASE values in reduced precision (built once per locked matrix from
// the plan's FP64 values, in the plan's order).
void ensure_xphase_vals(spira_dcsr *A, int storage)
{
    ensure_xphase(A);
    spira_plan &P = A->plan;
    const uint64_t n = A->nnz ? A->nnz : 1;
    const unsigned g = grid_for(n, kThreads);
    hipStream_t st = A->dev->stream;
    if (storage == SPIRA_VAL_F32 && !P.ph_v32) {
        float *o = (float *)plan_alloc(A, n * sizeof(float));
        k_ph_to_f32<<<g, kThreads, 0, st>>>(P.ph_val, o, A->nnz);
        P.ph_v32 = o;
    } else if (storage == SPIRA_VAL_BF16 && !P.ph_b16) {
        uint16_t *o = (uint16_t *)plan_alloc(A, n * sizeof(uint16_t));
        k_ph_to_bf16<<<g, kThreads, 0, st>>>(P.ph_val, o, A->nnz);
        P.ph_b16 = o;
    } else if (storage == SPIRA_VAL_SPLIT_BF16 && !P.ph_s32) {
        uint32_t *o = (uint32_t *)plan_alloc(A, n * sizeof(uint32_t));
        k_ph_to_split<<<g, kThreads, 0, st>>>(P.ph_val, o, A->nnz);
        P.ph_s32 = o;
    }
    SPIRA_HIP(hipGetLastError());
}

// part: kSlices * n_rows partial sums (every one written before it is read).
void launch_xstream(spira_dcsr *A, const double *x, double *y, double *part)
{
    const spira_plan &P = A->plan;
    const uint64_t grid = (uint64_t)P.xs_nblk_max * kSlices;
    if (grid > 0x7fffffffull) fail(SPIRA_ERR_UNSUPPORTED, "spmv: matrix too large for XSTREAM");
    if (grid > 0)
        k_spmv_xstream<1024><<<(unsigned)grid, kThreads, 0, A->dev->stream>>>(
            P.xs_blk, P.xs_blk_start, A->n_rows, P.xs_off, P.xs_col, P.xs_val, x, part);
    k_xslice_sum<<<grid_for(A->n_rows, kThreads), kThreads, 0, A->dev->stream>>>(A->n_rows, part, y);
}

template <int G>
void launch_xslice(spira_dcsr *A, const double *x, double *y, double *part)
{
    const spira_plan &P = A->plan;
    const uint64_t nrb = (A->n_rows * (uint64_t)G + kThreads - 1) / kThreads;
    const uint64_t grid = nrb * kSlices;
    if (grid > 0x7fffffffull) fail(SPIRA_ERR_UNSUPPORTED, "spmv: matrix too large for XSLICE");
    k_spmv_xslice<G><<<(unsigned)grid, kThreads, 0, A->dev->stream>>>(A->n_rows, P.xs_off, P.xs_col, P.xs_val, x,
                                                                       part);
    k_xslice_sum<<<grid_for(A->n_rows, kThreads), kThreads, 0, A->dev->stream>>>(A->n_rows, part, y);
}

}  // namespace

void spmv_free_plan(spira_dcsr *A)
{
    spira_plan &P = A->plan;
    for (auto &o : P.owned) dev_release(A->dev, o.first, o.second);
    P = spira_plan{};  // a lock changes the matrix: the pinned variant is dropped too
}

void spmv_build_plan(spira_dcsr *A) { ensure_stats(A); }

void spmv_stream_blocks(spira_dcsr *A, int which, const uint32_t **blk, uint32_t *nblk)
{
    ensure_blocks(A, which);
    *blk = A->plan.blk[which];
    *nblk = A->plan.nblk[which];
}

int spmv_auto_choice(const spira_dcsr *A, bool repro)
{
    const spira_plan &P = A->plan;
    if (P.variant >= 0 && (!repro || spmv_order_fixed(P.variant))) return P.variant;
    return repro ? P.auto_repro : P.auto_variant;
}

int spmv_resolve_auto(spira_dcsr *A, bool repro)
{
    spira_plan &P = A->plan;
    ensure_stats(A);
    if (P.variant >= 0 && (!repro || spmv_order_fixed(P.variant))) return P.variant;
    int32_t &slot = repro ? P.auto_repro : P.auto_variant;
    if (slot >= 0) return slot;
    int best = -1;
    if (P.tuned) {  // the autotuner's times, where it has run on this matrix
        for (int v = 0; v < SPIRA_SPMV_N_VARIANTS; ++v)
            if (P.tuned_ms[v] >= 0 && (!repro || spmv_order_fixed(v)) && (best < 0 || P.tuned_ms[v] < P.tuned_ms[best]))
                best = v;
    }
    if (best < 0 && spmv_selector_scores(A)) {
        for (int v = 0; v < SPIRA_SPMV_N_VARIANTS; ++v)
            if (spmv_variant_fits(A, v) && (!repro || spmv_order_fixed(v)) &&
                (best < 0 || P.sel_score[v] < P.sel_score[best]))
                best = v;
    }
    slot = best >= 0 ? best : spmv_rule_variant(A);
    return slot;
}

bool spmv_variant_fits(const spira_dcsr *A, int v)
{
    if (v < 0 || v >= SPIRA_SPMV_N_VARIANTS) return false;
    if (v >= SPIRA_SPMV_XSLICE1 && v <= SPIRA_SPMV_XSTREAM) {
        if (A->nnz >= (1ull << 32)) return false;
        const uint64_t G = v == SPIRA_SPMV_XSLICE1 ? 1 : v == SPIRA_SPMV_XSLICE2 ? 2 : 4;
        return (A->n_rows * G + kThreads - 1) / kThreads * kSlices <= 0x7fffffffull;
    }
    if (v == SPIRA_SPMV_XPHASE || v == SPIRA_SPMV_XPHASE_PACED) {
        uint64_t rb, nb, S, width;
        xphase_geometry(A, &rb, &nb, &S, &width);
        return width <= (1ull << kPhaseColBits) && nb < (1ull << 31) &&
               (A->nnz + kPlanChunk - 1) / kPlanChunk < (1ull << 31);
    }
    return true;
}

// The rule never picks an XPHASE variant (so it also serves reproducible AUTO).
int spmv_rule_variant(const spira_dcsr *A)
{
    const spira_plan &P = A->plan;
    const double mean = P.mean_len;
    // Irregular lengths (power-law rows) or very short rows: LDS row blocks.
    if ((double)P.max_len > 4.0 * (mean > 8.0 ? mean : 8.0) || P.cv_len > 0.5) return SPIRA_SPMV_STREAM2K;
    if (mean <= 6.0) return SPIRA_SPMV_STREAM1K;
    int g = 2;
    while (g < 64 && g < mean) g <<= 1;
    switch (g) {
    case 2: return SPIRA_SPMV_VEC2;
    case 4: return SPIRA_SPMV_VEC4;
    case 8: return SPIRA_SPMV_VEC8;
    case 16: return SPIRA_SPMV_VEC16;
    case 32: return SPIRA_SPMV_VEC32;
    default: return SPIRA_SPMV_VEC64;
    }
}

namespace {

// Launch `variant` on a plan whose parts it needs are in place; the scratch
// of the XPHASE_PACED counters and the XSLICE/XSTREAM partial sums is passed
// in (plan memory, or the caller's fixed array).
void launch_planned(spira_dcsr *A, const double *x, double *y, int variant, uint32_t *ctr, double *part)
{
    const bool o32 = A->plan.off32 != nullptr;
    if (variant <= SPIRA_SPMV_VEC64) {
        const int g = 1 << variant;
        if (o32) launch_vec_g<uint32_t>(g, A, A->plan.off32, x, y);
        else launch_vec_g<uint64_t>(g, A, A->off, x, y);
    } else if (variant == SPIRA_SPMV_STREAM1K || variant == SPIRA_SPMV_STREAM2K) {
        const int which = variant == SPIRA_SPMV_STREAM1K ? 0 : 1;
        const uint32_t *blk = A->plan.blk[which];
        const uint32_t nb = A->plan.nblk[which];
        if (which == 0) {
            if (o32) launch_stream<1024, uint32_t>(A, blk, nb, A->plan.off32, x, y);
            else launch_stream<1024, uint64_t>(A, blk, nb, A->off, x, y);
        } else {
            if (o32) launch_stream<2048, uint32_t>(A, blk, nb, A->plan.off32, x, y);
            else launch_stream<2048, uint64_t>(A, blk, nb, A->off, x, y);
        }
    } else if (variant == SPIRA_SPMV_XSLICE1) {
        launch_xslice<1>(A, x, y, part);
    } else if (variant == SPIRA_SPMV_XSLICE2) {
        launch_xslice<2>(A, x, y, part);
    } else if (variant == SPIRA_SPMV_XSLICE4) {
        launch_xslice<4>(A, x, y, part);
    } else if (variant == SPIRA_SPMV_XSTREAM) {
        launch_xstream(A, x, y, part);
    } else if (variant == SPIRA_SPMV_XPHASE) {
        launch_xphase<6, false>(A, x, y, ctr);
    } else if (variant == SPIRA_SPMV_XPHASE_PACED) {
        launch_xphase<20, true>(A, x, y, ctr);
    } else {
        fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: unknown kernel variant");
    }
    SPIRA_HIP(hipGetLastError());
}

}  // namespace

bool spmv_plan_ready(const spira_dcsr *A, int variant)
{
    const spira_plan &P = A->plan;
    if (!P.stats_valid || (A->nnz < (1ull << 32) && !P.off32)) return false;
    if (variant < 0) variant = spmv_auto_choice(A);
    if (variant < 0) return false;
    if (variant <= SPIRA_SPMV_VEC64) return true;
    if (variant == SPIRA_SPMV_STREAM1K) return P.blk[0] != nullptr;
    if (variant == SPIRA_SPMV_STREAM2K) return P.blk[1] != nullptr;
    if (variant >= SPIRA_SPMV_XSLICE1 && variant <= SPIRA_SPMV_XSTREAM) return P.xs_off != nullptr;
    return P.ph_goff != nullptr;
}

uint64_t spmv_scratch_bytes(const spira_dcsr *A, int variant)
{
    if (variant < 0) variant = spmv_auto_choice(A);
    if (variant == SPIRA_SPMV_XPHASE_PACED) return 8 * kPhaseSyncStride * sizeof(uint32_t);
    if (variant >= SPIRA_SPMV_XSLICE1 && variant <= SPIRA_SPMV_XSTREAM) return kSlices * A->n_rows * sizeof(double);
    return 0;
}

void spmv_launch_planned(spira_dcsr *A, const double *x, double *y, int variant, void *scratch)
{
    if (A->n_rows == 0) return;
    if (variant < 0) variant = spmv_auto_choice(A);
    if (variant < 0) fail(SPIRA_ERR_LOGIC, "spmv: AUTO's choice is not resolved for this matrix");
    launch_planned(A, x, y, variant, (uint32_t *)scratch, (double *)scratch);
}

namespace {

// XT = float: x rounded to FP32 beforehand; XT = double: rounded where gathered.
template <class XT>
void xphase_reduced(spira_dcsr *A, int storage, const XT *x, double *y, bool paced)
{
    if (A->n_rows == 0) return;
    ensure_stats(A);
    ensure_xphase_vals(A, storage);
    const spira_plan &P = A->plan;
    uint32_t *ctr = P.ph_sync;
    constexpr bool RX = std::is_same_v<XT, double>;
    if (storage == SPIRA_VAL_F32) {
        if (paced) launch_xphase<20, true, ph_f32, XT, double, RX>(A, x, y, ctr, ph_f32{P.ph_v32});
        else launch_xphase<6, false, ph_f32, XT, double, RX>(A, x, y, ctr, ph_f32{P.ph_v32});
    } else if (storage == SPIRA_VAL_BF16) {
        if (paced) launch_xphase<20, true, ph_bf16, XT, double, RX>(A, x, y, ctr, ph_bf16{P.ph_b16});
        else launch_xphase<6, false, ph_bf16, XT, double, RX>(A, x, y, ctr, ph_bf16{P.ph_b16});
    } else {
        if (paced) launch_xphase<20, true, ph_split, XT, double, RX>(A, x, y, ctr, ph_split{P.ph_s32});
        else launch_xphase<6, false, ph_split, XT, double, RX>(A, x, y, ctr, ph_split{P.ph_s32});
    }
    SPIRA_HIP(hipGetLastError());
}

}  // namespace

void spmv_xphase_reduced(spira_dcsr *A, int storage, const float *x, double *y, bool paced)
{
    xphase_reduced(A, storage, x, y, paced);
}

void spmv_xphase_reduced(spira_dcsr *A, int storage, const double *x, double *y, bool paced)
{
    xphase_reduced(A, storage, x, y, paced);
}

void spmv_xphase_f32(spira_dcsr *A, const float *x, float *y, bool paced)
{
    if (A->n_rows == 0) return;
    ensure_stats(A);
    ensure_xphase_vals(A, SPIRA_VAL_F32);
    const spira_plan &P = A->plan;
    if (paced) launch_xphase<20, true>(A, x, y, P.ph_sync, ph_f32{P.ph_v32});
    else launch_xphase<6, false>(A, x, y, P.ph_sync, ph_f32{P.ph_v32});
    SPIRA_HIP(hipGetLastError());
}

namespace {

// The fastest variant that needs no copy of A (VEC: the offsets only; STREAM:
// a row-block list from one pass over the offsets), by the selector's scores
// where they exist, else by the row-length rule.
int spmv_light_variant(const spira_dcsr *A)
{
    const spira_plan &P = A->plan;
    if (!P.scored) return spmv_rule_variant(A);
    int best = SPIRA_SPMV_VEC1;
    for (int v = SPIRA_SPMV_VEC1; v <= SPIRA_SPMV_STREAM2K; ++v)
        if (P.sel_score[v] < P.sel_score[best]) best = v;
    return best;
}

}  // namespace

void spmv_launch_f64(spira_dcsr *A, const double *x, double *y, int variant, bool repro)
{
    if (repro && variant >= 0 && !spmv_order_fixed(variant))
        fail(SPIRA_ERR_INVALID_ARGUMENT,
             "spmv: XPHASE variants add in a varying order; drop SPIRA_REPRODUCIBLE or choose another variant");
    if (A->n_rows == 0) return;
    if (A->n_rows >= (1ull << 32) - 1) fail(SPIRA_ERR_UNSUPPORTED, "spmv: n_rows must be < 2^32 - 1");
    ensure_stats(A, variant < 0);
    if (variant < 0) {
        variant = spmv_resolve_auto(A, repro);
        // A choice that copies A into its own layout (XSLICE, XSTREAM, XPHASE:
        // several passes over A, worth ~5-10 calls) is built on the second AUTO
        // call after a lock; the first runs the fastest plan-free variant, so
        // a matrix multiplied once after each lock never pays for the copy.
        // Reproducible calls keep one variant from the first call on.
        if (!repro && variant >= SPIRA_SPMV_XSLICE1 && A->plan.auto_ran < 0 && !spmv_plan_ready(A, variant))
            variant = spmv_light_variant(A);
        A->plan.auto_ran = variant;
    }
    if (variant == SPIRA_SPMV_STREAM1K || variant == SPIRA_SPMV_STREAM2K)
        ensure_blocks(A, variant == SPIRA_SPMV_STREAM1K ? 0 : 1);
    else if (variant >= SPIRA_SPMV_XSLICE1 && variant <= SPIRA_SPMV_XSTREAM) ensure_xslice(A);
    else if (variant == SPIRA_SPMV_XPHASE || variant == SPIRA_SPMV_XPHASE_PACED) ensure_xphase(A);
    else if (variant > SPIRA_SPMV_VEC64) fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: unknown kernel variant");
    launch_planned(A, x, y, variant, A->plan.ph_sync, A->plan.xs_part);
}

}  // namespace spira_amd
