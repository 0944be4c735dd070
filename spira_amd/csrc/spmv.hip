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
//
// Values and columns are read with non-temporal loads so the x gathers keep
// L2 and the MALL.  Blocks are renumbered so that each XCD (8 on MI355X, which
// receive workgroups round-robin) processes one contiguous slice of rows,
// keeping a row neighbourhood's x lines in one L2.
#include "internal.hpp"

#include <cmath>
#include <cstring>

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

template <int C, class OffT>
__global__ __launch_bounds__(kThreads) void k_spmv_stream(const uint32_t *__restrict__ blk_row,
                                                           const OffT *__restrict__ off,
                                                           const uint32_t *__restrict__ col,
                                                           const double *__restrict__ val,
                                                           const double *__restrict__ x, double *__restrict__ y)
{
    __shared__ double s_prod[2 * C];
    __shared__ uint32_t s_off[C + 1];
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t r0 = blk_row[b], r1 = blk_row[b + 1];
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

// ---------------------------------------------------------------------------
// Plan kernels
// ---------------------------------------------------------------------------
__global__ void k_off32(const uint64_t *__restrict__ off, uint32_t *__restrict__ o32, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o32[i] = (uint32_t)off[i];
}

__global__ __launch_bounds__(kThreads) void k_row_stats(const uint64_t *__restrict__ off, uint64_t n_rows,
                                                         unsigned long long *__restrict__ out /* [max, sumsq] */)
{
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    uint64_t len = 0;
    if (i < n_rows) len = off[i + 1] - off[i];
    // wave reductions, then one atomic per wave
    uint64_t mx = len;
    double sq = (double)len * (double)len;
    for (int s = 32; s > 0; s >>= 1) {
        uint64_t o = __shfl_xor(mx, s, 64);
        mx = o > mx ? o : mx;
        sq += __shfl_xor(sq, s, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out[0], (unsigned long long)mx);
        atomicAdd(reinterpret_cast<double *>(&out[1]), sq);
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

void ensure_stats(spira_dcsr *A)
{
    spira_plan &P = A->plan;
    if (P.stats_valid) return;
    spira_dev *dev = A->dev;
    if (A->n_rows > 0) {
        unsigned long long *d = (unsigned long long *)dev->ws.get(WS_MISC, 2 * sizeof(unsigned long long));
        SPIRA_HIP(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), dev->stream));
        k_row_stats<<<grid_for(A->n_rows, kThreads), kThreads, 0, dev->stream>>>(A->off, A->n_rows, d);
        SPIRA_HIP(hipGetLastError());
        unsigned long long h[2];
        SPIRA_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, dev->stream));
        dev_sync(dev);
        double sumsq;
        std::memcpy(&sumsq, &h[1], sizeof(double));
        P.max_len = h[0];
        P.mean_len = (double)A->nnz / (double)A->n_rows;
        double var = sumsq / (double)A->n_rows - P.mean_len * P.mean_len;
        P.cv_len = P.mean_len > 0 ? std::sqrt(var > 0 ? var : 0.0) / P.mean_len : 0.0;
    }
    if (A->nnz < (1ull << 32) && !P.off32) {
        SPIRA_HIP(hipMalloc(&P.off32, (A->n_rows + 1) * sizeof(uint32_t)));
        k_off32<<<grid_for(A->n_rows + 1, kThreads), kThreads, 0, dev->stream>>>(A->off, P.off32, A->n_rows + 1);
        SPIRA_HIP(hipGetLastError());
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
    SPIRA_HIP(hipMalloc(&P.blk[which], (nb + 1) * sizeof(uint32_t)));
    k_plan_scatter<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(A->off, n, C, pos, P.blk[which]);
    SPIRA_HIP(hipGetLastError());
    P.nblk[which] = (uint32_t)nb;
}

}  // namespace

void spmv_free_plan(spira_dcsr *A)
{
    spira_plan &P = A->plan;
    if (P.off32) (void)hipFree(P.off32);
    for (int i = 0; i < 2; ++i)
        if (P.blk[i]) (void)hipFree(P.blk[i]);
    int keep = P.variant;
    P = spira_plan{};
    (void)keep;  // a lock changes the matrix: the pinned variant is dropped too
}

void spmv_build_plan(spira_dcsr *A) { ensure_stats(A); }

int spmv_heuristic_variant(const spira_dcsr *A)
{
    const spira_plan &P = A->plan;
    if (P.variant >= 0) return P.variant;
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

void spmv_launch_f64(spira_dcsr *A, const double *x, double *y, int variant)
{
    if (A->n_rows == 0) return;
    if (A->n_rows >= (1ull << 32) - 1) fail(SPIRA_ERR_UNSUPPORTED, "spmv: n_rows must be < 2^32 - 1");
    ensure_stats(A);
    if (variant < 0) variant = spmv_heuristic_variant(A);
    const bool o32 = A->plan.off32 != nullptr;
    if (variant <= SPIRA_SPMV_VEC64) {
        const int g = 1 << variant;
        if (o32) launch_vec_g<uint32_t>(g, A, A->plan.off32, x, y);
        else launch_vec_g<uint64_t>(g, A, A->off, x, y);
    } else if (variant == SPIRA_SPMV_STREAM1K || variant == SPIRA_SPMV_STREAM2K) {
        const int which = variant == SPIRA_SPMV_STREAM1K ? 0 : 1;
        ensure_blocks(A, which);
        const uint32_t *blk = A->plan.blk[which];
        const uint32_t nb = A->plan.nblk[which];
        if (which == 0) {
            if (o32) launch_stream<1024, uint32_t>(A, blk, nb, A->plan.off32, x, y);
            else launch_stream<1024, uint64_t>(A, blk, nb, A->off, x, y);
        } else {
            if (o32) launch_stream<2048, uint32_t>(A, blk, nb, A->plan.off32, x, y);
            else launch_stream<2048, uint64_t>(A, blk, nb, A->off, x, y);
        }
    } else {
        fail(SPIRA_ERR_INVALID_ARGUMENT, "spmv: unknown kernel variant");
    }
    SPIRA_HIP(hipGetLastError());
}

}  // namespace spira_amd
