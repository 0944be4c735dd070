// spmv_mixed.hip -- reduced-precision value storage with compensated FP32
// accumulation, and an FP64 recompute of rows that cancel.
//
// Values are stored as FP32 (4 B), BF16 (2 B) or split BF16 (hi + lo, 4 B),
// copied once from A's FP64 values and dropped at the next lock.  x is
// rounded to FP32 once per call (the XPHASE kernels of banded matrices round
// it where they gather it instead).  Each row is accumulated with the Dot2
// scheme of Ogita, Rump and Oishi ("Accurate sum and dot product", SIAM J. Sci.
// Comput. 26(6), 2005): p = a*x, e = fma(a, x, -p) (the exact product error),
// (s, q) = TwoSum(s, p) (Knuth), c += q + e, result s + c.  The same pass sums
// |a*x|; a row whose cancellation ratio sum|a*x| / |sum a*x| exceeds the
// caller's threshold (or whose result is not finite) is flagged (one byte per
// row, no atomics) and recomputed in FP64 from A's FP64 values and the FP64 x
// by a second kernel.
#include "internal.hpp"

#include <cmath>

namespace spira_amd {

namespace {

constexpr int kThreads = 256;

// ---- value decoders --------------------------------------------------------
struct vals_f32 {
    const float *v;
    __device__ __forceinline__ float operator()(uint64_t k) const { return ld_stream(v + k); }
};
struct vals_bf16 {
    const uint16_t *v;
    __device__ __forceinline__ float operator()(uint64_t k) const { return bf16_to_f32(ld_stream(v + k)); }
};
struct vals_split {
    const uint16_t *hi, *lo;
    __device__ __forceinline__ float operator()(uint64_t k) const
    {
        return bf16_to_f32(ld_stream(hi + k)) + bf16_to_f32(ld_stream(lo + k));
    }
};

// ---- compensated accumulation ---------------------------------------------
// Contraction is off inside these error-free transformations: fused into
// FMAs (s + a * x as one rounding) the correction would count the product's
// rounding error twice.
struct dot2 {
    float s = 0.0f, c = 0.0f, a = 0.0f;  // sum, compensation, sum of |a*x|
    __device__ __forceinline__ void add_prod(float p, float e)
    {
#pragma clang fp contract(off)
        const float t = s + p;
        const float bp = t - s;
        const float q = (s - (t - bp)) + (p - bp);  // TwoSum error
        s = t;
        c += q + e;
        a += fabsf(p);
    }
    __device__ __forceinline__ void add(float av, float xv)
    {
#pragma clang fp contract(off)
        const float p = av * xv;
        add_prod(p, fmaf(av, xv, -p));
    }
    __device__ __forceinline__ void merge(float s2, float c2, float a2)
    {
#pragma clang fp contract(off)
        const float t = s + s2;
        const float bp = t - s;
        const float q = (s - (t - bp)) + (s2 - bp);
        s = t;
        c += c2 + q;
        a += a2;
    }
    // Merge across aligned groups of g lanes (g a power of two <= 64).
    __device__ __forceinline__ void group_reduce(int g)
    {
        for (int o = g / 2; o > 0; o >>= 1) {
            const float s2 = __shfl_xor(s, o, 64), c2 = __shfl_xor(c, o, 64), a2 = __shfl_xor(a, o, 64);
            merge(s2, c2, a2);
        }
    }
};

template <class YT>
__device__ __forceinline__ void finish_row(const dot2 &d, uint64_t row, YT *y, float ratio,
                                           uint8_t *flags, uint32_t *count)
{
    const double v = (double)d.s + (double)d.c;
    y[row] = (YT)v;
    if (ratio > 0.0f) {
        const bool bad = !isfinite(v) || (double)d.a > (double)ratio * fabs(v);
        flags[row] = bad ? 1u : 0u;
    }
}

// ---- kernels ---------------------------------------------------------------
__global__ void k_x_to_f32(const double *__restrict__ x, float *__restrict__ x32, uint64_t n, uint32_t *count)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x32[i] = (float)x[i];
    if (i == 0 && count) *count = 0u;  // reset the recompute count (no memset: graph-safe)
}

// VEC<G>: G lanes per row.  y is FP64 (mixed precision) or FP32 (the float
// instantiation).
template <int G, class Vals, class YT>
__global__ __launch_bounds__(kThreads) void k_mixed_vec(uint64_t n_rows, const uint32_t *__restrict__ off,
                                                         const uint32_t *__restrict__ col, Vals vals,
                                                         const float *__restrict__ x, YT *__restrict__ y,
                                                         float ratio, uint8_t *flags, uint32_t *count)
{
    const uint64_t gid = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    const uint64_t row = gid / G;
    const uint32_t lane = (uint32_t)(gid % G);
    if (row >= n_rows) return;
    const uint32_t b = off[row], e = off[row + 1];
    dot2 d;
    for (uint32_t k = b + lane; k < e; k += G) d.add(vals(k), x[ld_stream(col + k)]);
    d.group_reduce(G);
    if (lane == 0) finish_row(d, row, y, ratio, flags, count);
}

// STREAM<C>: LDS-staged row blocks (same plan as the FP64 kernel); the LDS
// holds each product and its exact error.
template <int C, class Vals, class YT>
__global__ __launch_bounds__(kThreads) void k_mixed_stream(const uint32_t *__restrict__ blk_row,
                                                            const uint32_t *__restrict__ off,
                                                            const uint32_t *__restrict__ col, Vals vals,
                                                            const float *__restrict__ x, YT *__restrict__ y,
                                                            float ratio, uint8_t *flags, uint32_t *count)
{
    __shared__ float s_p[2 * C];
    __shared__ float s_e[2 * C];
    __shared__ uint32_t s_off[C + 1];
    __shared__ float s_red[3][kThreads / 64];
    const uint32_t r0 = blk_row[blockIdx.x], r1 = blk_row[blockIdx.x + 1];
    const uint32_t nr = r1 - r0;
    const uint32_t k0 = off[r0], k1 = off[r1];
    if (nr == 1) {  // one (possibly very long) row: all lanes
        dot2 d;
        for (uint32_t k = k0 + threadIdx.x; k < k1; k += kThreads) d.add(vals(k), x[ld_stream(col + k)]);
        d.group_reduce(64);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            s_red[0][w] = d.s;
            s_red[1][w] = d.c;
            s_red[2][w] = d.a;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            dot2 t;
            for (int i = 0; i < kThreads / 64; ++i) t.merge(s_red[0][i], s_red[1][i], s_red[2][i]);
            finish_row(t, r0, y, ratio, flags, count);
        }
        return;
    }
    const uint32_t nk = k1 - k0;  // < 2C
    for (uint32_t i = threadIdx.x; i < nk; i += kThreads) {
        const float a = vals(k0 + i);
        const float xv = x[ld_stream(col + k0 + i)];
        const float p = a * xv;
        s_p[i] = p;
        s_e[i] = fmaf(a, xv, -p);
    }
    for (uint32_t j = threadIdx.x; j <= nr; j += kThreads) s_off[j] = off[r0 + j] - k0;
    __syncthreads();
    int T = 1;
    while (T < 64 && (uint32_t)(2 * T) * nr <= (uint32_t)kThreads) T <<= 1;
    const uint32_t teams = kThreads / T;
    const uint32_t lane = threadIdx.x & (T - 1);
    for (uint32_t r = threadIdx.x / T; r < nr; r += teams) {
        const uint32_t a = s_off[r], e = s_off[r + 1];
        dot2 d;
        for (uint32_t k = a + lane; k < e; k += T) d.add_prod(s_p[k], s_e[k]);
        d.group_reduce(T);
        if (lane == 0) finish_row(d, r0 + r, y, ratio, flags, count);
    }
}

// FP64 recompute of flagged rows.  A wavefront takes 64 rows at a time,
// finds the flagged ones with a ballot and recomputes them eight at a time
// (8 lanes per row); *count (reset by k_x_to_f32) gets the number of
// recomputed rows with one atomic per block.
__global__ __launch_bounds__(kThreads) void k_recompute_f64(const uint8_t *__restrict__ flags, uint64_t n_rows,
                                                             uint32_t *__restrict__ count,
                                                             const uint64_t *__restrict__ off,
                                                             const uint32_t *__restrict__ col,
                                                             const double *__restrict__ val,
                                                             const double *__restrict__ x, double *__restrict__ y)
{
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0u;
    __syncthreads();
    const uint64_t waves = (uint64_t)gridDim.x * (kThreads / 64);
    const int lane = threadIdx.x & 63;
    const int g = lane >> 3, gl = lane & 7;
    uint32_t mine = 0;
    for (uint64_t c = (uint64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); c * 64 < n_rows; c += waves) {
        const uint64_t r = c * 64 + (uint64_t)lane;
        const bool f = r < n_rows && flags[r] != 0;
        const uint64_t rb = f ? off[r] : 0, re = f ? off[r + 1] : 0;
        uint64_t m = __ballot(f);
        mine += (uint32_t)__popcll(m);
        // rows of more than 64 nonzeros: the whole wavefront, one at a time
        uint64_t ml = __ballot(f && re - rb > 64);
        m &= ~ml;
        while (ml) {
            const int bit = __ffsll((long long)ml) - 1;
            ml &= ml - 1;
            const uint64_t row = c * 64 + (uint64_t)bit;
            const uint64_t b = __shfl(rb, bit, 64), e = __shfl(re, bit, 64);
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            uint64_t q = b + (uint64_t)lane;
            for (; q + 192 < e; q += 256) {
                a0 += val[q] * x[col[q]];
                a1 += val[q + 64] * x[col[q + 64]];
                a2 += val[q + 128] * x[col[q + 128]];
                a3 += val[q + 192] * x[col[q + 192]];
            }
            for (; q < e; q += 64) a0 += val[q] * x[col[q]];
            const double acc = group_sum<64>((a0 + a1) + (a2 + a3));
            if (lane == 0) y[row] = acc;
        }
        // the others: lane group g takes flagged rows g, g + 8, g + 16, ... of
        // the chunk (in bit order), so the eight groups work concurrently
        for (int s = 0; s < g && m; ++s) m &= m - 1;
        while (m) {  // divergent across groups: bounds come from memory, not a shuffle
            const uint64_t row = c * 64 + (uint64_t)(__ffsll((long long)m) - 1);
            const uint64_t b = off[row], e = off[row + 1];
            double acc = 0.0;
            for (uint64_t q = b + gl; q < e; q += 8) acc += val[q] * x[col[q]];
            acc = group_sum<8>(acc);
            if (gl == 0) y[row] = acc;
            for (int s = 0; s < 8 && m; ++s) m &= m - 1;
        }
    }
    if (lane == 0 && mine) atomicAdd(&s_cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) atomicAdd(count, s_cnt);
}

// ---- storage conversion ------------------------------------------------------
__global__ void k_to_f32(const double *__restrict__ v, float *__restrict__ o, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (float)v[i];
}

__global__ void k_to_bf16(const double *__restrict__ v, uint16_t *__restrict__ o, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = f32_to_bf16((float)v[i]);
}

__global__ void k_to_split(const double *__restrict__ v, uint16_t *__restrict__ hi, uint16_t *__restrict__ lo,
                           uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t h = f32_to_bf16((float)v[i]);
    hi[i] = h;
    lo[i] = f32_to_bf16((float)(v[i] - (double)bf16_to_f32(h)));
}

void ensure_storage(spira_dcsr *A, int storage)
{
    spira_dev *dev = A->dev;
    const uint64_t n = A->nnz, cap = A->capacity ? A->capacity : 1;
    const unsigned grid = grid_for(n ? n : 1, kThreads);
    if (storage == SPIRA_VAL_F32 && !A->val_f32) {
        A->val_f32 = (float *)dev_alloc(dev, cap * sizeof(float));
        k_to_f32<<<grid, kThreads, 0, dev->stream>>>(A->val, A->val_f32, n);
    } else if (storage == SPIRA_VAL_BF16 && !A->val_bf16) {
        A->val_bf16 = (uint16_t *)dev_alloc(dev, cap * sizeof(uint16_t));
        k_to_bf16<<<grid, kThreads, 0, dev->stream>>>(A->val, A->val_bf16, n);
    } else if (storage == SPIRA_VAL_SPLIT_BF16 && !A->val_split_hi) {
        uint16_t *hi = (uint16_t *)dev_alloc(dev, cap * sizeof(uint16_t));
        uint16_t *lo = (uint16_t *)dev_alloc(dev, cap * sizeof(uint16_t));
        k_to_split<<<grid, kThreads, 0, dev->stream>>>(A->val, hi, lo, n);
        A->val_split_hi = hi;
        A->val_bf16_lo = lo;
    }
    SPIRA_HIP(hipGetLastError());
}

template <class Vals, class YT>
void launch_mixed(spira_dcsr *A, int variant, Vals vals, const float *x32, YT *y, float ratio, uint8_t *flags,
                  uint32_t *count, const uint32_t *blk, uint32_t nblk)
{
    hipStream_t st = A->dev->stream;
    const uint32_t *off = A->plan.off32;
    const uint64_t n = A->n_rows;
    switch (variant) {
    case SPIRA_SPMV_VEC4:
        k_mixed_vec<4, Vals, YT><<<grid_for(n * 4, kThreads), kThreads, 0, st>>>(n, off, A->col, vals, x32, y,
                                                                                ratio, flags, count);
        break;
    case SPIRA_SPMV_VEC8:
        k_mixed_vec<8, Vals, YT><<<grid_for(n * 8, kThreads), kThreads, 0, st>>>(n, off, A->col, vals, x32, y,
                                                                                ratio, flags, count);
        break;
    case SPIRA_SPMV_VEC16:
        k_mixed_vec<16, Vals, YT><<<grid_for(n * 16, kThreads), kThreads, 0, st>>>(n, off, A->col, vals, x32, y,
                                                                                  ratio, flags, count);
        break;
    case SPIRA_SPMV_VEC32:
        k_mixed_vec<32, Vals, YT><<<grid_for(n * 32, kThreads), kThreads, 0, st>>>(n, off, A->col, vals, x32, y,
                                                                                  ratio, flags, count);
        break;
    case SPIRA_SPMV_STREAM1K:
        k_mixed_stream<1024, Vals, YT><<<nblk, kThreads, 0, st>>>(blk, off, A->col, vals, x32, y, ratio, flags,
                                                                  count);
        break;
    default:
        k_mixed_stream<2048, Vals, YT><<<nblk, kThreads, 0, st>>>(blk, off, A->col, vals, x32, y, ratio, flags,
                                                                  count);
        break;
    }
    SPIRA_HIP(hipGetLastError());
}

bool row_major_variant(int variant)
{
    return variant == SPIRA_SPMV_VEC4 || variant == SPIRA_SPMV_VEC8 || variant == SPIRA_SPMV_VEC16 ||
           variant == SPIRA_SPMV_VEC32 || variant == SPIRA_SPMV_STREAM1K || variant == SPIRA_SPMV_STREAM2K;
}

bool xphase_variant(int variant) { return variant == SPIRA_SPMV_XPHASE || variant == SPIRA_SPMV_XPHASE_PACED; }

// The FP64 AUTO choice mapped onto the reduced-precision kernels (the XPHASE
// layout kept, unpaced: with FP32 x a slice covers twice the columns, and the
// unpaced kernel measured faster on every matrix tried -- random 4M rows 298
// against 306 us, power-law 8.4M rows 867 against 1,075 us, the 2048^2
// Laplacian 39.8 against 108 us; the XCD-slice copies go to the LDS row
// blocks; VEC widths outside 4..32 to the nearest one).
int reduced_auto(spira_dcsr *A)
{
    const int v = spmv_resolve_auto(A);
    if (xphase_variant(v)) return SPIRA_SPMV_XPHASE;
    if (row_major_variant(v)) return v;
    if (v >= SPIRA_SPMV_XSLICE1 && v <= SPIRA_SPMV_XSTREAM) return SPIRA_SPMV_STREAM2K;
    return v < SPIRA_SPMV_VEC4 ? SPIRA_SPMV_VEC4 : SPIRA_SPMV_VEC32;
}

}  // namespace

bool mixed_variant_supported(int variant) { return row_major_variant(variant) || xphase_variant(variant); }

// WS_MIXQ: one flag byte per row, then (16-byte aligned) the recompute count.
static uint8_t *mixed_flags(spira_dcsr *A)
{
    return (uint8_t *)A->dev->ws.get(WS_MIXQ, ((A->n_rows + 15) & ~(uint64_t)15) + 16);
}

uint32_t *mixed_queue_count(spira_dcsr *A)
{
    return (uint32_t *)(mixed_flags(A) + ((A->n_rows + 15) & ~(uint64_t)15));
}

void spmv_mixed_launch(spira_dcsr *A, int storage, const double *x, double *y, int variant, float ratio)
{
    if (A->n_rows == 0) return;
    if (A->nnz >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "mixed spmv: nnz must be < 2^32");
    spmv_build_plan(A);
    spira_dev *dev = A->dev;
    hipStream_t st = dev->stream;
    if (variant < 0) {
        // the recompute needs each row's sum of |a*x|, which only the
        // row-major kernels form (the XPHASE rows are summed by LDS atomics)
        variant = reduced_auto(A);
        if (ratio > 0.0f && !row_major_variant(variant)) {
            variant = A->plan.variant >= 0 ? A->plan.variant : spmv_rule_variant(A);
            if (!row_major_variant(variant)) variant = SPIRA_SPMV_STREAM2K;
        }
    }
    if (!mixed_variant_supported(variant))
        fail(SPIRA_ERR_INVALID_ARGUMENT,
             "mixed spmv: variant must be VEC4..VEC32, STREAM1K, STREAM2K, XPHASE or XPHASE_PACED");
    if (xphase_variant(variant)) {
        if (ratio > 0.0f)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "mixed spmv: the FP64 recompute needs a VEC or STREAM variant");
        // A matrix that AUTO runs on unpaced XPHASE (a row block meets few
        // slices) gathers FP64 x and rounds it there: no pass of its own
        // (2048^2 Laplacian, FP32 values: 62.6 -> 46.9 us).  Scattered
        // columns take x rounded to FP32 first: half the bytes behind each
        // gathered line (random 4M rows: 313 against 410 us gathering FP64).
        const bool paced = variant == SPIRA_SPMV_XPHASE_PACED;
        if (spmv_resolve_auto(A) == SPIRA_SPMV_XPHASE) {
            spmv_xphase_reduced(A, storage, x, y, paced);
        } else {
            float *x32 = (float *)dev->ws.get(WS_X32, (A->n_cols ? A->n_cols : 1) * sizeof(float));
            k_x_to_f32<<<grid_for(A->n_cols ? A->n_cols : 1, kThreads), kThreads, 0, st>>>(x, x32, A->n_cols,
                                                                                             mixed_queue_count(A));
            SPIRA_HIP(hipGetLastError());
            spmv_xphase_reduced(A, storage, x32, y, paced);
        }
        return;
    }
    ensure_storage(A, storage);
    const uint32_t *blk = nullptr;
    uint32_t nblk = 0;
    if (variant == SPIRA_SPMV_STREAM1K || variant == SPIRA_SPMV_STREAM2K) {
        const int which = variant == SPIRA_SPMV_STREAM1K ? 0 : 1;
        spmv_stream_blocks(A, which, &blk, &nblk);
    }
    uint8_t *flags = mixed_flags(A);
    uint32_t *count = mixed_queue_count(A);
    float *x32 = (float *)dev->ws.get(WS_X32, (A->n_cols ? A->n_cols : 1) * sizeof(float));
    k_x_to_f32<<<grid_for(A->n_cols ? A->n_cols : 1, kThreads), kThreads, 0, st>>>(x, x32, A->n_cols, count);
    SPIRA_HIP(hipGetLastError());
    if (storage == SPIRA_VAL_F32)
        launch_mixed(A, variant, vals_f32{A->val_f32}, x32, y, ratio, flags, count, blk, nblk);
    else if (storage == SPIRA_VAL_BF16)
        launch_mixed(A, variant, vals_bf16{A->val_bf16}, x32, y, ratio, flags, count, blk, nblk);
    else
        launch_mixed(A, variant, vals_split{A->val_split_hi, A->val_bf16_lo}, x32, y, ratio, flags, count, blk,
                     nblk);
    if (ratio > 0.0f) {
        const unsigned g = std::min<uint64_t>(grid_for(A->n_rows, kThreads / 64), 2048u);
        k_recompute_f64<<<g, kThreads, 0, st>>>(flags, A->n_rows, count, A->off, A->col, A->val, x, y);
        SPIRA_HIP(hipGetLastError());
    }
}

void spmv_f32_launch(spira_dcsr *A, const float *x, float *y, int variant)
{
    if (A->n_rows == 0) return;
    if (A->nnz >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "spmv (float): nnz must be < 2^32");
    spmv_build_plan(A);
    if (variant < 0) variant = reduced_auto(A);
    if (!mixed_variant_supported(variant))
        fail(SPIRA_ERR_INVALID_ARGUMENT,
             "spmv (float): variant must be VEC4..VEC32, STREAM1K, STREAM2K, XPHASE or XPHASE_PACED");
    if (xphase_variant(variant)) {
        spmv_xphase_f32(A, x, y, variant == SPIRA_SPMV_XPHASE_PACED);
        return;
    }
    ensure_storage(A, SPIRA_VAL_F32);
    const uint32_t *blk = nullptr;
    uint32_t nblk = 0;
    if (variant == SPIRA_SPMV_STREAM1K || variant == SPIRA_SPMV_STREAM2K)
        spmv_stream_blocks(A, variant == SPIRA_SPMV_STREAM1K ? 0 : 1, &blk, &nblk);
    launch_mixed(A, variant, vals_f32{A->val_f32}, x, y, 0.0f, nullptr, nullptr, blk, nblk);
}

}  // namespace spira_amd
