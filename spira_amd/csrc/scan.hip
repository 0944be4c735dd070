// scan.hip -- exclusive prefix sum of u64 words.
//
// Tiles of 4,096 words (256 threads x 16).  k_tile_sums writes each tile's
// sum.  k_tile_scan then gives every tile its prefix without waiting on any
// other block: with at most kDirectTiles tiles a block adds up the sums of
// the tiles before it itself (at most 8 KiB read from L2, spread over its
// 256 threads), otherwise it reads the prefix that k_sums_scan (one block)
// wrote over the sums.  The tile is staged in LDS, scanned and written; the
// last tile also writes the total at out[n] (and, if asked, a second copy).
// Two launches up to 4,194,304 words, three above.  Input and output may be the same array: each tile is
// read whole before it is written and no block reads another tile's words.
#include "internal.hpp"

namespace spira_amd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 4,096 words per tile
constexpr uint64_t kDirectTiles = 1024;

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const uint64_t t = __shfl_up(v, s, 64);
        if (lane >= s) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across the block; *total = the sum.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t *total)
{
    __shared__ uint64_t s_wave[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t inc = wave_incl_scan(v);
    if (lane == 63) s_wave[wid] = inc;
    __syncthreads();
    uint64_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const uint64_t s = s_wave[w];
        if (w < wid) base += s;
        tot += s;
    }
    __syncthreads();  // s_wave may be reused by the caller's next scan
    *total = tot;
    return base + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_tile_sums(const uint64_t *__restrict__ in, uint64_t n,
                                                            uint64_t *__restrict__ sums)
{
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t k = base + (uint64_t)i * kScanThreads + threadIdx.x;
        if (k < n) acc += in[k];
    }
    uint64_t tot;
    (void)block_excl_scan(acc, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// One block: sums[0..nt) replaced by their exclusive prefix.
__global__ __launch_bounds__(kScanThreads) void k_sums_scan(uint64_t *sums, uint64_t nt)
{
    uint64_t carry = 0;
    for (uint64_t base = 0; base < nt; base += kScanThreads) {
        const uint64_t k = base + threadIdx.x;
        const uint64_t v = k < nt ? sums[k] : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan(v, &tot);
        if (k < nt) sums[k] = carry + ex;
        carry += tot;
    }
}

// DIRECT: sums holds the tile sums and each block adds up those before it;
// otherwise sums holds their exclusive prefix.
template <bool DIRECT>
__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const uint64_t *in, uint64_t *out, uint64_t n,
                                                            const uint64_t *__restrict__ sums, uint64_t nt,
                                                            uint64_t *__restrict__ total_copy)
{
    __shared__ uint64_t s_tile[kScanTile];
    __shared__ uint64_t s_prefix;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const uint32_t j = (uint32_t)i * kScanThreads + threadIdx.x;
        const uint64_t k = base + j;
        s_tile[j] = k < n ? in[k] : 0ull;
    }
    if constexpr (DIRECT) {
        uint64_t acc = 0;
        for (uint64_t t = threadIdx.x; t < blockIdx.x; t += kScanThreads) acc += sums[t];
        uint64_t tot;
        (void)block_excl_scan(acc, &tot);
        if (threadIdx.x == 0) s_prefix = tot;
    } else {
        if (threadIdx.x == 0) s_prefix = sums[blockIdx.x];
    }
    __syncthreads();
    uint64_t v[kScanItems];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = s_tile[threadIdx.x * kScanItems + i];
        acc += v[i];
    }
    uint64_t tot;
    uint64_t run = s_prefix + block_excl_scan(acc, &tot);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        s_tile[threadIdx.x * kScanItems + i] = run;
        run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const uint32_t j = (uint32_t)i * kScanThreads + threadIdx.x;
        const uint64_t k = base + j;
        if (k < n) out[k] = s_tile[j];
    }
    if (threadIdx.x == 0 && blockIdx.x == nt - 1) {
        out[n] = s_prefix + tot;
        if (total_copy) *total_copy = s_prefix + tot;
    }
}

}  // namespace

void scan_excl_u64(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_copy)
{
    hipStream_t s = dev->stream;
    const uint64_t nt = n == 0 ? 1 : (n + kScanTile - 1) / kScanTile;
    if (nt >= (1ull << 31)) fail(SPIRA_ERR_UNSUPPORTED, "scan: 2^43 words or more");
    uint64_t *sums = (uint64_t *)dev->ws.get(WS_SCAN, nt * sizeof(uint64_t));
    k_tile_sums<<<(unsigned)nt, kScanThreads, 0, s>>>(in, n, sums);
    SPIRA_HIP(hipGetLastError());
    if (nt <= kDirectTiles) {
        k_tile_scan<true><<<(unsigned)nt, kScanThreads, 0, s>>>(in, out, n, sums, nt, total_copy);
    } else {
        k_sums_scan<<<1, kScanThreads, 0, s>>>(sums, nt);
        SPIRA_HIP(hipGetLastError());
        k_tile_scan<false><<<(unsigned)nt, kScanThreads, 0, s>>>(in, out, n, sums, nt, total_copy);
    }
    SPIRA_HIP(hipGetLastError());
}

}  // namespace spira_amd
