// scan.hip -- exclusive prefix sum of u64 words.
//
// Tiles of 4,096 words (256 threads x 16).  k_tile_sums writes each tile's
// sum.  k_tile_scan then gives every tile its prefix without waiting on any
// other block: while tiles x sums stays at most kDirectReads (2^21) a block
// adds up the sums before its tile itself (8 MiB read from L2 in all at
// most), otherwise it reads the prefix that k_sums_scan (one block) wrote
// over the sums.  The tile is staged in LDS, scanned and written; the
// last tile also writes the total at out[n] (and, if asked, a second copy).
// Two launches up to 5,931,008 words (1,448 tiles), three above.  Input and output may be
// the same array: each tile is read whole before it is written and no block
// reads another tile's words.
//
// scan_excl_u64_chunks skips k_tile_sums: the kernel that produced the input
// wrote the sum of every chunk of `chunk` words (chunk divides 4,096) into the
// array scan_chunk_sums returned, so one launch (two above kDirectReads)
// finishes the scan.
#include "internal.hpp"

namespace spira_amd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 4,096 words per tile
constexpr uint64_t kDirectReads = 1ull << 21;  // tiles x sums up to which each block adds its predecessors itself

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

// One block: sums[0..nt) replaced by their exclusive prefix, 4,096 sums
// (16 consecutive per thread) per step.
__global__ __launch_bounds__(kScanThreads) void k_sums_scan(uint64_t *sums, uint64_t nt)
{
    uint64_t carry = 0;
    for (uint64_t base = 0; base < nt; base += kScanTile) {
        uint64_t v[kScanItems];
        uint64_t acc = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            const uint64_t k = base + (uint64_t)threadIdx.x * kScanItems + i;
            v[i] = k < nt ? sums[k] : 0;
            acc += v[i];
        }
        uint64_t tot;
        uint64_t run = carry + block_excl_scan(acc, &tot);
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            const uint64_t k = base + (uint64_t)threadIdx.x * kScanItems + i;
            if (k < nt) sums[k] = run;
            run += v[i];
        }
        carry += tot;
    }
}

// sums holds one sum per chunk of kScanTile / per_tile words.  DIRECT: the
// chunk sums themselves, each block adds up those before its tile; otherwise
// their exclusive prefix.
template <bool DIRECT>
__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const uint64_t *in, uint64_t *out, uint64_t n,
                                                            const uint64_t *__restrict__ sums, uint32_t per_tile,
                                                            uint64_t nt, uint64_t *__restrict__ total_copy)
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
        const uint64_t before = (uint64_t)blockIdx.x * per_tile;
        for (uint64_t t = threadIdx.x; t < before; t += kScanThreads) acc += sums[t];
        uint64_t tot;
        (void)block_excl_scan(acc, &tot);
        if (threadIdx.x == 0) s_prefix = tot;
    } else {
        if (threadIdx.x == 0) s_prefix = sums[(uint64_t)blockIdx.x * per_tile];
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

namespace {

void launch_tile_scan(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *sums,
                      uint32_t per_tile, uint64_t *total_copy)
{
    hipStream_t s = dev->stream;
    const uint64_t nt = n == 0 ? 1 : (n + kScanTile - 1) / kScanTile;
    const uint64_t nsums = n == 0 ? 1 : (n + kScanTile / per_tile - 1) / (kScanTile / per_tile);
    // direct: the blocks read nt * nsums / 2 sums from L2 in all
    if (nt * nsums <= kDirectReads) {
        k_tile_scan<true><<<(unsigned)nt, kScanThreads, 0, s>>>(in, out, n, sums, per_tile, nt, total_copy);
    } else {
        k_sums_scan<<<1, kScanThreads, 0, s>>>(sums, nsums);
        SPIRA_HIP(hipGetLastError());
        k_tile_scan<false><<<(unsigned)nt, kScanThreads, 0, s>>>(in, out, n, sums, per_tile, nt, total_copy);
    }
    SPIRA_HIP(hipGetLastError());
}

}  // namespace

void scan_excl_u64(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_copy)
{
    const uint64_t nt = n == 0 ? 1 : (n + kScanTile - 1) / kScanTile;
    if (nt >= (1ull << 31)) fail(SPIRA_ERR_UNSUPPORTED, "scan: 2^43 words or more");
    uint64_t *sums = (uint64_t *)dev->ws.get(WS_SCAN, nt * sizeof(uint64_t));
    k_tile_sums<<<(unsigned)nt, kScanThreads, 0, dev->stream>>>(in, n, sums);
    SPIRA_HIP(hipGetLastError());
    launch_tile_scan(dev, in, out, n, sums, 1, total_copy);
}

uint64_t *scan_chunk_sums(spira_dev *dev, uint64_t n, uint32_t chunk)
{
    if (chunk == 0 || kScanTile % chunk) fail(SPIRA_ERR_DEVICE, "scan: the chunk must divide the tile");
    const uint64_t nt = n == 0 ? 1 : (n + kScanTile - 1) / kScanTile;
    if (nt >= (1ull << 31)) fail(SPIRA_ERR_UNSUPPORTED, "scan: 2^43 words or more");
    return (uint64_t *)dev->ws.get(WS_SCAN, nt * (kScanTile / chunk) * sizeof(uint64_t));
}

void scan_excl_u64_chunks(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *sums,
                          uint32_t chunk, uint64_t *total_copy)
{
    if (n == 0) {
        // no producer wrote a chunk sum: the total is 0
        k_tile_sums<<<1, kScanThreads, 0, dev->stream>>>(in, 0, sums);
        SPIRA_HIP(hipGetLastError());
        launch_tile_scan(dev, in, out, 0, sums, 1, total_copy);
        return;
    }
    launch_tile_scan(dev, in, out, n, sums, kScanTile / chunk, total_copy);
}

}  // namespace spira_amd
