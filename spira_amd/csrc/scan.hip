// scan.hip -- exclusive prefix sum over u64 (reduce-then-scan, three kernels).
// Used for CSR offsets in lock() (csr_build.hpp:57-59 is the serial form) and
// for the SpMV row-block plan.
#include "internal.hpp"

namespace spira_amd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048 items per block

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        uint64_t t = __shfl_up(v, s, 64);
        if (lane >= s) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across the block; returns the block
// total in *total.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t *total)
{
    __shared__ uint64_t s_wave[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t inc = wave_incl_scan(v);
    if (lane == 63) s_wave[wid] = inc;
    __syncthreads();
    uint64_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        uint64_t s = s_wave[w];
        if (w < wid) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_tile_reduce(const uint64_t *__restrict__ in, uint64_t n,
                                                              uint64_t *__restrict__ sums)
{
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        uint64_t k = base + (uint64_t)i * kScanThreads + threadIdx.x;
        if (k < n) acc += in[k];
    }
    uint64_t tot;
    (void)block_excl_scan(acc, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// One block scans the tile sums in place (exclusive) and writes the grand
// total to sums[nb].
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(uint64_t *sums, uint64_t nb)
{
    uint64_t carry = 0;
    for (uint64_t base = 0; base < nb; base += kScanThreads) {
        uint64_t k = base + threadIdx.x;
        uint64_t v = k < nb ? sums[k] : 0;
        uint64_t tot;
        uint64_t ex = block_excl_scan(v, &tot);
        if (k < nb) sums[k] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const uint64_t *in, uint64_t *out, uint64_t n,
                                                            const uint64_t *__restrict__ sums)
{
    __shared__ uint64_t s_tile[kScanTile];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int j = i * kScanThreads + threadIdx.x;
        uint64_t k = base + j;
        s_tile[j] = k < n ? in[k] : 0;
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
    uint64_t run = block_excl_scan(acc, &tot) + sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        s_tile[threadIdx.x * kScanItems + i] = run;
        run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int j = i * kScanThreads + threadIdx.x;
        uint64_t k = base + j;
        if (k < n) out[k] = s_tile[j];
    }
}

__global__ void k_write_total(uint64_t *out, uint64_t n, const uint64_t *sums, uint64_t nb)
{
    out[n] = sums[nb];
}

}  // namespace

void scan_excl_u64(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n)
{
    hipStream_t s = dev->stream;
    const uint64_t nb = n == 0 ? 0 : (n + kScanTile - 1) / kScanTile;
    uint64_t *sums = (uint64_t *)dev->ws.get(WS_SCAN, (nb + 1) * sizeof(uint64_t));
    if (nb > 0) {
        k_tile_reduce<<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, sums);
        SPIRA_HIP(hipGetLastError());
    }
    k_scan_sums<<<1, kScanThreads, 0, s>>>(sums, nb);
    SPIRA_HIP(hipGetLastError());
    if (nb > 0) {
        k_tile_scan<<<(unsigned)nb, kScanThreads, 0, s>>>(in, out, n, sums);
        SPIRA_HIP(hipGetLastError());
    }
    k_write_total<<<1, 1, 0, s>>>(out, n, sums, nb);
    SPIRA_HIP(hipGetLastError());
}

}  // namespace spira_amd
