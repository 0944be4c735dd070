// probe.hip -- two measured ceilings of the part, for the SpMV roofline.
//
//   gather   random 8-byte loads that hit L2: every block (one 1024-thread
//            block per CU) gathers at hashed positions inside its XCD's own
//            2 MiB slice of a 32 MiB array (the XPHASE slice size), 20 loads
//            in flight per thread, so each load is its own L1-to-L2 request.
//            Result: requests per second.
//   stream   a read-only sweep of 1 GiB with 16-byte non-temporal loads.
//            Result: bytes per second.
//
// The XPHASE kernels issue both kinds of request from the same per-CU memory
// pipeline: their x gathers (one request per distinct 128-byte x line per
// (row block, slice) group) and their entry stream.  These two rates bound
// how fast a given number of each can be served.
#include "internal.hpp"

namespace spira_amd {

namespace {

constexpr int kProbeThreads = 1024;
constexpr int kProbeU = 20;
constexpr uint64_t kProbeSliceElems = (2ull << 20) / sizeof(double);  // 2 MiB
constexpr uint64_t kProbeXElems = 8 * kProbeSliceElems;                  // 32 MiB
constexpr uint64_t kProbeStreamBytes = 1ull << 30;

__device__ __forceinline__ uint32_t probe_hash(uint32_t v)
{
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

// Block b gathers from slice b % 8 (its XCD's, as hardware blocks are dealt
// to the 8 XCDs round robin).  Positions stay inside the slice.
__global__ __launch_bounds__(kProbeThreads) void k_probe_gather(const double *__restrict__ x, uint32_t rounds,
                                                                double *__restrict__ out)
{
    const double *xs = x + (uint64_t)(blockIdx.x % 8) * kProbeSliceElems;
    uint32_t h = probe_hash(blockIdx.x * kProbeThreads + threadIdx.x + 1u);
    double acc = 0.0;
    for (uint32_t r = 0; r < rounds; ++r) {
        double v[kProbeU];
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            h = probe_hash(h + (uint32_t)u);
            v[u] = xs[h & (uint32_t)(kProbeSliceElems - 1)];
        }
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) acc += v[u];
    }
    if (acc == 1.0e300) out[blockIdx.x] = acc;  // keeps the loads live; x holds zeros
}

__global__ void k_probe_stream(const u32x4 *__restrict__ a, uint64_t n, uint32_t *__restrict__ out)
{
    u32x4 acc{0u, 0u, 0u, 0u};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        acc ^= ld_stream(a + i);
    if ((acc.x | acc.y | acc.z | acc.w) == 0x9e3779b9u) out[0] = acc.x;  // keeps the loads live; a holds zeros
}

template <class F>
float time_ms(hipStream_t st, int reps, F &&launch)
{
    hipEvent_t e0, e1;
    SPIRA_HIP(hipEventCreate(&e0));
    SPIRA_HIP(hipEventCreate(&e1));
    launch();  // warm
    SPIRA_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < reps; ++i) launch();
    SPIRA_HIP(hipEventRecord(e1, st));
    SPIRA_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    SPIRA_HIP(hipEventElapsedTime(&ms, e0, e1));
    SPIRA_HIP(hipEventDestroy(e0));
    SPIRA_HIP(hipEventDestroy(e1));
    return ms / (float)reps;
}

}  // namespace

void probe_ceilings(spira_dev *dev, double *gather_req_per_s, double *stream_bytes_per_s)
{
    hipStream_t st = dev->stream;
    const int n_cu = dev->n_cu > 0 ? dev->n_cu : 256;
    struct temp {
        spira_dev *d;
        void *p;
        size_t bytes;
        ~temp() { dev_release(d, p, bytes); }
    } tx{dev, dev_alloc(dev, kProbeXElems * sizeof(double)), kProbeXElems * sizeof(double)},
        ts{dev, dev_alloc(dev, kProbeStreamBytes), kProbeStreamBytes},
        to{dev, dev_alloc(dev, 4096 * sizeof(double)), 4096 * sizeof(double)};
    SPIRA_HIP(hipMemsetAsync(tx.p, 0, tx.bytes, st));
    SPIRA_HIP(hipMemsetAsync(ts.p, 0, ts.bytes, st));
    const uint32_t rounds = 64;
    const uint32_t grid = (uint32_t)std::min(n_cu, 4096);
    const float g_ms = time_ms(st, 5, [&] {
        k_probe_gather<<<grid, kProbeThreads, 0, st>>>((const double *)tx.p, rounds, (double *)to.p);
        SPIRA_HIP(hipGetLastError());
    });
    const double gathers = (double)grid * kProbeThreads * rounds * kProbeU;
    const uint64_t n16 = kProbeStreamBytes / 16;
    const float s_ms = time_ms(st, 5, [&] {
        k_probe_stream<<<(unsigned)(n_cu * 8), 256, 0, st>>>((const u32x4 *)ts.p, n16, (uint32_t *)to.p);
        SPIRA_HIP(hipGetLastError());
    });
    dev_sync(dev);
    if (gather_req_per_s) *gather_req_per_s = gathers / (g_ms * 1e-3);
    if (stream_bytes_per_s) *stream_bytes_per_s = (double)kProbeStreamBytes / (s_ms * 1e-3);
}

}  // namespace spira_amd
