// scan.hip -- exclusive prefix sum of u64 words: one pass, one launch.
//
// A block takes the next tile of 4,096 words from a tile counter (so every
// tile it may wait on was taken earlier by a block that is running), sums it,
// publishes the sum, then looks back: wavefront 0 reads the flags of the 64
// preceding tiles at once and adds their sums up to the nearest tile that has
// published its inclusive prefix, moving the window back until it finds one
// (decoupled look-back; Merrill and Garland, "Single-pass Parallel Prefix Scan
// with Decoupled Look-back", NVIDIA tech report NVR-2016-002).  The block
// then publishes its own inclusive prefix and writes its tile.  Input and
// output may be the same array: a tile is read whole before it is written,
// and no block reads another tile's words.
//
// A tile's state is four 64-bit words, each written once per call with one
// atomic store and read with one atomic load (no flag/value pair, so no
// cache maintenance between them): the low and high halves of its sum and of
// its inclusive prefix, each word = call epoch << 32 | 32-bit half.  A value
// counts as published when both of its halves carry the current epoch; words
// left by earlier calls carry older epochs, so the state buffer is cleared
// only when the epoch wraps.  The tile counter carries the epoch too.  A wait that exceeds a bound (it cannot in a
// correct run: every tile it waits on belongs to a running block) sets an
// error word and stops waiting instead of hanging; scan_errors() reads and
// clears it.
#include "internal.hpp"

namespace spira_amd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 4,096 words per tile
constexpr uint32_t kAggregate = 1u, kInclusive = 2u;
constexpr uint32_t kSpinMax = 1u << 20;
constexpr uint32_t kEpochMax = 0xfffffff0u;
// state buffer: tile counter (64 bits), error word, then 4 words per tile
constexpr size_t kCtrBytes = 256;
constexpr int kWordsPerTile = 4;  // sum lo, sum hi, inclusive lo, inclusive hi

__device__ __forceinline__ void publish(uint64_t *w, uint32_t epoch, uint64_t v)
{
    const uint64_t e = (uint64_t)epoch << 32;
    __hip_atomic_store(w, e | (v & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 1, e | (v >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// true (and *v) when both halves at w carry epoch
__device__ __forceinline__ bool fetch(const uint64_t *w, uint32_t epoch, uint64_t *v)
{
    const uint64_t lo = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t hi = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *v = (hi << 32) | (lo & 0xffffffffull);
    return (uint32_t)(lo >> 32) == epoch && (uint32_t)(hi >> 32) == epoch;
}

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

__device__ __forceinline__ uint64_t wave_sum(uint64_t v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    return v;
}

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
    *total = tot;
    return base + inc - v;
}

// The tile counter is one 64-bit word, epoch << 32 | tiles taken.  The first
// block of a call finds an older epoch there and swaps in (epoch, 1), taking
// tile 0; every other block adds one.  So a call never depends on the
// previous one having reset the counter.
__device__ __forceinline__ uint32_t take_tile(uint32_t *ctr32, uint32_t epoch)
{
    uint64_t *ctr = (uint64_t *)ctr32;
    uint64_t old = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if ((uint32_t)(old >> 32) != epoch) {
            const uint64_t want = ((uint64_t)epoch << 32) | 1ull;
            if (__hip_atomic_compare_exchange_strong(ctr, &old, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                return 0u;
            continue;  // old now holds the current word
        }
        old = __hip_atomic_fetch_add(ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return (uint32_t)old;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_lookback(const uint64_t *in, uint64_t *out, uint64_t n,
                                                                uint32_t tiles, uint32_t *__restrict__ ctr,
                                                                uint64_t *__restrict__ state, uint32_t epoch)
{
    __shared__ uint64_t s_tile[kScanTile];
    __shared__ uint32_t s_id;
    __shared__ uint64_t s_prefix;
    if (threadIdx.x == 0) s_id = take_tile(ctr, epoch);
    __syncthreads();
    const uint32_t t = s_id;
    const uint64_t base = (uint64_t)t * kScanTile;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const uint32_t j = (uint32_t)i * kScanThreads + threadIdx.x;
        const uint64_t k = base + j;
        s_tile[j] = k < n ? in[k] : 0ull;
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
    const uint64_t ex = block_excl_scan(acc, &tot);

    const uint32_t lane = threadIdx.x & 63;
    if (threadIdx.x < 64) {
        uint64_t prefix = 0;
        bool bad = false;
        uint64_t *mine = state + (uint64_t)t * kWordsPerTile;
        if (t == 0) {
            if (lane == 0) publish(mine + 2, epoch, tot);
        } else {
            if (lane == 0) publish(mine, epoch, tot);
            int64_t p = (int64_t)t - 1;  // nearest tile of the window
            uint32_t spins = 0;
            for (;;) {
                const int64_t q = p - (int64_t)lane;
                uint32_t st = kInclusive;  // before tile 0: an inclusive prefix of 0
                uint64_t val = 0;
                if (q >= 0) {
                    const uint64_t *w = state + (uint64_t)q * kWordsPerTile;
                    if (fetch(w + 2, epoch, &val)) st = kInclusive;
                    else if (fetch(w, epoch, &val)) st = kAggregate;
                    else st = 0u;
                }
                const uint64_t incl_lanes = __ballot(st == kInclusive);
                const uint32_t first = incl_lanes ? (uint32_t)__builtin_ctzll(incl_lanes) : 64u;
                const uint64_t before = first >= 64 ? ~0ull : ((1ull << first) - 1ull);  // lanes nearer than it
                if (__ballot(st == 0u) & before) {
                    if (++spins >= kSpinMax) {  // cannot happen in a correct run
                        bad = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += wave_sum(lane <= first ? val : 0ull);
                if (first < 64) break;
                p -= 64;
            }
            if (lane == 0) publish(mine + 2, epoch, prefix + tot);
        }
        if (lane == 0) {
            s_prefix = prefix;
            if (bad) __hip_atomic_store(&ctr[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    uint64_t run = s_prefix + ex;
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
    if (threadIdx.x == 0 && t == tiles - 1) out[n] = s_prefix + tot;
}

size_t state_bytes(uint64_t tiles) { return kCtrBytes + tiles * kWordsPerTile * sizeof(uint64_t); }

}  // namespace

void scan_excl_u64(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n)
{
    hipStream_t s = dev->stream;
    const uint64_t tiles = n == 0 ? 1 : (n + kScanTile - 1) / kScanTile;
    if (tiles >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "scan: 2^44 words or more");
    if (tiles > dev->scan_tiles_cap || dev->scan_epoch >= kEpochMax) {
        // (re)build the state: zeroed flags and counters (queued work that
        // used the old buffer finishes first)
        if (tiles > dev->scan_tiles_cap) {
            uint64_t cap = std::max<uint64_t>(tiles + tiles / 4, 1024);
            if (dev->scan_state) {
                SPIRA_HIP(hipStreamSynchronize(s));
                SPIRA_HIP(hipFree(dev->scan_state));
                dev->scan_state = nullptr;
                dev->scan_tiles_cap = 0;
            }
            SPIRA_HIP(hipMalloc(&dev->scan_state, state_bytes(cap)));
            dev->scan_tiles_cap = cap;
        }
        SPIRA_HIP(hipMemsetAsync(dev->scan_state, 0, state_bytes(dev->scan_tiles_cap), s));
        dev->scan_epoch = 0;
    }
    const uint32_t epoch = ++dev->scan_epoch;
    unsigned char *b = (unsigned char *)dev->scan_state;
    uint32_t *ctr = (uint32_t *)b;
    uint64_t *state = (uint64_t *)(b + kCtrBytes);
    k_scan_lookback<<<(unsigned)tiles, kScanThreads, 0, s>>>(in, out, n, (uint32_t)tiles, ctr, state, epoch);
    SPIRA_HIP(hipGetLastError());
}

bool scan_errors(spira_dev *dev)
{
    if (!dev->scan_state) return false;
    uint32_t *err = (uint32_t *)dev->scan_state + 2;
    const bool bad = read_scalar(dev, (const uint32_t *)err) != 0u;
    if (bad) SPIRA_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), dev->stream));
    return bad;
}

}  // namespace spira_amd
