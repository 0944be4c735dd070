// lock.hip -- device lock(): sort+dedup of the open-mode row buffers, then
// build_csr (first lock) or merge_csr (later locks), bit-identical to
// csr_build.hpp:32-98 / :126-488 and aos_array_buffer.hpp:76-101.
//
// Pipeline (all on the handle's stream):
//   1. check      stage_ptr well formed, columns < n_cols, rows binned by the
//                 number of staged inserts (<= 32, <= 8192, longer).
//   2. dedup      per row: last write wins, ascending columns, zeros kept.
//                 <= 32 inserts: one lane per row, O(k^2) ranking in registers
//                 and L1.  <= 8192: one block per row, bitonic sort of
//                 (column, insertion index) keys in LDS.  Longer rows (any
//                 length): one stable radix sort in HBM of all of them by
//                 (row in their list, column), run only when such rows exist.
//   3. flags      per deduped entry: nz = "value is not zero and the row is
//                 dirty", match = "column already present in the old row"
//                 (binary search); packed (match << 32 | nz) and scanned.
//   4. counts     new row length: clean rows keep their length; dirty rows
//                 old_len - matches + nz (- old zeros, if the old CSR can hold
//                 explicit zeros); exclusive scan -> new 64-bit offsets.
//   5. scatter    every old entry and every surviving new entry computes its
//                 own output slot from the prefix sums (merge by ranking), so
//                 the work is element-parallel and independent of row lengths;
//                 a re-lock does both in one pass over "row r's old entries,
//                 then its staged slots" so each row is written by one block.
// The host waits once, at the end, for the error bits and the new nnz; a
// failed check makes every later kernel return at once.  Rows longer than the
// LDS sort holds set a retry bit that acts like an error in that first pass;
// the host, which then knows how many such rows there are and how many
// inserts they hold, sorts them and runs steps 2-5 again (so matrices without
// such rows pay nothing for them).  The new arrays hold nnz_old + m entries
// (merge_csr's total_ub); the matrix reports nnz.
#include "internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstring>

namespace spira_amd {

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kSmallSeg = 32;
constexpr uint32_t kMaxSeg = 8192;                          // LDS bitonic limit of k_dedup_big
constexpr int kItems = 4;                                   // elements per thread in scatter passes
constexpr uint32_t kPerBlock = kThreads * kItems;           // elements per block of the staged-slot passes
constexpr int kItemsMerge = 8;                              // positions per thread in the re-lock scatter
constexpr uint32_t kPerBlockMerge = kThreads * kItemsMerge;

enum err_bits : uint32_t {
    ERR_OUT_OF_RANGE = 1u,
    ERR_BAD_PTR = 2u,
    ERR_LONG_ROWS = 4u,  // not an error: rows with > kMaxSeg inserts are listed, the host sorts them and reruns
    ERR_INTERNAL = 8u,   // a scatter slot fell outside its row (never expected)
};

// Control words of one lock, read back by the host in one copy at the end.
enum ctl_word : int {
    CTL_ERR = 0,       // err_bits
    CTL_NBIG = 1,      // rows listed for k_dedup_big
    CTL_NNZ_LO = 2,    // new nnz (written by the offsets scan; 8-byte aligned)
    CTL_NNZ_HI = 3,
    CTL_NLONG = 4,     // rows with more than kMaxSeg staged inserts
    CTL_LONG_INS = 5,  // staged inserts in those rows (< m < 2^32)
    CTL_WORDS = 8,
};

__device__ __forceinline__ bool row_dirty(const uint8_t *dirty, const uint64_t *sptr, uint64_t r)
{
    if (dirty) return dirty[r] != 0;
    return sptr[r + 1] != sptr[r];
}

// Row table of an element-parallel pass whose block b covers the elements
// [b*per, (b+1)*per) of a CSR-like pointer array: tab[b] = row owning element
// b*per, tab[nb] = row owning the last element (nb = ceil(total / per)), so a
// block reads its row range [tab[b], tab[b+1]] instead of searching for it.
// Called once for every row r (elements [e0, e1)) by a row-parallel kernel.
__device__ __forceinline__ void mark_block_rows(uint64_t r, uint64_t e0, uint64_t e1, uint64_t total, uint32_t per,
                                                uint64_t *__restrict__ tab)
{
    if (e0 == e1) return;
    for (uint64_t b = (e0 + per - 1) / per; b * per < e1; ++b) tab[b] = r;
    if (e1 == total) tab[(total + per - 1) / per] = r;  // the last non-empty row owns the last element
}

// ---------------------------------------------------------------------------
// 1. checks and binning
// ---------------------------------------------------------------------------
// One thread per staged pointer (rows 0..n) and per staged column (0..m-1):
// pointers well formed, columns < n_cols, rows binned by insert count.
__global__ void k_check_staged(const uint64_t *__restrict__ sptr, uint64_t n, uint64_t m,
                               const uint32_t *__restrict__ scol, uint64_t n_cols, uint32_t *ctl,
                               uint32_t *big_rows, uint32_t *long_rows)
{
    uint32_t *err = ctl + CTL_ERR;
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // Columns are read only when the caller's count agrees with stage_ptr[n]
    // (a count given by the caller may be wrong: then only ERR_BAD_PTR is set).
    if (r < m && sptr[n] == m && (uint64_t)scol[r] >= n_cols) atomicOr(err, ERR_OUT_OF_RANGE);
    if (r > n) return;
    if (r == 0 && sptr[0] != 0) atomicOr(err, ERR_BAD_PTR);
    if (r == n) {
        if (sptr[n] != m) atomicOr(err, ERR_BAD_PTR);
        return;
    }
    const uint64_t a = sptr[r], b = sptr[r + 1];
    if (b < a || b > m) {
        atomicOr(err, ERR_BAD_PTR);
        return;
    }
    const uint64_t k = b - a;
    if (k > kMaxSeg) {
        atomicOr(err, ERR_LONG_ROWS);
        long_rows[atomicAdd(ctl + CTL_NLONG, 1u)] = (uint32_t)r;
        atomicAdd(ctl + CTL_LONG_INS, (uint32_t)k);  // sums to at most m < 2^32
    } else if (k > kSmallSeg) {
        big_rows[atomicAdd(ctl + CTL_NBIG, 1u)] = (uint32_t)r;
    }
}

// ---------------------------------------------------------------------------
// 2. dedup
// ---------------------------------------------------------------------------

// Rows with <= 32 staged inserts: one lane per row.  keep_i = "no later insert
// to the same column" (the reverse + stable_sort + keep-first of
// aos_array_buffer.hpp:81-96 keeps exactly the last write); the output slot of
// a kept entry is the number of kept entries with a smaller column.
__global__ __launch_bounds__(kThreads) void k_dedup_small(uint64_t n, const uint64_t *__restrict__ sptr,
                                                           const uint32_t *__restrict__ scol,
                                                           const double *__restrict__ sval,
                                                           uint32_t *__restrict__ dcol, double *__restrict__ dval,
                                                           uint32_t *__restrict__ dcnt, uint64_t *__restrict__ stab,
                                                           const uint64_t *__restrict__ old_off,
                                                           uint64_t *__restrict__ mtab, const uint32_t *__restrict__ ctl)
{
    if (ctl[CTL_ERR]) return;  // the checks failed: nothing below may trust the staged arrays
    const uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    const uint64_t s0 = sptr[r];
    const uint64_t s1 = sptr[r + 1];
    // row tables of the later element-parallel passes: staged slots, and the
    // re-lock scatter's combined index space (old entries, then staged slots)
    mark_block_rows(r, s0, s1, sptr[n], kPerBlock, stab);
    if (mtab) mark_block_rows(r, old_off[r] + s0, old_off[r + 1] + s1, old_off[n] + sptr[n], kPerBlockMerge, mtab);
    const uint32_t k = (uint32_t)(s1 - s0);
    if (k > kSmallSeg) return;  // handled by the block kernel
    const uint32_t *c = scol + s0;
    uint32_t keep = 0;
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t ci = c[i];
        bool last = true;
        for (uint32_t j = i + 1; j < k; ++j) last &= (c[j] != ci);
        keep |= (last ? 1u : 0u) << i;
    }
    for (uint32_t i = 0; i < k; ++i) {
        if (!((keep >> i) & 1u)) continue;
        const uint32_t ci = c[i];
        uint32_t pos = 0;
        for (uint32_t j = 0; j < k; ++j) pos += (((keep >> j) & 1u) && c[j] < ci) ? 1u : 0u;
        dcol[s0 + pos] = ci;
        dval[s0 + pos] = sval[s0 + i];
    }
    dcnt[r] = (uint32_t)__builtin_popcount(keep);
}

__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t s_w[kThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        uint32_t t = __shfl_up(inc, s, 64);
        if (lane >= s) inc += t;
    }
    if (lane == 63) s_w[wid] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        if (w < wid) base += s_w[w];
        tot += s_w[w];
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

// Rows with 33..8192 staged inserts: one block per row (grid-stride over the
// list).  Keys (column << 32 | insertion index) are unique, so any sort gives
// the order of the reference's stable sort; the last key of each column run is
// the last write.
__global__ __launch_bounds__(kThreads) void k_dedup_big(const uint32_t *__restrict__ rows,
                                                         const uint32_t *__restrict__ ctl,  // [error bits, rows listed]
                                                         const uint64_t *__restrict__ sptr,
                                                         const uint32_t *__restrict__ scol,
                                                         const double *__restrict__ sval,
                                                         uint32_t *__restrict__ dcol, double *__restrict__ dval,
                                                         uint32_t *__restrict__ dcnt)
{
    extern __shared__ uint64_t s_key[];  // kMaxSeg entries
    if (ctl[CTL_ERR]) return;
    const uint32_t count = ctl[CTL_NBIG];
    for (uint32_t li = blockIdx.x; li < count; li += gridDim.x) {
        const uint32_t r = rows[li];
        const uint64_t s0 = sptr[r];
        const uint32_t k = (uint32_t)(sptr[r + 1] - s0);
        uint32_t N = 64;
        while (N < k) N <<= 1;
        for (uint32_t i = threadIdx.x; i < N; i += kThreads)
            s_key[i] = i < k ? (((uint64_t)scol[s0 + i]) << 32) | i : ~0ull;
        __syncthreads();
        for (uint32_t size = 2; size <= N; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t i = threadIdx.x; i < N / 2; i += kThreads) {
                    const uint32_t lo = 2 * i - (i & (stride - 1));
                    const uint32_t hi = lo + stride;
                    const bool asc = (lo & size) == 0;
                    const uint64_t a = s_key[lo], b = s_key[hi];
                    if ((a > b) == asc) {
                        s_key[lo] = b;
                        s_key[hi] = a;
                    }
                }
                __syncthreads();
            }
        }
        uint32_t carry = 0;
        for (uint32_t base = 0; base < k; base += kThreads) {
            const uint32_t i = base + threadIdx.x;
            bool keep = false;
            uint64_t key = 0;
            if (i < k) {
                key = s_key[i];
                keep = (i + 1 == k) || ((uint32_t)(s_key[i + 1] >> 32) != (uint32_t)(key >> 32));
            }
            uint32_t tot;
            const uint32_t ex = block_excl_scan_u32(keep ? 1u : 0u, &tot);
            if (keep) {
                dcol[s0 + carry + ex] = (uint32_t)(key >> 32);
                dval[s0 + carry + ex] = sval[s0 + (uint32_t)key];
            }
            carry += tot;
        }
        if (threadIdx.x == 0) dcnt[r] = carry;
        __syncthreads();
    }
}

// Rows with more than kMaxSeg staged inserts (any length).  Their number h
// and their insert count H are known on the host when these run.  Every such
// row's inserts are laid out back to back (list order), keyed
// (list index << cb | column) with the insertion index inside the row as the
// payload; one stable LSD radix sort orders them by row, then column, with
// equal columns in insertion order (the order the reference's reverse +
// stable_sort + keep-first walks, aos_array_buffer.hpp:81-96), so the last
// key of each run is the last write.
__global__ void k_long_lens(const uint32_t *__restrict__ rows, uint32_t h, const uint64_t *__restrict__ sptr,
                            uint64_t *__restrict__ lens)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < h) lens[i] = sptr[rows[i] + 1] - sptr[rows[i]];
}

__global__ __launch_bounds__(kThreads) void k_long_keys(uint64_t H, uint32_t h, const uint64_t *__restrict__ hoff,
                                                         const uint32_t *__restrict__ rows,
                                                         const uint64_t *__restrict__ sptr,
                                                         const uint32_t *__restrict__ scol, uint32_t cb,
                                                         uint64_t *__restrict__ key, uint32_t *__restrict__ idx)
{
    const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (p >= H) return;
    const uint64_t i = row_of(hoff, 0, h, p);  // every listed row is non-empty
    const uint64_t j = p - hoff[i];
    key[p] = (i << cb) | scol[sptr[rows[i]] + j];
    idx[p] = (uint32_t)j;
}

__global__ __launch_bounds__(kThreads) void k_long_keep(uint64_t H, const uint64_t *__restrict__ key,
                                                         uint64_t *__restrict__ keep)
{
    const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (p < H) keep[p] = (p + 1 == H || key[p + 1] != key[p]) ? 1u : 0u;
}

// Kept entry p goes to its row's staged base + (kept entries of the row
// before it); P is the exclusive prefix of the keep flags (H + 1 words).
__global__ __launch_bounds__(kThreads) void k_long_scatter(uint64_t H, const uint64_t *__restrict__ key,
                                                            const uint32_t *__restrict__ idx,
                                                            const uint64_t *__restrict__ P,
                                                            const uint64_t *__restrict__ hoff,
                                                            const uint32_t *__restrict__ rows,
                                                            const uint64_t *__restrict__ sptr,
                                                            const double *__restrict__ sval, uint32_t cb,
                                                            uint32_t *__restrict__ dcol, double *__restrict__ dval)
{
    const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (p >= H || P[p + 1] == P[p]) return;
    const uint64_t k = key[p];
    const uint64_t i = k >> cb;
    const uint64_t s0 = sptr[rows[i]];
    const uint64_t rank = P[p] - P[hoff[i]];
    dcol[s0 + rank] = (uint32_t)(k & ((1ull << cb) - 1));
    dval[s0 + rank] = sval[s0 + idx[p]];
}

__global__ void k_long_counts(uint32_t h, const uint32_t *__restrict__ rows, const uint64_t *__restrict__ hoff,
                              const uint64_t *__restrict__ P, uint32_t *__restrict__ dcnt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < h) dcnt[rows[i]] = (uint32_t)(P[hoff[i + 1]] - P[hoff[i]]);
}

// ---------------------------------------------------------------------------
// Row lookup for element-parallel passes: each block narrows the row range of
// its first and last element once, then every element searches that range.
// ---------------------------------------------------------------------------
struct row_range {
    uint64_t lo, hi;  // rows [lo, hi] own the block's elements
};

// Called by every thread of the block (whole wavefronts): each wavefront runs
// the two searches 64 lanes wide.
__device__ __forceinline__ row_range block_rows(const uint64_t *ptr, uint64_t n_rows, uint64_t first,
                                                uint64_t last)
{
    row_range rr;
    rr.lo = wave_row_of(ptr, 0, n_rows, first);
    rr.hi = wave_row_of(ptr, rr.lo, n_rows, last);
    return rr;
}

// The block's row range with its pointers staged in LDS when the range is
// short (the common case); lookups then stay in LDS.  Built by every thread of
// the block before any of them leaves.
constexpr uint32_t kRowCache = 2048;

// The block's sum of v, written by thread 0 to *dst (every thread calls it).
__device__ __forceinline__ void block_sum_to(uint64_t v, uint64_t *dst)
{
    __shared__ uint64_t s_w[kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kThreads / 64; ++w) t += s_w[w];
        *dst = t;
    }
}

struct row_cache {
    uint64_t lo, hi;
    bool cached;
    const uint64_t *s;  // LDS copy of ptr[lo .. hi+1] when cached
    const uint64_t *g;  // the pointer array itself
    __device__ __forceinline__ uint64_t find(uint64_t k) const
    {
        if (!cached) return row_of(g, lo, hi + 1, k);
        uint32_t a = 0, b = (uint32_t)(hi - lo + 1);
        while (b - a > 1) {
            const uint32_t mid = (a + b) >> 1;
            if (s[mid] <= k) a = mid;
            else b = mid;
        }
        return lo + a;
    }
    __device__ __forceinline__ uint64_t at(uint64_t r) const { return cached ? s[r - lo] : g[r]; }  // r in [lo, hi+1]
};

__device__ __forceinline__ row_cache make_row_cache(const row_range rr, const uint64_t *ptr, uint64_t *s_buf)
{
    row_cache rc;
    rc.lo = rr.lo;
    rc.hi = rr.hi;
    rc.g = ptr;
    rc.s = s_buf;
    rc.cached = rr.hi - rr.lo + 1 <= kRowCache;
    if (rc.cached)
        for (uint64_t j = threadIdx.x; j <= rr.hi - rr.lo + 1; j += blockDim.x) s_buf[j] = ptr[rr.lo + j];
    __syncthreads();
    return rc;
}

// The row range of this block from a table built by mark_block_rows.
__device__ __forceinline__ row_range table_rows(const uint64_t *__restrict__ tab)
{
    return row_range{tab[blockIdx.x], tab[blockIdx.x + 1]};
}

// ---------------------------------------------------------------------------
// 3. flags of deduped entries
// ---------------------------------------------------------------------------
// One thread per kItems staged slots.  The row of each slot: the non-empty
// rows of the block mark their first slot with their index in the block's row
// range and a running maximum spreads the marks (blocks touching more than
// kRowCache rows search sptr in global memory).  Index arithmetic inside a
// row is 32-bit (slot counts and old-row lengths are below 2^32).
__global__ __launch_bounds__(kThreads) void k_ded_flags(uint64_t n, uint64_t m, const uint64_t *__restrict__ sptr,
                                                         const uint32_t *__restrict__ dcnt,
                                                         const uint32_t *__restrict__ dcol,
                                                         const double *__restrict__ dval,
                                                         const uint8_t *__restrict__ dirty, int merge,
                                                         const uint64_t *__restrict__ old_off,
                                                         const uint32_t *__restrict__ old_col,
                                                         uint64_t *__restrict__ flags, uint32_t *__restrict__ qpos,
                                                         const uint64_t *__restrict__ stab,
                                                         const uint32_t *__restrict__ ctl,
                                                         uint64_t *__restrict__ chunk_sums)
{
    static_assert(kRowCache <= 65536, "16-bit local rows");
    __shared__ uint16_t s_rid[kPerBlock];  // local row of each slot of the block
    __shared__ uint32_t s_w[kThreads / kWave];
    const uint64_t first = (uint64_t)blockIdx.x * kPerBlock;
    if (first >= m || ctl[CTL_ERR]) return;
    const uint32_t nq = (uint32_t)(min<uint64_t>(first + kPerBlock, m) - first);
    // slot data first: these loads are in flight while the row marks fill
    uint32_t cs[kItems];
    double vs[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t q = (uint32_t)it * kThreads + threadIdx.x;
        cs[it] = q < nq ? dcol[first + q] : 0u;
        vs[it] = q < nq ? dval[first + q] : 0.0;
    }
    const row_range rr = table_rows(stab);
    const uint64_t nr = rr.hi - rr.lo + 1;
    const bool cached = nr <= kRowCache;  // the same for every thread of the block
    if (cached) {
        for (uint32_t q = threadIdx.x; q < nq; q += kThreads) s_rid[q] = 0;  // row lo owns slot `first`
        __syncthreads();
        // rows after lo start after `first`; a non-empty one marks its first slot
        for (uint32_t j = threadIdx.x + 1; j < nr; j += kThreads) {
            const uint64_t a = sptr[rr.lo + j];
            if (a < first + nq && sptr[rr.lo + j + 1] > a) s_rid[a - first] = (uint16_t)j;
        }
        __syncthreads();
        // running maximum over the kItems consecutive slots of each thread, then across threads
        const uint32_t a0 = threadIdx.x * kItems;
        uint32_t v[kItems], mx = 0;
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
            v[i] = a0 + i < nq ? s_rid[a0 + i] : 0u;
            mx = max(mx, v[i]);
            v[i] = mx;
        }
        const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
        uint32_t inc = mx;
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
        for (int i = 0; i < kItems; ++i)
            if (a0 + i < nq) s_rid[a0 + i] = (uint16_t)max(v[i], ex);
        __syncthreads();
    }
    uint64_t fsum = 0;  // this thread's flags, for the block's chunk sum
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t q = (uint32_t)it * kThreads + threadIdx.x;
        if (q >= nq) break;
        const uint64_t s = first + q;
        const uint64_t r = cached ? rr.lo + s_rid[q] : row_of(sptr, rr.lo, rr.hi + 1, s);
        const uint32_t j = (uint32_t)s - (uint32_t)sptr[r];  // slot in the row (< 2^32)
        uint64_t f = 0;
        uint32_t qp = 0;
        if (j < dcnt[r]) {
            const bool d = merge ? row_dirty(dirty, sptr, r) : true;
            const uint32_t c = cs[it];
            const bool nz = d && !is_zero_f64(vs[it]);
            bool match = false;
            if (merge && d) {
                const uint64_t ob = old_off[r];
                const uint32_t *row = old_col + ob;
                uint32_t lo = 0, hi = (uint32_t)(old_off[r + 1] - ob);  // lower bound of c in the old row
                const uint32_t len = hi;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (row[mid] < c) lo = mid + 1;
                    else hi = mid;
                }
                match = lo < len && row[lo] == c;
                qp = lo;
            }
            f = ((uint64_t)(match ? 1u : 0u) << 32) | (nz ? 1u : 0u);
        }
        flags[s] = f;
        fsum += f;
        if (merge) qpos[s] = qp;
    }
    block_sum_to(fsum, chunk_sums + blockIdx.x);  // the scan's chunk of kPerBlock flags
}

// Old entries that a dirty row drops because they are explicit zeros (and not
// overwritten): z = 1.  Only run when the old CSR can hold zeros.
__global__ __launch_bounds__(kThreads) void k_old_zero_flags(uint64_t n, uint64_t nnz_old,
                                                              const uint64_t *__restrict__ old_off,
                                                              const uint32_t *__restrict__ old_col,
                                                              const double *__restrict__ old_val,
                                                              const uint64_t *__restrict__ sptr,
                                                              const uint32_t *__restrict__ dcnt,
                                                              const uint32_t *__restrict__ dcol,
                                                              const uint8_t *__restrict__ dirty,
                                                              uint64_t *__restrict__ zf,
                                                              const uint32_t *__restrict__ ctl)
{
    __shared__ uint64_t s_rows[kRowCache + 1];
    const uint64_t first = (uint64_t)blockIdx.x * kThreads * kItems;
    if (first >= nnz_old || ctl[CTL_ERR]) return;
    const uint64_t last = min(first + kThreads * kItems, nnz_old) - 1;
    const row_cache rc = make_row_cache(block_rows(old_off, n, first, last), old_off, s_rows);
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t k = first + (uint64_t)it * kThreads + threadIdx.x;
        if (k >= nnz_old) break;
        const uint64_t r = rc.find(k);
        uint64_t z = 0;
        if (row_dirty(dirty, sptr, r) && is_zero_f64(old_val[k])) {
            const uint64_t db = sptr[r], de = db + dcnt[r];
            const uint32_t c = old_col[k];
            const uint64_t p = lower_bound_u32(dcol, db, de, c);
            const bool over = p < de && dcol[p] == c;
            z = over ? 0 : 1;
        }
        zf[k] = z;
    }
}

// ---------------------------------------------------------------------------
// 4. row counts
// ---------------------------------------------------------------------------
// One thread per row; each block also writes the sum of its kThreads counts
// (the offsets scan's chunk).
__global__ __launch_bounds__(kThreads) void k_row_counts(uint64_t n, const uint64_t *__restrict__ sptr,
                                                          const uint64_t *__restrict__ P,
                                                          const uint8_t *__restrict__ dirty, int merge,
                                                          const uint64_t *__restrict__ old_off,
                                                          const uint64_t *__restrict__ PZ, uint64_t *__restrict__ cnt,
                                                          const uint32_t *__restrict__ ctl,
                                                          uint64_t *__restrict__ chunk_sums)
{
    if (ctl[CTL_ERR]) return;  // the same word for every block: the whole grid returns
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t c = 0;
    if (r < n) {
        const uint64_t pa = P[sptr[r]], pb = P[sptr[r + 1]];
        const uint64_t nz = (pb & 0xffffffffull) - (pa & 0xffffffffull);
        const uint64_t mt = (pb >> 32) - (pa >> 32);
        if (!merge) {
            c = nz;
        } else {
            const uint64_t ob = old_off[r], oe = old_off[r + 1];
            if (!row_dirty(dirty, sptr, r)) {
                c = oe - ob;
            } else {
                const uint64_t zeros = PZ ? (PZ[oe] - PZ[ob]) : 0;
                c = (oe - ob) - mt - zeros + nz;
            }
        }
        cnt[r] = c;
    }
    block_sum_to(c, chunk_sums + blockIdx.x);
}

// ---------------------------------------------------------------------------
// 5. scatter passes
// ---------------------------------------------------------------------------
constexpr uint32_t kSoRows = 384;   // rows of one re-lock scatter block staged in LDS
constexpr uint32_t kSoIns = 1024;   // deduped inserts of those rows staged in LDS (< 2^16: 16-bit counts)

// Largest r in [lo, hi) with old_off[r] + sptr[r] <= p (both non-decreasing).
__device__ __forceinline__ uint64_t row_of_key(const uint64_t *old_off, const uint64_t *sptr, uint64_t lo, uint64_t hi,
                                               uint64_t p)
{
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (old_off[mid] + sptr[mid] <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Re-lock scatter (merge_csr, csr_build.hpp:126-488) as one element-parallel
// pass over a combined index space in which row r holds its old entries and
// then its staged slots, starting at key(r) = old_off[r] + sptr[r].  Every old
// entry and every surviving deduped insert computes its own output slot from
// the prefix sums, and both parts of a row's new range are written by the
// same block close together in time, so its cache lines leave L2 whole.
//
// Position p of row r is old entry k = p - sptr[r] when k < old_off[r + 1],
// else deduped slot s = p - old_off[r + 1] (a surviving insert when the row is
// dirty and s - sptr[r] < dcnt[r]; later slots hold nothing).  A block whose
// rows and deduped inserts fit in LDS (the common case) stages the rows
// (sptr and old-end offsets relative to the block, new offsets, dirty bit,
// deduped count), their deduped columns and 16-bit prefix counts and the row
// of each position, and does its index arithmetic in 32 bits; other blocks
// read the same quantities from global memory (scatter_merge_global).

// One block of the re-lock scatter with every row quantity read from global
// memory (a binary search finds each position's row).
__device__ __forceinline__ void scatter_merge_global(uint64_t first, uint64_t last, uint64_t rlo, uint64_t rhi,
                                                  const uint64_t *__restrict__ old_off,
                                                  const uint32_t *__restrict__ old_col,
                                                  const double *__restrict__ old_val,
                                                  const uint64_t *__restrict__ sptr,
                                                  const uint32_t *__restrict__ dcnt,
                                                  const uint32_t *__restrict__ dcol,
                                                  const double *__restrict__ dval, const uint64_t *__restrict__ P,
                                                  const uint32_t *__restrict__ qpos, const uint64_t *__restrict__ PZ,
                                                  const uint8_t *__restrict__ dirty,
                                                  const uint64_t *__restrict__ new_off, uint32_t *__restrict__ new_col,
                                                  double *__restrict__ new_val, uint32_t *__restrict__ err)
{
    for (int it = 0; it < kItemsMerge; ++it) {
        const uint64_t p = first + (uint64_t)it * kThreads + threadIdx.x;
        if (p > last) break;
        const uint64_t r = row_of_key(old_off, sptr, rlo, rhi + 1, p);
        const uint64_t ob = old_off[r], db = sptr[r], key = ob + db, olen = old_off[r + 1] - ob;
        const uint64_t no = new_off[r], nlen = new_off[r + 1] - no;
        const bool dty = row_dirty(dirty, sptr, r);
        const uint64_t dn = dcnt[r];
        const uint64_t i = p - key;
        uint64_t out;
        uint32_t c;
        double v;
        if (i < olen) {
            const uint64_t k = ob + i;
            c = old_col[k];
            v = old_val[k];
            if (!dty) {
                out = no + i;  // clean row: verbatim (csr_build.hpp:362-377)
            } else {
                const uint64_t de = db + dn;
                uint64_t lo = db, hi = de;  // lower bound of c in the row's deduped columns
                while (lo < hi) {
                    const uint64_t mid = (lo + hi) >> 1;
                    if (dcol[mid] < c) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < de && dcol[lo] == c) continue;  // buffer wins (:430-436)
                if (PZ && is_zero_f64(v)) continue;      // old zero dropped (:420-421)
                const uint64_t a = P[db], b = P[lo];
                const uint64_t before_z = PZ ? (PZ[k] - PZ[ob]) : 0;
                out = no + i - (uint32_t)((b >> 32) - (a >> 32)) - before_z + (uint32_t)((uint32_t)b - (uint32_t)a);
            }
        } else if (dty && i - olen < dn) {  // clean rows keep no insert; slots past the run hold nothing
            const uint64_t s = db + (i - olen);
            const uint64_t a = P[db], sw = P[s], tw = P[s + 1];
            if ((uint32_t)tw == (uint32_t)sw) continue;  // zero insert: a delete marker, not an entry
            c = dcol[s];
            v = dval[s];
            const uint64_t q = qpos[s];  // old entries of the row before this column
            const uint64_t before_z = PZ ? (PZ[ob + q] - PZ[ob]) : 0;
            out = no + (uint32_t)((uint32_t)sw - (uint32_t)a) + q - (uint32_t)((sw >> 32) - (a >> 32)) - before_z;
        } else {
            continue;
        }
        if (out - no >= nlen) {  // every slot must fall inside the row's new range; never write outside
            atomicOr(err, ERR_INTERNAL);
            continue;
        }
        new_col[out] = c;
        new_val[out] = v;
    }
}

__global__ __launch_bounds__(kThreads) void k_scatter_merge(uint64_t total, const uint64_t *__restrict__ old_off,
                                                             const uint32_t *__restrict__ old_col,
                                                             const double *__restrict__ old_val,
                                                             const uint64_t *__restrict__ sptr,
                                                             const uint32_t *__restrict__ dcnt,
                                                             const uint32_t *__restrict__ dcol,
                                                             const double *__restrict__ dval,
                                                             const uint64_t *__restrict__ P,
                                                             const uint32_t *__restrict__ qpos,
                                                             const uint64_t *__restrict__ PZ,
                                                             const uint8_t *__restrict__ dirty,
                                                             const uint64_t *__restrict__ new_off,
                                                             uint32_t *__restrict__ new_col,
                                                             double *__restrict__ new_val, uint32_t *__restrict__ err,
                                                             const uint64_t *__restrict__ mtab)
{
    __shared__ uint64_t s_no[kSoRows + 1];   // new_off[lo .. hi+1]
    __shared__ int64_t s_ob[kSoRows];        // old_off[r] - K0 (old index of the row's first entry, block-relative)
    __shared__ uint32_t s_sp[kSoRows];       // sptr[r] - ibase
    __shared__ int32_t s_oe[kSoRows];        // old_off[r + 1] - K0, clamped to +-2^30
    __shared__ uint8_t s_dirty[kSoRows];
    __shared__ uint16_t s_dn[kSoRows];       // dcnt[r]
    __shared__ uint32_t s_dcol[kSoIns];      // dcol[ibase ..)
    __shared__ uint32_t s_P[kSoIns + 1];     // (matches << 16 | nonzeros) of P[ibase + t] - P[ibase]
    __shared__ uint16_t s_rid[kPerBlockMerge];  // local row of each position of the block
    static_assert(kSoRows <= 65536 && kSoIns < 65536, "local rows and 16-bit prefix counts");
    const uint64_t first = (uint64_t)blockIdx.x * kPerBlockMerge;
    if (first >= total || *err) return;
    const uint64_t last = min(first + kPerBlockMerge, total) - 1;
    const row_range rr = table_rows(mtab);
    const uint64_t nr = rr.hi - rr.lo + 1;
    const uint64_t ibase = sptr[rr.lo];
    const uint64_t ni = sptr[rr.hi] + dcnt[rr.hi] - ibase;  // deduped inserts of rows lo..hi lie in [ibase, ibase+ni)
    if (nr > kSoRows || ni > kSoIns) {  // the same for every thread of the block
        scatter_merge_global(first, last, rr.lo, rr.hi, old_off, old_col, old_val, sptr, dcnt, dcol, dval, P, qpos,
                             PZ, dirty, new_off, new_col, new_val, err);
        return;
    }
    // Block-relative indices: position q = p - first; old entry k = K0 + d
    // with d = q - (sptr[r] - ibase) in [-kSoIns, kPerBlockMerge); deduped slot
    // s = ibase + (q - s_oe[r]).  first >= key(lo) >= sptr[lo], so K0 >= 0.
    const uint64_t K0 = first - ibase;
    for (uint32_t j = threadIdx.x; j <= nr; j += kThreads) {
        const uint64_t r = rr.lo + j;
        s_no[j] = new_off[r];
        if (j < nr) {
            const uint64_t oo = old_off[r], oe = old_off[r + 1], sp = sptr[r];
            const int64_t oe_rel = (int64_t)(oe - K0);
            s_ob[j] = (int64_t)(oo - K0);
            s_sp[j] = (uint32_t)(sp - ibase);
            s_oe[j] = (int32_t)max<int64_t>(-(1ll << 30), min<int64_t>(oe_rel, 1ll << 30));
            s_dirty[j] = row_dirty(dirty, sptr, r) ? 1 : 0;
            s_dn[j] = (uint16_t)dcnt[r];
            // the block's positions that row j owns
            const uint64_t a = max(oo + sp, first), b = min(oe + sptr[r + 1], last + 1);
            for (uint64_t p = a; p < b; ++p) s_rid[p - first] = (uint16_t)j;
        }
    }
    const uint64_t p0 = P[ibase];
    for (uint32_t t = threadIdx.x; t <= ni; t += kThreads) {
        const uint64_t pt = P[ibase + t];
        const uint32_t dm = (uint32_t)(pt >> 32) - (uint32_t)(p0 >> 32);
        const uint32_t dz = (uint32_t)pt - (uint32_t)p0;
        s_P[t] = (dm << 16) | dz;
        if (t < ni) s_dcol[t] = dcol[ibase + t];
    }
    __syncthreads();
    const uint32_t nq = (uint32_t)(last - first + 1);

    // Pass 1: classify every position and issue the loads of its entry.
    uint32_t cs[kItemsMerge];
    double vs[kItemsMerge];
    uint32_t kinds = 0;  // 2 bits per position: 0 nothing, 1 old entry, 2 deduped insert
#pragma unroll
    for (int it = 0; it < kItemsMerge; ++it) {
        const uint32_t q = (uint32_t)it * kThreads + threadIdx.x;
        cs[it] = 0;
        vs[it] = 0.0;
        if (q >= nq) continue;
        const uint32_t j = s_rid[q];
        const int32_t sp = (int32_t)s_sp[j], oe = s_oe[j];
        const int32_t d = (int32_t)q - sp;
        if (d < oe) {
            const uint64_t k = K0 + (int64_t)d;
            kinds |= 1u << (2 * it);
            cs[it] = ld_stream(old_col + k);
            vs[it] = ld_stream(old_val + k);
        } else {
            const int32_t srel = (int32_t)q - oe;  // deduped slot, relative to ibase
            if (s_dirty[j] && srel - sp < (int32_t)s_dn[j]) {  // clean rows keep no insert; slots past the run hold nothing
                kinds |= 2u << (2 * it);
                cs[it] = s_dcol[srel];
                vs[it] = dval[ibase + (uint32_t)srel];
            }
        }
    }
    // Pass 2: output slots and stores.
#pragma unroll
    for (int it = 0; it < kItemsMerge; ++it) {
        const uint32_t kind = (kinds >> (2 * it)) & 3u;
        if (kind == 0) continue;
        const uint32_t q = (uint32_t)it * kThreads + threadIdx.x;
        const uint32_t j = s_rid[q];
        const uint32_t sp = s_sp[j];
        const uint64_t no = s_no[j], nlen = s_no[j + 1] - no;
        const uint32_t c = cs[it];
        const double v = vs[it];
        const uint32_t wa = s_P[sp];  // prefix counts at the row's first deduped slot
        uint64_t out;
        if (kind == 1) {
            const int32_t d = (int32_t)q - (int32_t)sp;
            const uint64_t i = (uint64_t)((int64_t)d - s_ob[j]);  // index in the row's old entries
            if (!s_dirty[j]) {
                out = no + i;  // clean row: verbatim (csr_build.hpp:362-377)
            } else {
                const uint32_t de = sp + s_dn[j];
                uint32_t lo = sp, hi = de;  // lower bound of c in the row's deduped columns
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_dcol[mid] < c) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < de && s_dcol[lo] == c) continue;  // buffer wins (:430-436)
                if (PZ && is_zero_f64(v)) continue;        // old zero dropped (:420-421)
                const uint32_t wb = s_P[lo];
                const uint64_t ob = K0 + (uint64_t)s_ob[j];
                const uint64_t before_z = PZ ? (PZ[ob + i] - PZ[ob]) : 0;
                out = no + i - ((wb >> 16) - (wa >> 16)) - before_z + ((wb & 0xffffu) - (wa & 0xffffu));
            }
        } else {
            const uint32_t srel = q - (uint32_t)s_oe[j];
            const uint32_t ws = s_P[srel], wt = s_P[srel + 1];
            if ((wt & 0xffffu) == (ws & 0xffffu)) continue;  // zero insert: a delete marker, not an entry
            const uint64_t qo = qpos[ibase + srel];  // old entries of the row before this column
            const uint64_t ob = K0 + (uint64_t)s_ob[j];
            const uint64_t before_z = PZ ? (PZ[ob + qo] - PZ[ob]) : 0;
            out = no + ((ws & 0xffffu) - (wa & 0xffffu)) + qo - ((ws >> 16) - (wa >> 16)) - before_z;
        }
        // Every slot must fall inside the row's new range; never write outside.
        if (out - no >= nlen) {
            atomicOr(err, ERR_INTERNAL);
            continue;
        }
        new_col[out] = c;
        new_val[out] = v;
    }
}

// First lock (build_csr, csr_build.hpp:32-98): the surviving deduped inserts
// are the whole matrix; each finds its slot from the nonzero prefix.
__global__ __launch_bounds__(kThreads) void k_scatter_ded(uint64_t m, const uint64_t *__restrict__ sptr,
                                                           const uint32_t *__restrict__ dcol,
                                                           const double *__restrict__ dval,
                                                           const uint64_t *__restrict__ P,
                                                           const uint64_t *__restrict__ new_off,
                                                           uint32_t *__restrict__ new_col, double *__restrict__ new_val,
                                                           uint32_t *__restrict__ err, const uint64_t *__restrict__ stab)
{
    __shared__ uint64_t s_rows[kRowCache + 1];
    const uint64_t first = (uint64_t)blockIdx.x * kPerBlock;
    if (first >= m || *err) return;
    // prefix words first: these loads are in flight while the row cache fills
    uint64_t pw[kItems], pw1[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t s = first + (uint64_t)it * kThreads + threadIdx.x;
        pw[it] = s < m ? P[s] : 0;
        pw1[it] = s < m ? P[s + 1] : 0;  // P holds m + 1 words
    }
    const row_cache rc = make_row_cache(table_rows(stab), sptr, s_rows);
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t s = first + (uint64_t)it * kThreads + threadIdx.x;
        if (s >= m) break;
        const uint64_t ps = pw[it];
        const uint64_t ps1 = pw1[it];
        if (((ps1 - ps) & 0xffffffffull) == 0) continue;  // not a surviving entry
        const uint64_t r = rc.find(s);
        const uint64_t db = rc.at(r);
        const uint64_t pa = P[db];
        const uint64_t before_nz = (ps & 0xffffffffull) - (pa & 0xffffffffull);
        const uint64_t out = new_off[r] + before_nz;
        if (out - new_off[r] >= new_off[r + 1] - new_off[r]) {
            atomicOr(err, ERR_INTERNAL);
            continue;
        }
        new_col[out] = dcol[s];
        new_val[out] = dval[s];
    }
}

// The new nnz next to the error bits, so one copy brings both back.
// ---------------------------------------------------------------------------
// load_csr validation (matrix.hpp:524-575), first failure in the reference's
// check order wins: key = row << 32 | step, step 0 = offsets of the row,
// step 1 + 2t = bounds of entry t, step 2 + 2t = order of entry t.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_validate_rows(uint64_t n, uint64_t n_cols,
                                                             const uint64_t *__restrict__ off,
                                                             unsigned long long *__restrict__ first_bad)
{
    const uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    if (off[r] > off[r + 1]) atomicMin(first_bad, (unsigned long long)(r << 32));
}

__global__ __launch_bounds__(kThreads) void k_validate_entries(uint64_t n, uint64_t nnz, uint64_t n_cols,
                                                                const uint64_t *__restrict__ off,
                                                                const uint32_t *__restrict__ col,
                                                                const double *__restrict__ val,
                                                                unsigned long long *__restrict__ first_bad,
                                                                uint32_t *__restrict__ zero_seen)
{
    __shared__ uint64_t s_rows[kRowCache + 1];
    const uint64_t first = (uint64_t)blockIdx.x * kThreads * kItems;
    if (first >= nnz) return;
    const uint64_t last = min(first + kThreads * kItems, nnz) - 1;
    // Entry data first (in flight during the row search).  Only an entry that
    // is out of range or not above its predecessor needs its row: to tell a
    // row start from an order violation, and for the failure key.
    uint32_t cur[kItems], prev[kItems];
    bool z = false;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t k = first + (uint64_t)it * kThreads + threadIdx.x;
        cur[it] = k < nnz ? col[k] : 0u;
        prev[it] = k < nnz && k > 0 ? col[k - 1] : 0u;
        z |= k < nnz && is_zero_f64(val[k]);
    }
    const row_cache rc = make_row_cache(block_rows(off, n, first, last), off, s_rows);
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t k = first + (uint64_t)it * kThreads + threadIdx.x;
        if (k >= nnz) break;
        const bool oob = (uint64_t)cur[it] >= n_cols;
        if (!oob && (k == 0 || cur[it] > prev[it])) continue;
        const uint64_t r = rc.find(k);
        const uint64_t t = k - rc.at(r);
        const uint64_t tt = t < 0x7fffffffull ? t : 0x7fffffffull;
        if (oob) atomicMin(first_bad, (unsigned long long)((r << 32) | (1 + 2 * tt)));
        else if (t > 0) atomicMin(first_bad, (unsigned long long)((r << 32) | (2 + 2 * tt)));
    }
    if (z) atomicOr(zero_seen, 1u);
}

}  // namespace

void validate_csr_device(spira_dcsr *A, uint64_t nnz, const uint64_t *d_off, const uint32_t *d_col,
                         const double *d_val, bool *has_zeros)
{
    spira_dev *dev = A->dev;
    const uint64_t n = A->n_rows;
    uint64_t ends[2];
    {  // both words into the pinned buffer, one wait
        uint64_t *hw = (uint64_t *)dev->host_words;
        SPIRA_HIP(hipMemcpyAsync(hw, d_off, sizeof(uint64_t), hipMemcpyDeviceToHost, dev->stream));
        SPIRA_HIP(hipMemcpyAsync(hw + 1, d_off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, dev->stream));
        dev_sync(dev);
        ends[0] = hw[0];
        ends[1] = hw[1];
    }
    if (ends[0] != 0) fail(SPIRA_ERR_INVALID_ARGUMENT, "spira::matrix::load_csr: CSR offsets[0] must be zero");
    if (ends[1] != nnz)
        fail(SPIRA_ERR_INVALID_ARGUMENT, "spira::matrix::load_csr: CSR offsets[n_rows] does not equal nnz");
    unsigned long long *bad = (unsigned long long *)dev->ws.get(WS_MISC, 2 * sizeof(unsigned long long));
    uint32_t *zs = (uint32_t *)(bad + 1);
    const unsigned long long init[2] = {~0ull, 0ull};
    SPIRA_HIP(hipMemcpyAsync(bad, init, sizeof(init), hipMemcpyHostToDevice, dev->stream));
    if (n > 0) {
        k_validate_rows<<<grid_for(n, kThreads), kThreads, 0, dev->stream>>>(n, A->n_cols, d_off, bad);
        SPIRA_HIP(hipGetLastError());
    }
    // Entries are only checked once offsets are known to be monotone (the
    // reference stops at the first bad row as well).
    unsigned long long h[2];
    read_back(dev, h, bad, sizeof(h));
    if (h[0] == ~0ull && nnz > 0) {
        k_validate_entries<<<grid_for(nnz, kThreads * kItems), kThreads, 0, dev->stream>>>(n, nnz, A->n_cols, d_off,
                                                                                           d_col, d_val, bad, zs);
        SPIRA_HIP(hipGetLastError());
        read_back(dev, h, bad, sizeof(h));
    }
    if (h[0] != ~0ull) {
        const uint64_t step = h[0] & 0xffffffffull;
        if (step == 0) fail(SPIRA_ERR_INVALID_ARGUMENT, "spira::matrix::load_csr: CSR offsets are not non-decreasing");
        if (step & 1ull) fail(SPIRA_ERR_INVALID_ARGUMENT, "spira::matrix::load_csr: column index out of bounds");
        fail(SPIRA_ERR_INVALID_ARGUMENT,
             "spira::matrix::load_csr: CSR row entries are not strictly sorted by column");
    }
    uint32_t zflag = 0;
    std::memcpy(&zflag, &h[1], sizeof(uint32_t));
    *has_zeros = zflag != 0;
}

namespace {

uint32_t bits_for(uint64_t v)  // bits needed to hold every value in [0, v]
{
    uint32_t b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

// Dedup of the h listed rows with more than kMaxSeg staged inserts (H in all):
// deduped (column, value) pairs at each row's staged base and dcnt[row].
void dedup_long_rows(spira_dev *dev, uint64_t n_cols, const uint32_t *rows, uint32_t h, uint64_t H,
                     const uint64_t *sptr, const uint32_t *scol, const double *sval, uint32_t *dcol, double *dval,
                     uint32_t *dcnt)
{
    hipStream_t st = dev->stream;
    const uint32_t cb = std::max<uint32_t>(1, bits_for(n_cols ? n_cols - 1 : 0));
    const uint32_t kb = cb + bits_for(h - 1);  // key bits: list index, then column
    if (kb > 64) fail(SPIRA_ERR_UNSUPPORTED, "lock: too many long rows for the long-row sort");
    uint64_t *hoff = (uint64_t *)dev->ws.get(WS_LONG_OFF, (h + 1) * sizeof(uint64_t));
    uint64_t *key = (uint64_t *)dev->ws.get(WS_LONG_KEY, 2 * H * sizeof(uint64_t));
    uint32_t *idx = (uint32_t *)dev->ws.get(WS_LONG_IDX, 2 * H * sizeof(uint32_t));
    uint64_t *P = (uint64_t *)dev->ws.get(WS_LONG_P, (H + 1) * sizeof(uint64_t));
    k_long_lens<<<grid_for(h, kThreads), kThreads, 0, st>>>(rows, h, sptr, hoff);
    SPIRA_HIP(hipGetLastError());
    scan_excl_u64(dev, hoff, hoff, h);
    k_long_keys<<<grid_for(H, kThreads), kThreads, 0, st>>>(H, h, hoff, rows, sptr, scol, cb, key, idx);
    SPIRA_HIP(hipGetLastError());
    size_t tb = 0;
    SPIRA_HIP(rocprim::radix_sort_pairs(nullptr, tb, key, key + H, idx, idx + H, (unsigned)H, 0, kb, st));
    void *tmp = dev->ws.get(WS_SORT, tb ? tb : 8);
    SPIRA_HIP(rocprim::radix_sort_pairs(tmp, tb, key, key + H, idx, idx + H, (unsigned)H, 0, kb, st));
    k_long_keep<<<grid_for(H, kThreads), kThreads, 0, st>>>(H, key + H, P);
    SPIRA_HIP(hipGetLastError());
    scan_excl_u64(dev, P, P, H);
    k_long_scatter<<<grid_for(H, kThreads), kThreads, 0, st>>>(H, key + H, idx + H, P, hoff, rows, sptr, sval, cb,
                                                               dcol, dval);
    SPIRA_HIP(hipGetLastError());
    k_long_counts<<<grid_for(h, kThreads), kThreads, 0, st>>>(h, rows, hoff, P, dcnt);
    SPIRA_HIP(hipGetLastError());
}

}  // namespace

uint64_t lock_device_f64(spira_dcsr *A, const uint64_t *sptr, const uint32_t *scol, const double *sval,
                         uint64_t m, const uint8_t *dirty)
{
    spira_dev *dev = A->dev;
    hipStream_t st = dev->stream;
    const uint64_t n = A->n_rows;
    const bool merge = A->built;
    if (m >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "lock: more than 2^32-1 staged inserts in one call");

    // ---- 1. checks + binning -------------------------------------------------
    // ctl (ctl_word): error bits, listed rows, new nnz.  There is no host round
    // trip here: every later kernel returns at once when a bit is set, and the
    // host reads the words with the new nnz at the end, so a failed check
    // still leaves the matrix untouched.
    uint32_t *ctl = (uint32_t *)dev->ws.get(WS_MISC, CTL_WORDS * sizeof(uint32_t));
    uint32_t *big = (uint32_t *)dev->ws.get(WS_BIG, (n + 1) * sizeof(uint32_t));
    uint32_t *long_rows = (uint32_t *)dev->ws.get(WS_LONG_ROWS, (m / (kMaxSeg + 1) + 1) * sizeof(uint32_t));
    SPIRA_HIP(hipMemsetAsync(ctl, 0, CTL_WORDS * sizeof(uint32_t), st));
    k_check_staged<<<grid_for(std::max<uint64_t>(n + 1, m), kThreads), kThreads, 0, st>>>(sptr, n, m, scol, A->n_cols,
                                                                                     ctl, big, long_rows);
    SPIRA_HIP(hipGetLastError());

    uint32_t *dcol = (uint32_t *)dev->ws.get(WS_DED_COL, m * sizeof(uint32_t));
    double *dval = (double *)dev->ws.get(WS_DED_VAL, m * sizeof(double));
    uint32_t *dcnt = (uint32_t *)dev->ws.get(WS_ROWS, (n + 1) * sizeof(uint32_t) + 8);
    // Row tables (mark_block_rows) of the staged-slot passes and of the
    // old-entry scatter, filled by the dedup pass that reads both pointers.
    const uint64_t total_merge = merge ? A->nnz + m : 0;  // re-lock scatter positions: old entries + staged slots
    const uint64_t nb_s = (m + kPerBlock - 1) / kPerBlock + 1;
    const uint64_t nb_m = total_merge ? (total_merge + kPerBlockMerge - 1) / kPerBlockMerge + 1 : 0;
    uint64_t *stab = (uint64_t *)dev->ws.get(WS_BLK, (nb_s + nb_m) * sizeof(uint64_t));
    uint64_t *mtab = total_merge ? stab + nb_s : nullptr;
    uint64_t *P = (uint64_t *)dev->ws.get(WS_FLAGS, (m + 1) * sizeof(uint64_t));
    uint32_t *qpos = merge ? (uint32_t *)dev->ws.get(WS_QPOS, m * sizeof(uint32_t)) : nullptr;
    const bool zeros = merge && A->has_zeros && A->nnz > 0;
    uint64_t *PZ = zeros ? (uint64_t *)dev->ws.get(WS_PZ, (A->nnz + 1) * sizeof(uint64_t)) : nullptr;
    uint64_t *cnt = (uint64_t *)dev->ws.get(WS_CNT, (n + 1) * sizeof(uint64_t));

    // The new arrays are sized by the upper bound nnz_old + m (as merge_csr's
    // total_ub, csr_build.hpp:139-145) so no host round trip is needed before
    // the scatter; the exact nnz comes back with the final status word.
    const uint64_t ub = (merge ? A->nnz : 0) + m;
    const size_t ob = (n + 1) * sizeof(uint64_t), cb = (ub ? ub : 1) * sizeof(uint32_t),
                 vb = (ub ? ub : 1) * sizeof(double);
    uint64_t *new_off = (uint64_t *)dev_alloc(dev, ob);
    uint32_t *new_col = nullptr;
    double *new_val = nullptr;
    try {
        new_col = (uint32_t *)dev_alloc(dev, cb);
        new_val = (double *)dev_alloc(dev, vb);

        // Steps 2-5; every kernel returns at once while an error (or the
        // long-row retry) bit is set.
        uint32_t hctl[CTL_WORDS];
        auto steps = [&]() {
            // ---- 2. dedup -------------------------------------------------
            if (n > 0) {
                k_dedup_small<<<grid_for(n, kThreads), kThreads, 0, st>>>(n, sptr, scol, sval, dcol, dval, dcnt,
                                                                           stab, A->off, mtab, ctl);
                SPIRA_HIP(hipGetLastError());
                // the number of listed rows is on the device: blocks stride over it
                const unsigned g = (unsigned)std::min<uint64_t>(n, 2048u);
                k_dedup_big<<<g, kThreads, kMaxSeg * sizeof(uint64_t), st>>>(big, ctl, sptr, scol, sval, dcol,
                                                                              dval, dcnt);
                SPIRA_HIP(hipGetLastError());
            }
            // ---- 3. flags + scan ------------------------------------------
            uint64_t *psums = scan_chunk_sums(dev, m, kPerBlock);
            if (m > 0) {
                k_ded_flags<<<grid_for(m, kPerBlock), kThreads, 0, st>>>(n, m, sptr, dcnt, dcol, dval, dirty,
                                                                         merge ? 1 : 0, A->off, A->col, P, qpos,
                                                                         stab, ctl, psums);
                SPIRA_HIP(hipGetLastError());
            }
            scan_excl_u64_chunks(dev, P, P, m, psums, kPerBlock);
            // Old explicit zeros (only possible after load_csr with zeros).
            if (zeros) {
                k_old_zero_flags<<<grid_for(A->nnz, kThreads * kItems), kThreads, 0, st>>>(
                    n, A->nnz, A->off, A->col, A->val, sptr, dcnt, dcol, dirty, PZ, ctl);
                SPIRA_HIP(hipGetLastError());
                scan_excl_u64(dev, PZ, PZ, A->nnz);
            }
            // ---- 4. counts + offsets --------------------------------------
            uint64_t *csums = scan_chunk_sums(dev, n, kThreads);
            if (n > 0) {
                k_row_counts<<<grid_for(n, kThreads), kThreads, 0, st>>>(n, sptr, P, dirty, merge ? 1 : 0, A->off,
                                                                         PZ, cnt, ctl, csums);
                SPIRA_HIP(hipGetLastError());
            }
            // new offsets; the total (the new nnz) also goes to the control words
            scan_excl_u64_chunks(dev, cnt, new_off, n, csums, kThreads, (uint64_t *)(ctl + CTL_NNZ_LO));
            // ---- 5. scatter ------------------------------------------------
            if (total_merge > 0) {
                k_scatter_merge<<<grid_for(total_merge, kPerBlockMerge), kThreads, 0, st>>>(
                    total_merge, A->off, A->col, A->val, sptr, dcnt, dcol, dval, P, qpos, PZ, dirty, new_off,
                    new_col, new_val, ctl, mtab);
                SPIRA_HIP(hipGetLastError());
            } else if (!merge && m > 0) {
                k_scatter_ded<<<grid_for(m, kPerBlock), kThreads, 0, st>>>(m, sptr, dcol, dval, P, new_off,
                                                                           new_col, new_val, ctl, stab);
                SPIRA_HIP(hipGetLastError());
            }
            read_back(dev, hctl, ctl, sizeof(hctl));
        };
        steps();
        if (hctl[CTL_ERR] == ERR_LONG_ROWS) {
            // Rows longer than the LDS sort: sort them, clear the bit, rerun.
            dedup_long_rows(dev, A->n_cols, long_rows, hctl[CTL_NLONG], hctl[CTL_LONG_INS], sptr, scol, sval, dcol,
                            dval, dcnt);
            SPIRA_HIP(hipMemsetAsync(ctl + CTL_ERR, 0, sizeof(uint32_t), st));
            steps();
        }
        if (hctl[CTL_ERR] & ERR_BAD_PTR)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "lock: stage_ptr must start at 0, be non-decreasing and end at m");
        if (hctl[CTL_ERR] & ERR_OUT_OF_RANGE) fail(SPIRA_ERR_OUT_OF_RANGE, "spira::matrix: col_index out of range");
        if (hctl[CTL_ERR] & ERR_INTERNAL)
            fail(SPIRA_ERR_DEVICE, "lock: internal error, a merged entry fell outside its row");
        uint64_t nnz_new;
        std::memcpy(&nnz_new, hctl + CTL_NNZ_LO, sizeof(nnz_new));
        // install: the old arrays go back to the cache (stream order keeps
        // them intact for anything still queued)
        free_dcsr_arrays(A);
        A->off = new_off;
        A->col = new_col;
        A->val = new_val;
        A->nnz = nnz_new;
        A->capacity = ub;
        A->built = true;
        // Dirty rows come out without zeros; only clean rows can keep the
        // explicit zeros an upload brought in.
        if (!merge) A->has_zeros = false;
        return nnz_new;
    } catch (...) {
        (void)hipStreamSynchronize(st);
        dev_release(dev, new_off, ob);
        dev_release(dev, new_col, cb);
        dev_release(dev, new_val, vb);
        throw;
    }
}

}  // namespace spira_amd
