// algebra.hip -- C = A + B, C = A - B and C = A * B on locked device CSRs,
// bit-identical to the reference's serial and parallel algorithms.
//
// Addition (serial/matrix_addition.hpp:34-112, parallel/algorithms/
// matrix_addition.hpp:24-99): every row is the column-ordered union of A's
// and B's row; a shared column holds a + b, a lone one its value; the output's
// first lock() drops every zero.  A - B adds B scaled by -1
// (matrix_operators.hpp:32-41, scalars.hpp:23-35).  Rows are classified by
// la + lb: up to 64 entries are merged in registers (a bitonic merge of A's
// row and B's reversed row on column << 1 | source, the shared column's two
// entries then neighbours), longer rows along the merge path (k_add_long).
// A first pass counts each row's kept entries from the columns
// (k_add_count_lanes; values only where a + b may cancel or a matrix may hold
// explicit zeros), a scan gives C's offsets, and the merge pass
// (k_add_lanes) writes every row straight into C.
//
// SpGEMM (serial/spgemm.hpp:12-50, parallel/algorithms/spgemm.hpp:24-69):
// Gustavson rows.  For output (i, j) the reference adds the products
// a_ik * b_kj into acc[j] = 0.0 in the order of k (A's row) and, inside B's
// row, of j; each product is rounded, then added (no fused multiply-add);
// entries with acc[j] != 0 are kept.  The device keeps that order exactly:
//   * k_spgemm_entry_len + a scan: products before each entry of A, so each
//     row has a product count and a scratch area with room for one entry per
//     product; k_spgemm_classify gives the row a size class;
//   * k_spgemm_lanes<G> (rows of <= 16, 32, 64 products): 64 / G rows per
//     wavefront, all in registers -- lane g owns product g in the reference's
//     order, a shuffle bitonic network sorts (column, g), and the first lane
//     of each column run adds the run from 0.0 in order;
//   * k_spgemm_rows<T, CAP, NA> (<= 256, 1024 products): one workgroup per row
//     through LDS; the B rows are already sorted runs, so log2(A entries)
//     merge levels (a binary search per key into the partner run) replace
//     the sort;
//   * longer rows, expand-sort-compress in HBM: products keyed by (row,
//     column) in the reference's order, a stable radix sort, one thread per
//     (row, column) run;
//   * every path writes a row's kept entries, sorted, to its scratch area and
//     its count; a scan of the counts gives C's offsets and k_spgemm_gather
//     moves the entries into C (one pass over the products in all).
// Contraction into fused multiply-adds is switched off in the kernels that
// round a product or a sum.
#include "internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <functional>

namespace spira_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;        // entries per thread in the addition passes
constexpr int kRowCache = 2048;  // row offsets a block keeps in LDS


// Rows of the entries k = first + it * kThreads + threadIdx.x (it < kItems)
// of a CSR: the block's row range from two wavefront-wide searches, then a
// search inside it -- over a copy of the range's offsets in LDS when the range
// has at most kRowCache rows.  Every thread of the block must call it.
__device__ __forceinline__ void entry_rows(const uint64_t *__restrict__ off, uint64_t n_rows, uint64_t nnz,
                                           uint64_t first, uint64_t *s_off, uint64_t (&rows)[kItems])
{
    const uint64_t last = min(first + (uint64_t)kThreads * kItems, nnz) - 1;
    const uint64_t lo = wave_row_of(off, 0, n_rows, first);
    const uint64_t hi = wave_row_of(off, lo, n_rows, last);
    const uint64_t nr = hi - lo + 1;  // the same in every wavefront
    if (nr <= kRowCache) {
        for (uint64_t i = threadIdx.x; i <= nr; i += kThreads) s_off[i] = off[lo + i];  // off[lo .. hi + 1]
        __syncthreads();
        uint32_t r = 0;  // a thread's entries ascend: each search starts at the previous row
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const uint64_t k = first + (uint64_t)it * kThreads + threadIdx.x;
            if (k < nnz) {
                uint32_t b = (uint32_t)nr;
                while (b - r > 1) {
                    const uint32_t mid = r + (b - r) / 2;
                    if (s_off[mid] <= k) r = mid;
                    else b = mid;
                }
            }
            rows[it] = lo + r;
        }
    } else {
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const uint64_t k = first + (uint64_t)it * kThreads + threadIdx.x;
            rows[it] = k < nnz ? row_of(off, lo, hi + 1, k) : lo;
        }
    }
}

__device__ __forceinline__ double scale_b(double b, int subtract)
{
#pragma clang fp contract(off)
    return subtract ? b * -1.0 : b;
}

// ---- SpGEMM ----------------------------------------------------------------

// Size classes of a row of C by its product count ub and A entries na:
//   0..2: ub, na <= 16 / 32 / 64: k_spgemm_lanes<G>, 64 / G rows per wavefront;
//   3, 4: ub, na <= 256 / 1024: k_spgemm_rows (LDS);
//   5: the rest, expand-sort-compress in HBM.
constexpr int kClasses = 6;
constexpr uint32_t kClassCap[kClasses - 1] = {16, 32, 64, 256, 1024};  // products
constexpr uint32_t kClassNa[kClasses - 1] = {16, 32, 64, 64, 256};     // entries of A

// len[k] = length of the B row that entry k of A selects (its products).
__global__ void k_spgemm_entry_len(uint64_t nnz_a, const uint32_t *__restrict__ a_col,
                                   const uint64_t *__restrict__ b_off, uint64_t *__restrict__ len)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz_a;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = a_col[k];
        len[k] = b_off[c + 1] - b_off[c];
    }
}

// Per row of A, from the prefix s of the per-entry product counts: base[i] =
// products before row i (the start of its scratch area; base[n_rows] = all),
// its size class; rows with products are appended to lists[cls * n_rows + ...]
// (order inside a list is free: rows are independent).  cnt[i] = 0 for rows
// without products.  A block takes kClassifyRows rows, counts its classes in
// LDS, reserves its ranges with one global atomic per class, then writes.
constexpr int kClassifyItems = 16;
constexpr uint64_t kClassifyRows = (uint64_t)kThreads * kClassifyItems;

// Appends the rows r0 + it * kThreads + threadIdx.x with class cls[it] >= 0 to
// lists[class * n_rows + ...]: counts in LDS (one LDS atomic per wavefront and
// class), one global atomic per class reserves the block's ranges, then the
// rows are written (order inside a list is free: rows are independent).
template <int NC>
__device__ __forceinline__ void append_to_lists(const int (&cls)[kClassifyItems], uint64_t r0, uint64_t n_rows,
                                                uint32_t *__restrict__ lists, uint32_t *__restrict__ list_n)
{
    __shared__ uint32_t s_n[NC], s_b[NC];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    if (threadIdx.x < NC) s_n[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kClassifyItems; ++it) {
        for (int q = 0; q < NC; ++q) {
            const uint64_t m = __ballot(cls[it] == q);
            if (m && lane == (uint32_t)(__ffsll((unsigned long long)m) - 1)) atomicAdd(&s_n[q], (uint32_t)__popcll(m));
        }
    }
    __syncthreads();
    if (threadIdx.x < NC) {
        const uint32_t q = threadIdx.x;
        s_b[q] = s_n[q] ? atomicAdd(&list_n[q], s_n[q]) : 0u;
        s_n[q] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kClassifyItems; ++it) {
        const uint64_t i = r0 + (uint64_t)it * kThreads + threadIdx.x;
        const int c = cls[it];
        for (int q = 0; q < NC; ++q) {
            const uint64_t m = __ballot(c == q);
            if (!m) continue;
            const uint32_t leader = __ffsll((unsigned long long)m) - 1;
            uint32_t b0 = 0;
            if (lane == leader) b0 = atomicAdd(&s_n[q], (uint32_t)__popcll(m));
            b0 = __shfl(b0, leader, kWave);
            if (c == q)
                lists[(uint64_t)q * n_rows + s_b[q] + b0 + __popcll(m & ((1ull << lane) - 1))] = (uint32_t)i;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_spgemm_classify(uint64_t n_rows, const uint64_t *__restrict__ a_off,
                                                               const uint64_t *__restrict__ s,
                                                               uint64_t *__restrict__ base,
                                                               uint64_t *__restrict__ cnt,
                                                               uint32_t *__restrict__ lists,
                                                               uint32_t *__restrict__ list_n)
{
    const uint64_t r0 = (uint64_t)blockIdx.x * kClassifyRows;
    int cls[kClassifyItems];
#pragma unroll
    for (int it = 0; it < kClassifyItems; ++it) {
        const uint64_t i = r0 + (uint64_t)it * kThreads + threadIdx.x;
        int c = -1;
        if (i < n_rows) {
            const uint64_t a0 = a_off[i], a1 = a_off[i + 1], na = a1 - a0;
            const uint64_t s0 = s[a0], u = s[a1] - s0;
            base[i] = s0;
            cnt[i] = 0;
            if (u > 0) {
                c = kClasses - 1;
                for (int q = kClasses - 2; q >= 0; --q)
                    if (na <= kClassNa[q] && u <= kClassCap[q]) c = q;
            }
        } else if (i == n_rows) {
            base[i] = s[a_off[i]];
        }
        cls[it] = c;
    }
    append_to_lists<kClasses>(cls, r0, n_rows, lists, list_n);
}

// v of lane (lane ^ m): masks 1, 2, 4, 8 stay inside a 16-lane row and go
// through DPP (quad_perm, row_shl / row_shr: VALU moves), larger ones through
// the LDS crossbar (ds_bpermute).  Every lane of the wavefront must be active.
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, uint32_t m)
{
    const int x = (int)v;
    switch (m) {
    case 1: return (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    case 2: return (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    case 4: {
        const int up = __builtin_amdgcn_update_dpp(0, x, 0x104, 0xF, 0xF, false);  // row_shl:4, lane i <- i + 4
        const int dn = __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4, lane i <- i - 4
        return (uint32_t)((threadIdx.x & 4) ? dn : up);
    }
    case 8: {
        const int up = __builtin_amdgcn_update_dpp(0, x, 0x108, 0xF, 0xF, false);
        const int dn = __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);
        return (uint32_t)((threadIdx.x & 8) ? dn : up);
    }
    default: return (uint32_t)__shfl_xor(x, (int)m, kWave);
    }
}

__device__ __forceinline__ uint64_t xor_lane64(uint64_t v, uint32_t m)
{
    return ((uint64_t)xor_lane((uint32_t)(v >> 32), m) << 32) | xor_lane((uint32_t)v, m);
}

// v of lane + 1 (wave_shl1: a DPP move across the wavefront; lane 63 reads 0).
__device__ __forceinline__ uint64_t next_lane64(uint64_t v)
{
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), 0x130, 0xF, 0xF, false);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x130, 0xF, 0xF, false);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// v of lane - 1 (wave_shr1; lane 0 reads 0).
__device__ __forceinline__ uint64_t prev_lane64(uint64_t v)
{
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), 0x138, 0xF, 0xF, false);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x138, 0xF, 0xF, false);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Rows with at most G products and G entries of A, 64 / G rows per
// wavefront, every step in registers: lane g of a row's G-lane group owns
// product g (in the reference's order), the group sorts (column, g) with a
// shuffle bitonic network, and the first lane of each column run adds the
// run's products from 0.0 in order.  The row's kept entries go to
// t_col/t_val at base[i]; cnt[i] = their number.
template <int G>
__global__ __launch_bounds__(256) void k_spgemm_lanes(const uint32_t *__restrict__ list, uint32_t n_list,
                                                       const uint64_t *__restrict__ a_off,
                                                       const uint32_t *__restrict__ a_col,
                                                       const double *__restrict__ a_val,
                                                       const uint64_t *__restrict__ b_off,
                                                       const uint32_t *__restrict__ b_col,
                                                       const double *__restrict__ b_val, const uint64_t *__restrict__ base,
                                                       uint64_t *__restrict__ cnt, uint32_t *__restrict__ t_col,
                                                       double *__restrict__ t_val)
{
#pragma clang fp contract(off)
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t g = lane & (G - 1);
    const uint32_t groups = blockDim.x / G;
    const uint64_t gmask = G == kWave ? ~0ull : (((1ull << G) - 1) << (lane & ~(uint32_t)(G - 1)));
    for (uint64_t l0 = (uint64_t)blockIdx.x * groups; l0 < n_list; l0 += (uint64_t)gridDim.x * groups) {
        const uint64_t li = l0 + threadIdx.x / G;
        const bool active = li < n_list;
        const uint32_t i = active ? list[li] : 0;
        uint32_t na = 0;
        uint64_t a0 = 0;
        if (active) {
            a0 = a_off[i];
            na = (uint32_t)(a_off[i + 1] - a0);  // <= G (classified)
        }
        uint32_t len = 0;
        uint64_t bs = 0;
        double av = 0.0;
        if (g < na) {
            const uint32_t k = a_col[a0 + g];
            bs = b_off[k];
            len = (uint32_t)(b_off[k + 1] - bs);
            av = a_val[a0 + g];
        }
        uint32_t incl = len;  // inclusive prefix of the B row lengths over the group
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
            const uint32_t u = __shfl_up(incl, d, G);
            if (g >= (uint32_t)d) incl += u;
        }
        const uint32_t pre = incl - len;
        const uint32_t ub = __shfl(incl, G - 1, G);
        // product g belongs to A entry e = the last e < na with pre[e] <= g
        uint32_t na_w = na;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) na_w = max(na_w, (uint32_t)__shfl_xor(na_w, d, kWave));
        uint32_t e = 0;
        for (uint32_t t = 1; t < na_w; ++t) {
            const uint32_t pt = __shfl(pre, t, G);
            if (t < na && pt <= g) e = t;
        }
        const uint64_t bse = __shfl(bs, e, G);
        const double ae = __shfl(av, e, G);
        const uint32_t pe = __shfl(pre, e, G);
        uint64_t key = ~0ull;
        double prod = 0.0;
        if (g < ub) {
            const uint64_t p = bse + (g - pe);
            prod = ae * b_val[p];
            key = ((uint64_t)b_col[p] << 32) | g;
        }
        // bitonic sort of (key, prod) over the group
#pragma unroll
        for (int size = 2; size <= G; size <<= 1) {
#pragma unroll
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                const uint64_t ok = xor_lane64(key, (uint32_t)stride);
                const double op = __builtin_bit_cast(double, xor_lane64(__builtin_bit_cast(uint64_t, prod), (uint32_t)stride));
                const bool up = (g & size) == 0, lower = (g & stride) == 0;
                if (lower == up ? ok < key : ok > key) {
                    key = ok;
                    prod = op;
                }
            }
        }
        // column runs: the head adds its run from 0.0 in order
        const uint32_t c = (uint32_t)(key >> 32);
        const uint32_t cprev = __shfl_up(c, 1, G);
        const bool head = g < ub && (g == 0 || cprev != c);
        double sum = 0.0;
        if (head) sum = sum + prod;
        bool open = head;
        uint64_t kd = key, pdb = __builtin_bit_cast(uint64_t, prod);  // key and product of lane + d
        for (int d = 1; d < G; ++d) {
            if (!__any(open)) break;
            kd = next_lane64(kd);
            pdb = next_lane64(pdb);
            if (open) {
                if (g + d < ub && (uint32_t)(kd >> 32) == c) sum = sum + __builtin_bit_cast(double, pdb);
                else open = false;
            }
        }
        const bool keep = head && sum != 0.0;
        const uint64_t bal = __ballot(keep) & gmask;
        if (keep) {
            const uint64_t pos = base[i] + __popcll(bal & ((1ull << lane) - 1));
            t_col[pos] = c;
            t_val[pos] = sum;
        }
        if (active && g == 0) cnt[i] = __popcll(bal);
    }
}

// Inclusive scan of v over a workgroup of T threads; total = the sum.
template <int T>
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t *s_w, uint32_t &total)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t u = __shfl_up(v, d, kWave);
        if (lane >= d) v += u;
    }
    if constexpr (T == kWave) {
        total = __shfl(v, kWave - 1, kWave);
        return v;
    } else {
        const int w = threadIdx.x / kWave;
        if (lane == kWave - 1) s_w[w] = v;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < T / kWave; ++i) {
            const uint32_t x = s_w[i];
            pre += i < w ? x : 0;
            tot += x;
        }
        __syncthreads();  // s_w free again
        total = tot;
        return v + pre;
    }
}

// Rows with at most CAP products and NA entries of A: one workgroup per row,
// products expanded into LDS, merged by (column, position) from the sorted
// runs the B rows give, one lane per column run adding from 0.0 in position
// order.  Kept entries go to t_col/t_val at base[i]; cnt[i] = their number.
template <int T, uint32_t CAP, uint32_t NA>
__global__ __launch_bounds__(T) void k_spgemm_rows(const uint32_t *__restrict__ list, uint32_t n_list,
                                                    const uint64_t *__restrict__ a_off,
                                                    const uint32_t *__restrict__ a_col,
                                                    const double *__restrict__ a_val,
                                                    const uint64_t *__restrict__ b_off,
                                                    const uint32_t *__restrict__ b_col,
                                                    const double *__restrict__ b_val, const uint64_t *__restrict__ base,
                                                    uint64_t *__restrict__ cnt, uint32_t *__restrict__ t_col,
                                                    double *__restrict__ t_val)
{
#pragma clang fp contract(off)
    // key = column << 32 | A entry << 16 | position in the reference's order
    // LDS is what limits how many rows a CU holds: the A-entry arrays are
    // needed only until the products are expanded, so the merge buffer reuses them
    constexpr uint32_t kAux = CAP > 2 * NA ? CAP : 2 * NA;
    __shared__ uint64_t s_key[CAP];
    __shared__ uint64_t s_aux[kAux];     // B row start and A value of each A entry; then the merge buffer
    __shared__ double s_prod[CAP];       // products by position
    __shared__ uint32_t s_pre[NA + 1];   // first position of each A entry's products
    __shared__ uint32_t s_w[T / kWave];
    uint64_t *const s_start = s_aux;
    double *const s_a = (double *)(s_aux + NA);
    uint64_t *const s_key2 = s_aux;
    const uint32_t tid = threadIdx.x;
    for (uint32_t li = blockIdx.x; li < n_list; li += gridDim.x) {
        const uint32_t i = list[li];
        const uint64_t a0 = a_off[i];
        const uint32_t na = (uint32_t)(a_off[i + 1] - a0);  // <= NA (classified)
        // 1. A's entries and the prefix of their B row lengths
        uint32_t carry = 0;
        for (uint32_t b0 = 0; b0 < na; b0 += T) {
            const uint32_t e = b0 + tid;
            uint32_t len = 0;
            if (e < na) {
                const uint32_t k = a_col[a0 + e];
                const uint64_t bs = b_off[k];
                len = (uint32_t)(b_off[k + 1] - bs);
                s_start[e] = bs;
                s_a[e] = a_val[a0 + e];
            }
            uint32_t tot;
            const uint32_t incl = block_incl_scan<T>(len, s_w, tot);
            if (e < na) s_pre[e] = carry + incl - len;
            carry += tot;
        }
        const uint32_t ub = carry;  // 1 .. CAP
        if (tid == 0) s_pre[na] = ub;
        __syncthreads();
        // 2. expand: position q -> A entry e (last e with s_pre[e] <= q), B entry
        {
            // every position's A entry first, then all B loads in flight at once
            constexpr int EU = CAP / T;
            uint64_t pp[EU];
            uint32_t ee[EU];
#pragma unroll
            for (int j = 0; j < EU; ++j) {
                const uint32_t q = tid + j * T;
                uint32_t lo = 0, hi = na;
                if (q < ub) {
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) / 2;
                        if (s_pre[mid] <= q) lo = mid;
                        else hi = mid;
                    }
                    pp[j] = s_start[lo] + (q - s_pre[lo]);
                } else {
                    pp[j] = 0;
                }
                ee[j] = lo;
            }
            uint32_t bc[EU];
            double bv[EU];
#pragma unroll
            for (int j = 0; j < EU; ++j) {
                if (tid + j * T < ub) {
                    bc[j] = b_col[pp[j]];
                    bv[j] = b_val[pp[j]];
                }
            }
#pragma unroll
            for (int j = 0; j < EU; ++j) {
                const uint32_t q = tid + j * T;
                if (q < ub) {
                    s_prod[q] = s_a[ee[j]] * bv[j];
                    s_key[q] = ((uint64_t)bc[j] << 32) | (ee[j] << 16) | q;
                }
            }
        }
        __syncthreads();
        // 3. merge sort: each A entry's products arrive sorted by column (a B
        // row), and in position order; runs of w lists merge pairwise, each key
        // going to (its index in its run) + (keys of the partner run below it)
        // -- keys are distinct, so the merge keeps position order for equal
        // columns
        uint64_t *kin = s_key, *kout = s_key2;
        for (uint32_t w = 1; w < na; w <<= 1) {
            for (uint32_t x = tid; x < ub; x += T) {
                const uint64_t key = kin[x];
                const uint32_t run = ((uint32_t)(key >> 16) & 0xffffu) / w;
                const uint32_t rs = s_pre[min(run * w, na)];
                const uint32_t pr = run ^ 1u;
                const uint32_t ps = s_pre[min(pr * w, na)], pe = s_pre[min((pr + 1) * w, na)];
                uint32_t lo = ps, hi = pe;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (kin[mid] < key) lo = mid + 1;
                    else hi = mid;
                }
                kout[min(rs, ps) + (x - rs) + (lo - ps)] = key;
            }
            __syncthreads();
            uint64_t *t = kin;
            kin = kout;
            kout = t;
        }
        // 4. one lane per column run: sum from 0.0 in position order
        uint32_t kept = 0;
        const uint64_t dst = base[i];
        for (uint32_t b0 = 0; b0 < ub; b0 += T) {
            const uint32_t s = b0 + tid;
            uint32_t keep = 0, c = 0;
            double sum = 0.0;
            if (s < ub) {
                c = (uint32_t)(kin[s] >> 32);
                if (s == 0 || (uint32_t)(kin[s - 1] >> 32) != c) {
                    for (uint32_t u = s; u < ub && (uint32_t)(kin[u] >> 32) == c; ++u)
                        sum = sum + s_prod[(uint32_t)kin[u] & 0xffffu];
                    keep = sum != 0.0;
                }
            }
            uint32_t tot;
            const uint32_t incl = block_incl_scan<T>(keep, s_w, tot);
            if (keep) {
                const uint64_t pos = dst + kept + incl - 1;
                t_col[pos] = c;
                t_val[pos] = sum;
            }
            kept += tot;
        }
        if (tid == 0) cnt[i] = kept;
        __syncthreads();  // LDS reused by the next row
    }
}

// ---- rows with many products: expand, sort, compress in HBM ----------------

// u[l] = product count of long row l, from the scanned counts.
__global__ void k_spgemm_big_ub(const uint32_t *__restrict__ list, uint32_t n, const uint64_t *__restrict__ base,
                                uint64_t *__restrict__ u)
{
    const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) u[l] = base[list[l] + 1] - base[list[l]];
}

__global__ __launch_bounds__(kThreads) void k_spgemm_big_expand(
    const uint32_t *__restrict__ list, uint32_t n, const uint64_t *__restrict__ u, unsigned cb,
    const uint64_t *__restrict__ a_off, const uint32_t *__restrict__ a_col, const double *__restrict__ a_val,
    const uint64_t *__restrict__ b_off, const uint32_t *__restrict__ b_col, const double *__restrict__ b_val,
    uint64_t *__restrict__ key, double *__restrict__ prod)
{
#pragma clang fp contract(off)
    for (uint32_t l = blockIdx.x; l < n; l += gridDim.x) {
        const uint32_t i = list[l];
        uint64_t pos = u[l];
        for (uint64_t ka = a_off[i]; ka < a_off[i + 1]; ++ka) {
            const uint32_t k = a_col[ka];
            const double a = a_val[ka];
            const uint64_t bs = b_off[k], len = b_off[k + 1] - bs;
            for (uint64_t t = threadIdx.x; t < len; t += kThreads) {
                key[pos + t] = ((uint64_t)l << cb) | b_col[bs + t];
                prod[pos + t] = a * b_val[bs + t];
            }
            pos += len;
        }
    }
}

// One thread per run of equal keys (a (row, column) pair): the sum from 0.0 in
// the sorted order, which the stable sort left in the reference's order.
__global__ void k_spgemm_big_runs(const uint64_t *__restrict__ skey, const double *__restrict__ sprod, uint64_t P,
                                  double *__restrict__ sum_out, uint64_t *__restrict__ keep)
{
#pragma clang fp contract(off)
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < P;
         s += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = skey[s];
        uint64_t kp = 0;
        if (s == 0 || skey[s - 1] != k) {
            double sum = 0.0;
            for (uint64_t v = s; v < P && skey[v] == k; ++v) sum = sum + sprod[v];
            sum_out[s] = sum;
            kp = sum != 0.0;
        }
        keep[s] = kp;
    }
}

__global__ void k_spgemm_big_counts(const uint32_t *__restrict__ list, uint32_t n, const uint64_t *__restrict__ u,
                                    const uint64_t *__restrict__ sk, uint64_t *__restrict__ cnt)
{
    const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) cnt[list[l]] = sk[u[l + 1]] - sk[u[l]];
}

// Kept runs of the long rows into the row's scratch area at base[row].
__global__ void k_spgemm_big_write(const uint64_t *__restrict__ skey, const double *__restrict__ sums, uint64_t P,
                                   unsigned cb, const uint32_t *__restrict__ list, const uint64_t *__restrict__ u,
                                   const uint64_t *__restrict__ sk, const uint64_t *__restrict__ base,
                                   uint32_t *__restrict__ t_col, double *__restrict__ t_val)
{
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < P;
         s += (uint64_t)gridDim.x * blockDim.x) {
        if (sk[s + 1] == sk[s]) continue;
        const uint64_t k = skey[s];
        const uint32_t l = (uint32_t)(k >> cb);
        const uint64_t pos = base[list[l]] + (sk[s] - sk[u[l]]);
        t_col[pos] = (uint32_t)(k & ((1ull << cb) - 1));
        t_val[pos] = sums[s];
    }
}

// C's entry p (row r) = scratch entry base[r] + (p - c_off[r]).
__global__ __launch_bounds__(kThreads) void k_spgemm_gather(uint64_t n_rows, uint64_t nnz,
                                                             const uint64_t *__restrict__ c_off,
                                                             const uint64_t *__restrict__ base,
                                                             const uint32_t *__restrict__ t_col,
                                                             const double *__restrict__ t_val,
                                                             uint32_t *__restrict__ c_col, double *__restrict__ c_val)
{
    __shared__ uint64_t s_off[kRowCache + 1];
    const uint64_t first = (uint64_t)blockIdx.x * kThreads * kItems;
    if (first >= nnz) return;
    uint64_t rows[kItems];
    entry_rows(c_off, n_rows, nnz, first, s_off, rows);
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t p = first + (uint64_t)it * kThreads + threadIdx.x;
        if (p >= nnz) break;
        const uint64_t r = rows[it];
        const uint64_t src = base[r] + (p - c_off[r]);
        c_col[p] = t_col[src];
        c_val[p] = t_val[src];
    }
}

// ---- addition ----------------------------------------------------------------

// Size classes of an output row by its input entries la + lb: at most 16 / 32
// / 64 (k_add_lanes<G>), or longer (k_add_long).
constexpr int kAddClasses = 4;
constexpr uint32_t kAddCap[kAddClasses - 1] = {16, 32, 64};
constexpr int kAddRows = 4;  // rows per lane group and step of k_add_lanes

// cnt[i] = 0 for empty rows; rows with entries go to their class list.
__global__ __launch_bounds__(kThreads) void k_add_classify(uint64_t n_rows, const uint64_t *__restrict__ a_off,
                                                            const uint64_t *__restrict__ b_off,
                                                            uint64_t *__restrict__ cnt, uint32_t *__restrict__ lists,
                                                            uint32_t *__restrict__ list_n)
{
    const uint64_t r0 = (uint64_t)blockIdx.x * kClassifyRows;
    int cls[kClassifyItems];
#pragma unroll
    for (int it = 0; it < kClassifyItems; ++it) {
        const uint64_t i = r0 + (uint64_t)it * kThreads + threadIdx.x;
        int c = -1;
        if (i < n_rows) {
            const uint64_t a0 = a_off[i], b0 = b_off[i];
            const uint64_t t = (a_off[i + 1] - a0) + (b_off[i + 1] - b0);
            cnt[i] = 0;
            if (t > 0) {
                c = kAddClasses - 1;
                for (int q = kAddClasses - 2; q >= 0; --q)
                    if (t <= kAddCap[q]) c = q;
            }
        }
        cls[it] = c;
    }
    append_to_lists<kAddClasses>(cls, r0, n_rows, lists, list_n);
}

// Rows with at most G entries in A's and B's row together, 64 / G rows per
// wavefront, in registers, keyed column << 1 | (0 for A, 1 for B): lanes
// 0 .. la-1 hold A's row in order, lanes G-lb .. G-1 hold B's row reversed and
// the lanes between hold the largest key, so the G keys rise and then fall
// (both rows are sorted with distinct columns) and one bitonic merge of
// log2(G) stages sorts them.  A column both rows hold then sits on two
// neighbouring lanes, A's first: that lane writes a + b (b * -1 for A - B),
// the next writes nothing; zeros are dropped.
// Each group of G lanes takes R rows per step: their list entries, offsets and
// entries are loaded together before the R networks run, so R rows' memory
// latency overlaps.
// The entries go straight to C's row (c_off from the counts of
// k_add_count_lanes, which keeps exactly the same entries); a row never writes
// past its counted length.
template <int G, int R>
__global__ __launch_bounds__(256) void k_add_lanes(const uint32_t *__restrict__ list, uint32_t n_list,
                                                    const uint64_t *__restrict__ a_off,
                                                    const uint32_t *__restrict__ a_col,
                                                    const double *__restrict__ a_val,
                                                    const uint64_t *__restrict__ b_off,
                                                    const uint32_t *__restrict__ b_col,
                                                    const double *__restrict__ b_val, int subtract, int zeros,
                                                    const uint64_t *__restrict__ c_off, uint32_t *__restrict__ c_col,
                                                    double *__restrict__ c_val)
{
#pragma clang fp contract(off)
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t g = lane & (G - 1);
    const uint32_t groups = blockDim.x / G;
    const uint64_t gmask = G == kWave ? ~0ull : (((1ull << G) - 1) << (lane & ~(uint32_t)(G - 1)));
    for (uint64_t l0 = (uint64_t)blockIdx.x * groups * R; l0 < n_list; l0 += (uint64_t)gridDim.x * groups * R) {
        const uint64_t lb0 = l0 + (uint64_t)(threadIdx.x / G) * R;  // this group's first list entry
        uint32_t row[R];
#pragma unroll
        for (int j = 0; j < R; ++j) row[j] = lb0 + j < n_list ? list[lb0 + j] : 0xffffffffu;
        uint64_t a0[R], b0[R];
        uint32_t la[R], lb[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            a0[j] = b0[j] = 0;
            la[j] = lb[j] = 0;
            if (row[j] != 0xffffffffu) {
                const uint32_t i = row[j];
                a0[j] = a_off[i];
                la[j] = (uint32_t)(a_off[i + 1] - a0[j]);
                b0[j] = b_off[i];
                lb[j] = (uint32_t)(b_off[i + 1] - b0[j]);  // la + lb <= G (classified)
            }
        }
        uint64_t key[R];
        double v[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            key[j] = ~0ull;
            v[j] = 0.0;
            if (g < la[j]) {
                key[j] = (uint64_t)a_col[a0[j] + g] << 1;
                v[j] = a_val[a0[j] + g];
            } else if (g >= G - lb[j]) {  // B's row reversed: lane G-1 holds its first entry
                const uint64_t k = b0[j] + (G - 1 - g);
                key[j] = ((uint64_t)b_col[k] << 1) | 1u;
                v[j] = scale_b(b_val[k], subtract);
            }
        }
#pragma unroll
        for (int stride = G >> 1; stride > 0; stride >>= 1) {
            const bool lower = (g & stride) == 0;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t ok = xor_lane64(key[j], (uint32_t)stride);
                const double ov =
                    __builtin_bit_cast(double, xor_lane64(__builtin_bit_cast(uint64_t, v[j]), (uint32_t)stride));
                if (lower ? ok < key[j] : ok > key[j]) {  // the lower lane keeps the smaller key
                    key[j] = ok;
                    v[j] = ov;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint32_t n = la[j] + lb[j];
            const uint64_t nk = next_lane64(key[j]), pk = prev_lane64(key[j]);
            const double nv = __builtin_bit_cast(double, next_lane64(__builtin_bit_cast(uint64_t, v[j])));
            const bool valid = g < n;
            const bool dup = (key[j] & 1u) && g > 0 && (pk >> 1) == (key[j] >> 1);  // B's entry of a shared column
            double out = v[j];
            const bool shared = !(key[j] & 1u) && g + 1 < n && (nk >> 1) == (key[j] >> 1);
            if (shared) out = v[j] + nv;
            // the counting pass's rule: a lone entry is tested for zero only
            // when a matrix may hold explicit zeros (otherwise it never is one)
            const bool keep = valid && !dup && (!(shared || zeros) || !is_zero_f64(out));
            const uint64_t bal = __ballot(keep) & gmask;
            if (keep) {
                const uint64_t c0 = c_off[row[j]], rank = __popcll(bal & ((1ull << lane) - 1));
                if (rank < c_off[row[j] + 1] - c0) {
                    c_col[c0 + rank] = (uint32_t)(key[j] >> 1);
                    c_val[c0 + rank] = out;
                }
            }
        }
    }
}

// The number of entries k_add_lanes keeps in each row, from the columns: the
// same bitonic merge on keys (column << 7 | source << 6 | index in its row),
// values read only for a column both rows hold (a + b may cancel) and, when
// either matrix may hold explicit zeros, for every entry.
template <int G, int R>
__global__ __launch_bounds__(256) void k_add_count_lanes(const uint32_t *__restrict__ list, uint32_t n_list,
                                                          const uint64_t *__restrict__ a_off,
                                                          const uint32_t *__restrict__ a_col,
                                                          const double *__restrict__ a_val,
                                                          const uint64_t *__restrict__ b_off,
                                                          const uint32_t *__restrict__ b_col,
                                                          const double *__restrict__ b_val, int subtract, int zeros,
                                                          uint64_t *__restrict__ cnt)
{
#pragma clang fp contract(off)
    static_assert(G <= 64, "6-bit entry index");
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t g = lane & (G - 1);
    const uint32_t groups = blockDim.x / G;
    const uint64_t gmask = G == kWave ? ~0ull : (((1ull << G) - 1) << (lane & ~(uint32_t)(G - 1)));
    for (uint64_t l0 = (uint64_t)blockIdx.x * groups * R; l0 < n_list; l0 += (uint64_t)gridDim.x * groups * R) {
        const uint64_t lb0 = l0 + (uint64_t)(threadIdx.x / G) * R;
        uint32_t row[R];
#pragma unroll
        for (int j = 0; j < R; ++j) row[j] = lb0 + j < n_list ? list[lb0 + j] : 0xffffffffu;
        uint64_t a0[R], b0[R];
        uint32_t la[R], lb[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            a0[j] = b0[j] = 0;
            la[j] = lb[j] = 0;
            if (row[j] != 0xffffffffu) {
                const uint32_t i = row[j];
                a0[j] = a_off[i];
                la[j] = (uint32_t)(a_off[i + 1] - a0[j]);
                b0[j] = b_off[i];
                lb[j] = (uint32_t)(b_off[i + 1] - b0[j]);
            }
        }
        uint64_t key[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            key[j] = ~0ull;
            if (g < la[j]) {
                key[j] = ((uint64_t)a_col[a0[j] + g] << 7) | g;
            } else if (g >= G - lb[j]) {
                const uint32_t k = G - 1 - g;
                key[j] = ((uint64_t)b_col[b0[j] + k] << 7) | 64u | k;
            }
        }
#pragma unroll
        for (int stride = G >> 1; stride > 0; stride >>= 1) {
            const bool lower = (g & stride) == 0;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t ok = xor_lane64(key[j], (uint32_t)stride);
                if (lower ? ok < key[j] : ok > key[j]) key[j] = ok;
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint32_t n = la[j] + lb[j];
            const uint64_t k = key[j], nk = next_lane64(k), pk = prev_lane64(k);
            const bool from_b = (k >> 6) & 1u;
            const bool shared = !from_b && g + 1 < n && (nk >> 7) == (k >> 7);  // A's entry; B holds the column too
            const bool dup = from_b && g > 0 && (pk >> 7) == (k >> 7);         // B's entry of a shared column
            bool keep = g < n && !dup;
            if (keep && (shared || zeros)) {
                double out;
                if (!from_b) {
                    out = a_val[a0[j] + (k & 63u)];
                    if (shared) out = out + scale_b(b_val[b0[j] + (nk & 63u)], subtract);
                } else {
                    out = scale_b(b_val[b0[j] + (k & 63u)], subtract);
                }
                keep = !is_zero_f64(out);
            }
            const uint64_t bal = __ballot(keep) & gmask;
            if (row[j] != 0xffffffffu && g == 0) cnt[row[j]] = __popcll(bal);
        }
    }
}

// Longer rows: one workgroup per row walks the merged order (A's entry before
// B's on a shared column) in chunks of kThreads * kAddItems positions; each
// thread finds its start on the merge path by a binary search, merges
// kAddItems positions (a shared column: a + b at A's entry, nothing at B's),
// and a workgroup scan of the kept counts places the entries.
constexpr int kAddItems = 8;

// COUNT: only the kept count of each row, into cnt; otherwise the entries go
// to C's row (c_off), never past its counted length.
template <bool COUNT>
__global__ __launch_bounds__(kThreads) void k_add_long(const uint32_t *__restrict__ list, uint32_t n_list,
                                                        const uint64_t *__restrict__ a_off,
                                                        const uint32_t *__restrict__ a_col,
                                                        const double *__restrict__ a_val,
                                                        const uint64_t *__restrict__ b_off,
                                                        const uint32_t *__restrict__ b_col,
                                                        const double *__restrict__ b_val, int subtract,
                                                        const uint64_t *__restrict__ c_off, uint64_t *__restrict__ cnt,
                                                        uint32_t *__restrict__ c_col, double *__restrict__ c_val)
{
#pragma clang fp contract(off)
    __shared__ uint32_t s_w[kThreads / kWave];
    for (uint32_t li = blockIdx.x; li < n_list; li += gridDim.x) {
        const uint32_t i = list[li];
        const uint64_t a0 = a_off[i], la = a_off[i + 1] - a0, b0 = b_off[i], lb = b_off[i + 1] - b0;
        const uint64_t n = la + lb, dst = COUNT ? 0 : c_off[i], room = COUNT ? 0 : c_off[i + 1] - dst;
        const uint32_t *ac = a_col + a0, *bc = b_col + b0;
        const double *av = a_val + a0, *bv = b_val + b0;
        uint64_t carry = 0;
        for (uint64_t c0 = 0; c0 < n; c0 += (uint64_t)kThreads * kAddItems) {
            const uint64_t d = min(n, c0 + (uint64_t)threadIdx.x * kAddItems);
            // A entries among the first d merged: A's key col << 1, B's col << 1 | 1
            uint64_t lo = d > lb ? d - lb : 0, hi = min(d, la);
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (((uint64_t)ac[mid] << 1) < (((uint64_t)bc[d - mid - 1] << 1) | 1u)) lo = mid + 1;
                else hi = mid;
            }
            uint64_t ia = lo, ib = d - lo;
            uint32_t oc[kAddItems];
            double ov[kAddItems];
            bool kp[kAddItems];
            uint32_t nk = 0;
#pragma unroll
            for (int t = 0; t < kAddItems; ++t) {
                kp[t] = false;
                oc[t] = 0;
                ov[t] = 0.0;
                if (d + t >= n) continue;
                if (ib >= lb || (ia < la && ac[ia] <= bc[ib])) {  // A's entry (first on a shared column)
                    const uint32_t c = ac[ia];
                    double v = av[ia];
                    if (ib < lb && bc[ib] == c) v = v + scale_b(bv[ib], subtract);
                    ++ia;
                    oc[t] = c;
                    ov[t] = v;
                    kp[t] = !is_zero_f64(v);
                } else {
                    const uint32_t c = bc[ib];
                    const double v = scale_b(bv[ib], subtract);
                    ++ib;
                    oc[t] = c;
                    ov[t] = v;
                    kp[t] = !(ia > 0 && ac[ia - 1] == c) && !is_zero_f64(v);
                }
                nk += kp[t] ? 1u : 0u;
            }
            uint32_t tot;
            const uint32_t incl = block_incl_scan<kThreads>(nk, s_w, tot);
            if (!COUNT) {
                uint64_t rank = carry + (incl - nk);
#pragma unroll
                for (int t = 0; t < kAddItems; ++t) {
                    if (kp[t]) {
                        if (rank < room) {
                            c_col[dst + rank] = oc[t];
                            c_val[dst + rank] = ov[t];
                        }
                        ++rank;
                    }
                }
            }
            carry += tot;
        }
        if (COUNT && threadIdx.x == 0) cnt[i] = carry;
    }
}

unsigned bits_for(uint64_t max_value)
{
    unsigned b = 1;
    while (b < 64 && (1ull << b) <= max_value) ++b;
    return b;
}

// Frees device blocks on every exit path.
struct scratch_set {
    spira_dev *dev;
    std::vector<std::pair<void *, size_t>> blocks;
    void *get(size_t bytes)
    {
        void *p = dev_alloc(dev, bytes ? bytes : 8);
        blocks.push_back({p, bytes ? bytes : 8});
        return p;
    }
    void keep(void *p)
    {
        for (auto &b : blocks)
            if (b.first == p) b.first = nullptr;
    }
    ~scratch_set()
    {
        for (auto &b : blocks)
            if (b.first) dev_release(dev, b.first, b.second);
    }
};

void install(spira_dcsr *C, uint64_t *off, uint32_t *col, double *val, uint64_t nnz)
{
    free_dcsr_arrays(C);
    C->off = off;
    C->col = col;
    C->val = val;
    C->nnz = nnz;
    C->capacity = nnz;
    C->built = true;
    C->has_zeros = false;
}

int run_guarded(spira_dev *dev, const std::function<void()> &body)
{
    try {
        SPIRA_HIP(hipSetDevice(dev->device));
        body();
        return SPIRA_OK;
    } catch (const status_error &e) {
        (void)hipStreamSynchronize(dev->stream);
        dev->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        (void)hipStreamSynchronize(dev->stream);
        dev->last_error = "host allocation failed";
        return SPIRA_ERR_NOMEM;
    } catch (const std::exception &e) {
        (void)hipStreamSynchronize(dev->stream);
        dev->last_error = e.what();
        return SPIRA_ERR_DEVICE;
    }
}

}  // namespace

}  // namespace spira_amd

using namespace spira_amd;

extern "C" int spira_add_f64(spira_dcsr *A, spira_dcsr *B, spira_dcsr *C, uint32_t flags)
{
    if (!A || !B || !C) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return run_guarded(dev, [&] {
        if (B->dev != dev || C->dev != dev) fail(SPIRA_ERR_INVALID_ARGUMENT, "add: A, B and C belong to different devices");
        if (A->n_rows != B->n_rows || A->n_cols != B->n_cols)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "Matrices aren't the same size");
        if (!A->built) fail(SPIRA_ERR_LOGIC, "MatrixAddition: A must be locked");
        if (!B->built) fail(SPIRA_ERR_LOGIC, "MatrixAddition: B must be locked");
        if (C->n_rows != A->n_rows || C->n_cols != A->n_cols)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "add: C must have the shape of A and B");
        if (flags & ~(uint32_t)SPIRA_ADD_SUBTRACT) fail(SPIRA_ERR_INVALID_ARGUMENT, "add: unknown flags");
        const uint64_t n = A->n_rows;
        if (n >= (1ull << 32) - 1) fail(SPIRA_ERR_UNSUPPORTED, "add: needs fewer than 2^32 - 1 rows");
        const int sub = (flags & SPIRA_ADD_SUBTRACT) ? 1 : 0;
        hipStream_t st = dev->stream;
        scratch_set tmp{dev, {}};
        uint64_t *cnt = (uint64_t *)tmp.get((n + 1) * sizeof(uint64_t));   // row counts, then C's offsets
        uint32_t *lists = (uint32_t *)tmp.get((uint64_t)kAddClasses * (n ? n : 1) * sizeof(uint32_t));
        uint32_t *list_n = (uint32_t *)dev->ws.get(WS_MISC, 8 * sizeof(uint32_t));
        SPIRA_HIP(hipMemsetAsync(list_n, 0, 8 * sizeof(uint32_t), st));
        k_add_classify<<<(unsigned)((n + kClassifyRows) / kClassifyRows), kThreads, 0, st>>>(n, A->off, B->off, cnt,
                                                                                         lists, list_n);
        SPIRA_HIP(hipGetLastError());
        uint32_t ln[8] = {};
        read_back(dev, ln, list_n, sizeof(ln));
        const uint32_t *l[kAddClasses];
        for (int c = 0; c < kAddClasses; ++c) l[c] = lists + (uint64_t)c * n;
        auto blocks = [](uint32_t rows, uint32_t per_block, uint32_t cap) {
            return (unsigned)std::min<uint64_t>((rows + per_block - 1) / per_block, cap);
        };
        // 1. the kept count of every row (empty rows were set to 0), 2. C's
        // offsets, 3. the merged rows written straight into C.
        const int zeros = (A->has_zeros || B->has_zeros) ? 1 : 0;
        if (ln[0])
            k_add_count_lanes<16, kAddRows><<<blocks(ln[0], 16 * kAddRows, 65536), 256, 0, st>>>(
                l[0], ln[0], A->off, A->col, A->val, B->off, B->col, B->val, sub, zeros, cnt);
        if (ln[1])
            k_add_count_lanes<32, kAddRows><<<blocks(ln[1], 8 * kAddRows, 65536), 256, 0, st>>>(
                l[1], ln[1], A->off, A->col, A->val, B->off, B->col, B->val, sub, zeros, cnt);
        if (ln[2])
            k_add_count_lanes<64, kAddRows><<<blocks(ln[2], 4 * kAddRows, 65536), 256, 0, st>>>(
                l[2], ln[2], A->off, A->col, A->val, B->off, B->col, B->val, sub, zeros, cnt);
        if (ln[3])
            k_add_long<true><<<blocks(ln[3], 1, 16384), kThreads, 0, st>>>(l[3], ln[3], A->off, A->col, A->val,
                                                                          B->off, B->col, B->val, sub, nullptr, cnt,
                                                                          nullptr, nullptr);
        SPIRA_HIP(hipGetLastError());
        scan_excl_u64(dev, cnt, cnt, n);
        const uint64_t nnz = read_scalar(dev, cnt + n);
        uint32_t *c_col = (uint32_t *)tmp.get(nnz * sizeof(uint32_t));
        double *c_val = (double *)tmp.get(nnz * sizeof(double));
        if (nnz > 0) {
            if (ln[0])
                k_add_lanes<16, kAddRows><<<blocks(ln[0], 16 * kAddRows, 65536), 256, 0, st>>>(
                    l[0], ln[0], A->off, A->col, A->val, B->off, B->col, B->val, sub, zeros, cnt, c_col, c_val);
            if (ln[1])
                k_add_lanes<32, kAddRows><<<blocks(ln[1], 8 * kAddRows, 65536), 256, 0, st>>>(
                    l[1], ln[1], A->off, A->col, A->val, B->off, B->col, B->val, sub, zeros, cnt, c_col, c_val);
            if (ln[2])
                k_add_lanes<64, kAddRows><<<blocks(ln[2], 4 * kAddRows, 65536), 256, 0, st>>>(
                    l[2], ln[2], A->off, A->col, A->val, B->off, B->col, B->val, sub, zeros, cnt, c_col, c_val);
            if (ln[3])
                k_add_long<false><<<blocks(ln[3], 1, 16384), kThreads, 0, st>>>(l[3], ln[3], A->off, A->col,
                                                                               A->val, B->off, B->col, B->val, sub,
                                                                               cnt, nullptr, c_col, c_val);
            SPIRA_HIP(hipGetLastError());
        }
        tmp.keep(cnt);
        tmp.keep(c_col);
        tmp.keep(c_val);
        install(C, cnt, c_col, c_val, nnz);  // C may be A or B: stream order protects their arrays
    });
}

extern "C" int spira_spgemm_f64(spira_dcsr *A, spira_dcsr *B, spira_dcsr *C)
{
    if (!A || !B || !C) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    return run_guarded(dev, [&] {
        if (B->dev != dev || C->dev != dev)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "spgemm: A, B and C belong to different devices");
        if (A->n_cols != B->n_rows) fail(SPIRA_ERR_INVALID_ARGUMENT, "A.cols must equal B.rows");
        if (!A->built) fail(SPIRA_ERR_LOGIC, "spgemm: A must be locked");
        if (!B->built) fail(SPIRA_ERR_LOGIC, "spgemm: B must be locked");
        if (C->n_rows != A->n_rows || C->n_cols != B->n_cols)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "spgemm: C must have A's rows and B's columns");
        const uint64_t n = A->n_rows, m = B->n_cols;
        if (n >= (1ull << 32) - 1) fail(SPIRA_ERR_UNSUPPORTED, "spgemm: needs fewer than 2^32 - 1 rows");
        hipStream_t st = dev->stream;
        scratch_set tmp{dev, {}};
        uint64_t *cnt = (uint64_t *)tmp.get((n + 1) * sizeof(uint64_t));   // row counts, then C's offsets
        uint64_t *base = (uint64_t *)tmp.get((n + 1) * sizeof(uint64_t));  // scratch area of each row
        uint32_t *lists = (uint32_t *)tmp.get((uint64_t)kClasses * (n ? n : 1) * sizeof(uint32_t));
        uint32_t *list_n = (uint32_t *)dev->ws.get(WS_MISC, 8 * sizeof(uint32_t));
        SPIRA_HIP(hipMemsetAsync(list_n, 0, 8 * sizeof(uint32_t), st));
        const uint64_t nnz_a = A->nnz;
        uint64_t *len = (uint64_t *)tmp.get((nnz_a + 1) * sizeof(uint64_t));
        if (nnz_a) {
            k_spgemm_entry_len<<<grid_for(nnz_a, kThreads, 16384), kThreads, 0, st>>>(nnz_a, A->col, B->off, len);
            SPIRA_HIP(hipGetLastError());
        }
        scan_excl_u64(dev, len, len, nnz_a);  // products before each entry of A
        k_spgemm_classify<<<(unsigned)((n + kClassifyRows) / kClassifyRows), kThreads, 0, st>>>(n, A->off, len, base,
                                                                                             cnt, lists, list_n);
        SPIRA_HIP(hipGetLastError());
        uint32_t ln[8] = {};
        SPIRA_HIP(hipMemcpyAsync(ln, list_n, sizeof(ln), hipMemcpyDeviceToHost, st));
        const uint64_t total = read_scalar(dev, base + n);  // synchronises ln too
        // scratch for the rows' kept entries (at most one per product)
        if (total > (uint64_t(48) << 30) / 12)
            fail(SPIRA_ERR_UNSUPPORTED, "spgemm: more than 4G products (scratch of 12 bytes per product)");
        uint32_t *t_col = (uint32_t *)tmp.get(total * sizeof(uint32_t));
        double *t_val = (double *)tmp.get(total * sizeof(double));
        const uint32_t *l[kClasses];
        for (int c = 0; c < kClasses; ++c) l[c] = lists + (uint64_t)c * n;
        auto blocks = [](uint32_t rows, uint32_t per_block, uint32_t cap) {
            return (unsigned)std::min<uint64_t>((rows + per_block - 1) / per_block, cap);
        };
        if (ln[0])
            k_spgemm_lanes<16><<<blocks(ln[0], 16, 65536), 256, 0, st>>>(l[0], ln[0], A->off, A->col, A->val, B->off,
                                                                        B->col, B->val, base, cnt, t_col, t_val);
        if (ln[1])
            k_spgemm_lanes<32><<<blocks(ln[1], 8, 65536), 256, 0, st>>>(l[1], ln[1], A->off, A->col, A->val, B->off,
                                                                       B->col, B->val, base, cnt, t_col, t_val);
        if (ln[2])
            k_spgemm_lanes<64><<<blocks(ln[2], 4, 65536), 256, 0, st>>>(l[2], ln[2], A->off, A->col, A->val, B->off,
                                                                       B->col, B->val, base, cnt, t_col, t_val);
        if (ln[3])
            k_spgemm_rows<64, kClassCap[3], kClassNa[3]><<<blocks(ln[3], 1, 65536), 64, 0, st>>>(
                l[3], ln[3], A->off, A->col, A->val, B->off, B->col, B->val, base, cnt, t_col, t_val);
        if (ln[4])
            k_spgemm_rows<256, kClassCap[4], kClassNa[4]><<<blocks(ln[4], 1, 32768), 256, 0, st>>>(
                l[4], ln[4], A->off, A->col, A->val, B->off, B->col, B->val, base, cnt, t_col, t_val);
        SPIRA_HIP(hipGetLastError());
        const uint32_t n5 = ln[5];
        if (n5) {  // rows with many products
            const uint32_t *l5 = l[5];
            uint64_t *u = (uint64_t *)tmp.get(((uint64_t)n5 + 1) * sizeof(uint64_t));
            k_spgemm_big_ub<<<grid_for(n5, kThreads), kThreads, 0, st>>>(l5, n5, base, u);
            SPIRA_HIP(hipGetLastError());
            scan_excl_u64(dev, u, u, n5);
            const uint64_t P = read_scalar(dev, u + n5);
            if (P >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "spgemm: more than 2^32 products in long rows");
            const unsigned cb = bits_for(m - 1), rb = bits_for(n5 - 1);
            uint64_t *key = (uint64_t *)tmp.get(P * sizeof(uint64_t));
            double *prod = (double *)tmp.get(P * sizeof(double));
            uint64_t *skey = (uint64_t *)tmp.get(P * sizeof(uint64_t));
            double *sprod = (double *)tmp.get(P * sizeof(double));
            uint64_t *keep = (uint64_t *)tmp.get((P + 1) * sizeof(uint64_t));
            k_spgemm_big_expand<<<std::min<uint32_t>(n5, 8192), kThreads, 0, st>>>(
                l5, n5, u, cb, A->off, A->col, A->val, B->off, B->col, B->val, key, prod);
            SPIRA_HIP(hipGetLastError());
            size_t tb = 0;
            SPIRA_HIP(rocprim::radix_sort_pairs(nullptr, tb, key, skey, prod, sprod, (unsigned)P, 0, cb + rb, st));
            void *t = dev->ws.get(WS_SORT, tb ? tb : 8);
            SPIRA_HIP(rocprim::radix_sort_pairs(t, tb, key, skey, prod, sprod, (unsigned)P, 0, cb + rb, st));
            double *sums = prod;  // the unsorted products are no longer needed
            k_spgemm_big_runs<<<grid_for(P, kThreads, 16384), kThreads, 0, st>>>(skey, sprod, P, sums, keep);
            SPIRA_HIP(hipGetLastError());
            scan_excl_u64(dev, keep, keep, P);
            k_spgemm_big_counts<<<grid_for(n5, kThreads), kThreads, 0, st>>>(l5, n5, u, keep, cnt);
            SPIRA_HIP(hipGetLastError());
            k_spgemm_big_write<<<grid_for(P, kThreads, 16384), kThreads, 0, st>>>(skey, sums, P, cb, l5, u, keep,
                                                                                base, t_col, t_val);
            SPIRA_HIP(hipGetLastError());
        }
        scan_excl_u64(dev, cnt, cnt, n);
        const uint64_t nnz = read_scalar(dev, cnt + n);
        uint32_t *c_col = (uint32_t *)tmp.get(nnz * sizeof(uint32_t));
        double *c_val = (double *)tmp.get(nnz * sizeof(double));
        if (nnz > 0) {
            const uint64_t nb = (nnz + (uint64_t)kThreads * kItems - 1) / ((uint64_t)kThreads * kItems);
            if (nb >= (1ull << 31)) fail(SPIRA_ERR_UNSUPPORTED, "spgemm: too many output entries");
            k_spgemm_gather<<<(unsigned)nb, kThreads, 0, st>>>(n, nnz, cnt, base, t_col, t_val, c_col, c_val);
            SPIRA_HIP(hipGetLastError());
        }
        tmp.keep(cnt);
        tmp.keep(c_col);
        tmp.keep(c_val);
        install(C, cnt, c_col, c_val, nnz);
    });
}
