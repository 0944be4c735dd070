// batch.hip -- many matrices of uneven size, k right-hand sides (k = 1, 4, 16
// or 32), one persistent launch fed by a device work queue; the repeated call
// can be captured in a hipGraph.
//
// Layout.  A batch is one block-diagonal CSR whose rows are stored, within
// each matrix, sorted by length class (class c holds rows of 2^(c-1) ..
// 2^c - 1 nonzeros; class 0 the empty rows), each class in ascending row
// order; rowid[p] maps a stored row back to its row of the block-diagonal
// system, so X (sum of n_cols x k, row major) and Y (sum of n_rows x k) keep
// the caller's order.  Columns are stored as u16 offsets inside their matrix
// when every matrix has at most 65536 columns (u32 batch columns otherwise);
// values as FP32, BF16 or split BF16.
//
// Work.  A schedule gives every matrix a lanes-per-row rule (2^plog2 target
// nonzeros per lane: lanes per row = 2^class / 2^plog2, at most 64 / k), a
// work-item size and whether its X block is staged in LDS.  A tile is a run
// of stored rows of one class (one lanes-per-row value: rows per wavefront =
// 64 / (lanes * k)); rows long enough to give a wavefront's lane slot more
// than 32 nonzeros (from 2048 / k nonzeros) are tiles of their own, worked on
// by the whole block.  A work item is a run of consecutive stored
// rows of one matrix (its tiles) of about item_nnz nonzeros + rows and is the
// unit of the device queue: items are queued heaviest first and every block
// of the persistent grid (one 1024-thread block per CU) takes the next one
// with one atomic add until the queue is empty.
//
// LDS.  A block first copies what its item reads into LDS with 16-byte
// loads, all in flight at once: the matrix's X block when the schedule
// stages it (n_cols * k floats, up to SPIRA_BATCH_XLDS_FLOATS), and the
// item's row pointers, row ids, columns and values when they fit beside it.
// The rows are then computed from LDS, so the only global reads left in the
// compute phase are the X gathers of unstaged X blocks.  Items that do not
// fit (long rows, or an X block that fills the LDS) read their entries from
// global memory with the next rows' bounds loaded a step ahead.
//
// Each output is accumulated with the compensated Dot2 scheme (Ogita, Rump,
// Oishi 2005) in FP32; the lanes of a row merge their (sum, correction) pairs
// in a fixed order, so a call is deterministic.
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace spira_amd {

namespace {

constexpr int kThreads = 1024;  // one block per CU
constexpr int kWaves = kThreads / 64;
constexpr uint32_t kLdsBytes = 155584;  // LDS for the X block and staged entries
static_assert(SPIRA_BATCH_XLDS_FLOATS * 4 + 32 <= kLdsBytes, "X block must fit the LDS area");
constexpr uint32_t kMinEntryLds = 65536;  // stage entries only with at least this much LDS left
constexpr uint32_t kClasses = 34;         // bit length of a u32 row length, plus one
constexpr uint32_t kTileLong = 0x80u;
constexpr uint32_t kItemStaged = 1u;   // item flag: entries staged in LDS
constexpr uint32_t kItemTwoPass = 2u;  // item flag: k = 1, X block staged in two column halves
constexpr size_t kArrayPad = 64;      // bytes after each entry array (16-byte staging reads)
// How a block stages its item, as 16-byte loads in flight per lane: 0 =
// straight into LDS (global_load_lds_dwordx4: no VGPRs, every chunk of the
// item in flight at once) at k = 1, where staging is most of the item's time;
// one register-staged load at k >= 4, where a deeper staging loop changes how
// the compiler schedules the compute phase and costs more than it saves
// (measured).
template <int K>
constexpr uint32_t stage_depth() { return K == 1 ? 0u : 1u; }

// Before the barrier that publishes staged data: direct-to-LDS loads complete
// only when the issuing wavefront has waited for them.
template <uint32_t D>
__device__ __forceinline__ void stage_wait()
{
    if constexpr (D == 0) __builtin_amdgcn_s_waitcnt(0);
}

inline __host__ __device__ uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }

// LDS bytes of an X block of n floats: it is copied as the 16-byte aligned
// span around its elements (see stage_array), so it may start up to 12 bytes
// into its area and needs up to 32 bytes more than n * 4.
inline __host__ __device__ uint32_t xblock_bytes(uint32_t n) { return n ? align16(n * 4) + 32 : 0u; }

// Two-pass items: each half of the X block (at most kHalfFloats floats) sits
// at the start of the LDS area, followed by one (sum, correction, cursor)
// record per row of the item.
constexpr uint32_t kHalfFloats = 32768;
constexpr uint32_t kHalfBytes = kHalfFloats * 4 + 32;
constexpr uint32_t kTwoPassRows = (kLdsBytes - kHalfBytes) / 12;

__device__ __forceinline__ float bf16f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

struct tile_t {
    uint32_t r0, r1;  // stored rows [r0, r1)
    uint32_t meta;    // bits 0-6: log2 lanes per row per right-hand side; kTileLong: one whole-block row
    uint32_t pad;
};

struct item_t {
    uint32_t t0, t1;  // tiles [t0, t1)
    uint32_t cbase;   // first batch column of the item's matrix
    uint32_t xcols;   // columns whose X rows are staged in LDS (0: gathers read global memory)
    uint32_t r0, r1;  // stored rows of the item
    uint32_t ebase;   // first entry of row r0
    uint32_t flags;   // kItemStaged, kItemTwoPass
};

// Byte layout of an item's staged entries in LDS (after the X block).  An
// entry array is copied as the 16-byte aligned span around its elements,
// at most align16(n * size) + 32 bytes.
template <int ST, typename CT>
struct lds_layout {
    static constexpr uint32_t n_val_arrays = ST == SPIRA_VAL_SPLIT_BF16 ? 2 : 1;
    static constexpr uint32_t val_size = ST == SPIRA_VAL_F32 ? 4 : 2;
    static __host__ __device__ uint32_t span(uint32_t n, uint32_t size) { return align16(n * size) + 32; }
    uint32_t off_o, rid_o, col_o, v1_o, v2_o, end_o;
    __host__ __device__ lds_layout(uint32_t xbytes, uint32_t rows, uint32_t nnz)
    {
        off_o = align16(xbytes);
        rid_o = off_o + 4 * (rows + 1);
        col_o = align16(rid_o + 4 * rows);
        v1_o = col_o + span(nnz, (uint32_t)sizeof(CT));
        v2_o = v1_o + span(nnz, val_size);
        end_o = n_val_arrays == 2 ? v2_o + span(nnz, val_size) : v2_o;
    }
};

struct acc2 {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float a, float x)
    {
        const float p = a * x;
        const float e = fmaf(a, x, -p);
        const float t = s + p;
        const float bp = t - s;
        c += ((s - (t - bp)) + (p - bp)) + e;
        s = t;
    }
    __device__ __forceinline__ void merge(float s2, float c2)
    {
        const float t = s + s2;
        const float bp = t - s;
        c += c2 + ((s - (t - bp)) + (s2 - bp));
        s = t;
    }
};

__global__ void k_reset_counter(uint32_t *counter) { *counter = 0u; }

// Source of entries (column, value): global arrays or the item's LDS copy,
// one storage format each (template parameters, so the unrolled loads stay
// straight-line code).  get() is masked: a lane with ok false reads entry
// min(q, last) -- inside the arrays -- and gets column 0, value 0.
template <int ST, typename CT, bool LDS>
struct esrc {
    const CT *col;
    const float *f32;
    const uint16_t *hi, *lo;
    uint32_t last;
    __device__ __forceinline__ void get(bool ok, uint32_t q, uint32_t &c, float &v) const
    {
        const uint32_t qc = q < last ? q : last;
        uint32_t cc;
        float vv;
        if constexpr (LDS) {
            cc = (uint32_t)col[qc];
            if constexpr (ST == SPIRA_VAL_F32) vv = f32[qc];
            else if constexpr (ST == SPIRA_VAL_BF16) vv = bf16f(hi[qc]);
            else vv = bf16f(hi[qc]) + bf16f(lo[qc]);
        } else {
            cc = (uint32_t)__builtin_nontemporal_load(col + qc);
            if constexpr (ST == SPIRA_VAL_F32) vv = __builtin_nontemporal_load(f32 + qc);
            else if constexpr (ST == SPIRA_VAL_BF16) vv = bf16f(__builtin_nontemporal_load(hi + qc));
            else vv = bf16f(__builtin_nontemporal_load(hi + qc)) + bf16f(__builtin_nontemporal_load(lo + qc));
        }
        c = ok ? cc : 0u;
        v = ok ? vv : 0.0f;
    }
};

// Masked X element of stored column c for right-hand side j, from the LDS
// block (XS) or global X.  A masked lane reads element j of the block
// (present whenever the batch has a nonzero) and returns 0.
template <int K, typename CT, bool XS>
struct xsrc {
    const float *X;  // global X, or the LDS block
    uint32_t cbase;
    __device__ __forceinline__ float operator()(bool ok, uint32_t c, int j) const
    {
        const uint32_t lc = (sizeof(CT) == 2) ? c : c - cbase;  // column inside the matrix
        float v;
        if constexpr (XS) {
            v = X[ok ? lc * K + j : (uint32_t)j];
        } else {
            const uint64_t gi = ((uint64_t)cbase + lc) * K + j;
            v = X[ok ? gi : (uint64_t)j];
        }
        return ok ? v : 0.0f;
    }
};

// One row's share of the lanes (entries beg + e, beg + e + E, ...), U
// independent loads in flight per step.
template <int K, int U, typename ES, typename XV>
__device__ __forceinline__ void row_dot(acc2 &a, uint32_t beg, uint32_t end, uint32_t e, uint32_t E, const ES &es,
                                        const XV &xv, int j)
{
    for (uint32_t i = beg + e; i < end; i += U * E) {
        uint32_t c[U];
        float av[U], xx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) es.get(i + u * E < end, i + u * E, c[u], av[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) xx[u] = xv(i + u * E < end, c[u], j);
#pragma unroll
        for (int u = 0; u < U; ++u) a.add(av[u], xx[u]);
    }
}

// Two rows of one lane group at once: per step U column/value loads of each
// row in flight, then U gathers of each.
template <int K, int U, typename ES, typename XV>
__device__ __forceinline__ void rows2_dot(acc2 &a0, acc2 &a1, uint32_t b0, uint32_t e0, uint32_t b1, uint32_t e1,
                                          uint32_t e, uint32_t E, const ES &es, const XV &xv, int j)
{
    uint32_t i0 = b0 + e, i1 = b1 + e;
    while (i0 < e0 || i1 < e1) {
        uint32_t c0[U], c1[U];
        float v0[U], v1[U], x0[U], x1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            es.get(i0 + u * E < e0, i0 + u * E, c0[u], v0[u]);
            es.get(i1 + u * E < e1, i1 + u * E, c1[u], v1[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            x0[u] = xv(i0 + u * E < e0, c0[u], j);
            x1[u] = xv(i1 + u * E < e1, c1[u], j);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a0.add(v0[u], x0[u]);
            a1.add(v1[u], x1[u]);
        }
        i0 += U * E;
        i1 += U * E;
    }
}

template <int K>
__device__ __forceinline__ void merge_lanes(acc2 &a, uint32_t E)
{
    for (uint32_t o = K; o < K * E; o <<= 1) a.merge(__shfl_xor(a.s, (int)o, 64), __shfl_xor(a.c, (int)o, 64));
}

// Copy nbytes (a multiple of 16) from 16-byte aligned global src to LDS dst
// with D 16-byte loads in flight per lane, or (D = 0) straight into LDS: one
// wavefront instruction moves 64 consecutive chunks, lane l's to dst + 16 l
// (lanes past the end masked off); the caller waits with stage_wait<0>().
template <uint32_t D>
__device__ __forceinline__ void stage16(unsigned char *dst, const unsigned char *src, uint32_t nbytes)
{
    const u32x4 *s = (const u32x4 *)src;
    u32x4 *d = (u32x4 *)dst;
    const uint32_t n = nbytes / 16;
    if constexpr (D == 0) {
        const uint32_t lane = threadIdx.x & (kWave - 1);
        for (uint32_t c0 = threadIdx.x & ~(uint32_t)(kWave - 1); c0 < n; c0 += kThreads)
            if (c0 + lane < n)
                __builtin_amdgcn_global_load_lds((const void *)(s + c0 + lane),
                                                 (__attribute__((address_space(3))) void *)(d + c0), 16, 0,
                                                 2 /* non-temporal */);
        return;
    }
    uint32_t i = threadIdx.x;
    if constexpr (D > 1) {
        for (; i + (D - 1) * kThreads < n; i += D * kThreads) {
            u32x4 v[D];
#pragma unroll
            for (uint32_t u = 0; u < D; ++u) v[u] = __builtin_nontemporal_load(s + i + u * kThreads);
#pragma unroll
            for (uint32_t u = 0; u < D; ++u) d[i + u * kThreads] = v[u];
        }
    }
    for (; i < n; i += kThreads) d[i] = __builtin_nontemporal_load(s + i);
}

// Stage one array: the 16-byte aligned span holding elements [base, base + n)
// of arr, D loads in flight per lane; returns the LDS pointer to element base.
template <uint32_t D, typename T>
__device__ __forceinline__ const T *stage_array(unsigned char *lds, uint32_t o, const T *arr, uint64_t base,
                                                uint32_t n)
{
    const uintptr_t a = (uintptr_t)(arr + base);
    const uintptr_t a0 = a & ~(uintptr_t)15;
    const uintptr_t a1 = ((uintptr_t)(arr + base + n) + 15) & ~(uintptr_t)15;
    stage16<D>(lds + o, (const unsigned char *)a0, (uint32_t)(a1 - a0));
    return (const T *)(lds + o + (uint32_t)(a - a0));
}

// Where stage_array put element base of arr in the LDS area at offset o.
template <typename T>
__device__ __forceinline__ const T *staged_at(const unsigned char *lds, uint32_t o, const T *arr, uint64_t base)
{
    return (const T *)(lds + o + (uint32_t)(((uintptr_t)(arr + base)) & 15u));
}

struct kparams {
    const item_t *items;
    uint32_t n_items;
    uint32_t *counter;
    const tile_t *tiles;
    const uint32_t *off;
    const uint32_t *rowid;
    const void *col;
    const float *f32;
    const uint16_t *hi, *lo;
    uint32_t last;  // index of the batch's last stored entry (0 when it has none)
    const float *X;
    float *Y;
    uint64_t *prof;
};

// All tiles of one item; XS: X block in LDS, ESTAGED: entries in LDS.
template <int K, int ST, typename CT, bool XS, bool ESTAGED>
__device__ __forceinline__ void run_item(const kparams &kp, const item_t &it, unsigned char *lds,
                                         float (*s_part)[kWaves][64], int lane, int w, int j)
{
    constexpr int U = K >= 16 ? 8 : 4;  // loads in flight per row and lane group
    const xsrc<K, CT, XS> xv{XS ? staged_at(lds, 0, kp.X, (uint64_t)it.cbase * K) : kp.X, it.cbase};
    const uint32_t *s_off = nullptr, *s_rid = nullptr;
    esrc<ST, CT, ESTAGED> es{};
    if constexpr (ESTAGED) {
        const uint32_t rows = it.r1 - it.r0;
        const uint32_t nnz = kp.off[it.r1] - it.ebase;
        const lds_layout<ST, CT> L(XS ? xblock_bytes(it.xcols * K) : 0u, rows, nnz);
        s_off = (const uint32_t *)(lds + L.off_o);
        s_rid = (const uint32_t *)(lds + L.rid_o);
        // the entry arrays start at their 16-byte aligned span (see stage_array)
        const uint32_t cs = (uint32_t)(((uintptr_t)((const CT *)kp.col + it.ebase)) & 15u);
        es.col = (const CT *)(lds + L.col_o + cs);
        if constexpr (ST == SPIRA_VAL_F32) {
            es.f32 = (const float *)(lds + L.v1_o + (uint32_t)(((uintptr_t)(kp.f32 + it.ebase)) & 15u));
        } else {
            es.hi = (const uint16_t *)(lds + L.v1_o + (uint32_t)(((uintptr_t)(kp.hi + it.ebase)) & 15u));
            if constexpr (ST == SPIRA_VAL_SPLIT_BF16)
                es.lo = (const uint16_t *)(lds + L.v2_o + (uint32_t)(((uintptr_t)(kp.lo + it.ebase)) & 15u));
        }
        es.last = nnz ? nnz - 1 : 0u;
    } else {
        es.col = (const CT *)kp.col;
        es.f32 = kp.f32;
        es.hi = kp.hi;
        es.lo = kp.lo;
        es.last = kp.last;
    }
    for (uint32_t tt = it.t0; tt < it.t1; ++tt) {
        const tile_t tl = kp.tiles[tt];
        if (tl.meta & kTileLong) {
            // one long row: every wave strides over it, then the waves merge
            const uint32_t p = tl.r0;
            uint32_t beg, end, id;
            if constexpr (ESTAGED) {
                beg = s_off[p - it.r0];
                end = s_off[p - it.r0 + 1];
                id = s_rid[p - it.r0];
            } else {
                beg = kp.off[p];
                end = kp.off[p + 1];
                id = kp.rowid[p];
            }
            constexpr int slots = 64 / K;
            const uint32_t e = (uint32_t)(lane / K) + (uint32_t)w * slots;
            acc2 a;
            row_dot<K, U>(a, beg, end, e, (uint32_t)(kWaves * slots), es, xv, j);
            for (int o = K; o < 64; o <<= 1) a.merge(__shfl_xor(a.s, o, 64), __shfl_xor(a.c, o, 64));
            if (lane < K) {
                s_part[0][w][j] = a.s;
                s_part[1][w][j] = a.c;
            }
            __syncthreads();
            if (w == 0 && lane < K) {
                acc2 b;
                for (int ww = 0; ww < kWaves; ++ww) b.merge(s_part[0][ww][j], s_part[1][ww][j]);
                kp.Y[(uint64_t)id * K + j] = b.s + b.c;
            }
            __syncthreads();
            continue;
        }
        // lanes = (row slot rs, entry slot e, rhs j): E entry lanes per row,
        // 64 / (E * K) rows per wavefront step; the group takes rows p and
        // p + S per step (S = rows per block sweep).  The lanes of a row share
        // its row and so its activity: the merge pairs only them.
        const uint32_t E = 1u << (tl.meta & 0x7fu);
        const uint32_t e = (uint32_t)(lane / K) & (E - 1);
        const uint32_t rs = (uint32_t)lane / (K * E);
        const uint32_t rpw = 64u / (K * E);
        const uint32_t S = kWaves * rpw;
        const uint32_t r1 = tl.r1;
        uint32_t p = tl.r0 + (uint32_t)w * rpw + rs;
        if constexpr (ESTAGED) {
            for (; p < r1; p += 2 * S) {
                const uint32_t l0 = p - it.r0, l1 = l0 + S;
                const bool has1 = p + S < r1;
                const uint32_t b0 = s_off[l0], e0 = s_off[l0 + 1];
                const uint32_t b1 = has1 ? s_off[l1] : 0u, e1 = has1 ? s_off[l1 + 1] : 0u;
                acc2 a0, a1;
                rows2_dot<K, U>(a0, a1, b0, e0, b1, e1, e, E, es, xv, j);
                merge_lanes<K>(a0, E);
                merge_lanes<K>(a1, E);
                if (e == 0) {
                    kp.Y[(uint64_t)s_rid[l0] * K + j] = a0.s + a0.c;
                    if (has1) kp.Y[(uint64_t)s_rid[l1] * K + j] = a1.s + a1.c;
                }
            }
        } else {
            // bounds and row ids of the next pair are loaded a step ahead
            uint32_t b0 = 0, e0 = 0, b1 = 0, e1 = 0, id0 = 0, id1 = 0;
            if (p < r1) {
                b0 = kp.off[p];
                e0 = kp.off[p + 1];
                id0 = kp.rowid[p];
            }
            if (p + S < r1) {
                b1 = kp.off[p + S];
                e1 = kp.off[p + S + 1];
                id1 = kp.rowid[p + S];
            }
            while (p < r1) {
                const uint32_t pn = p + 2 * S;
                uint32_t nb0 = 0, ne0 = 0, nb1 = 0, ne1 = 0, nid0 = 0, nid1 = 0;
                if (pn < r1) {
                    nb0 = kp.off[pn];
                    ne0 = kp.off[pn + 1];
                    nid0 = kp.rowid[pn];
                }
                if (pn + S < r1) {
                    nb1 = kp.off[pn + S];
                    ne1 = kp.off[pn + S + 1];
                    nid1 = kp.rowid[pn + S];
                }
                acc2 a0, a1;
                rows2_dot<K, U>(a0, a1, b0, e0, b1, e1, e, E, es, xv, j);
                merge_lanes<K>(a0, E);
                merge_lanes<K>(a1, E);
                if (e == 0) {
                    kp.Y[(uint64_t)id0 * K + j] = a0.s + a0.c;
                    if (p + S < r1) kp.Y[(uint64_t)id1 * K + j] = a1.s + a1.c;
                }
                p = pn;
                b0 = nb0;
                e0 = ne0;
                b1 = nb1;
                e1 = ne1;
                id0 = nid0;
                id1 = nid1;
            }
        }
    }
}

// k = 1 item whose X block is staged in two column halves.  Pass h stages
// columns [h W, min(n, (h + 1) W)) of the matrix (W = ceil(n / 2)); each row's
// lanes consume its entries from the row's cursor while their column is in the
// half (columns are sorted, so a row's entries of half 0 precede those of half
// 1), merge into the row's (sum, correction) record and advance the cursor.
// Entries are read from global memory; the rows carry no whole-block tiles.
template <int ST, typename CT>
__device__ __forceinline__ void run_item_two_pass(const kparams &kp, const item_t &it, unsigned char *lds, int lane,
                                                  int w)
{
    constexpr int U = 4;
    const uint32_t n = it.xcols, W = (n + 1) / 2;
    const uint32_t rows = it.r1 - it.r0;
    float *s_sum = (float *)(lds + kHalfBytes);
    float *s_cor = s_sum + kTwoPassRows;
    uint32_t *s_pos = (uint32_t *)(s_cor + kTwoPassRows);
    const esrc<ST, CT, false> es{(const CT *)kp.col, kp.f32, kp.hi, kp.lo, kp.last};
    for (uint32_t q = threadIdx.x; q < rows; q += kThreads) {
        s_sum[q] = 0.0f;
        s_cor[q] = 0.0f;
        s_pos[q] = kp.off[it.r0 + q];
    }
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t c0 = h * W, c1 = min(n, c0 + W);
        __syncthreads();  // the previous pass is done with the half and the records
        const float *xs = stage_array<stage_depth<1>()>(lds, 0, kp.X, (uint64_t)it.cbase + c0, c1 - c0);
        stage_wait<stage_depth<1>()>();
        __syncthreads();
        for (uint32_t tt = it.t0; tt < it.t1; ++tt) {
            const tile_t tl = kp.tiles[tt];
            const uint32_t E = 1u << (tl.meta & 0x7fu);
            const uint32_t e = (uint32_t)lane & (E - 1);
            const uint32_t rs = (uint32_t)lane / E;
            const uint32_t rpw = 64u / E;
            for (uint32_t p = tl.r0 + (uint32_t)w * rpw + rs; p < tl.r1; p += kWaves * rpw) {
                const uint32_t lq = p - it.r0;
                const uint32_t beg = s_pos[lq], end = kp.off[p + 1];
                acc2 a;
                uint32_t upto = beg;  // one past the last entry this lane consumed
                for (uint32_t i = beg + e; i < end; i += U * E) {
                    uint32_t c[U];
                    float v[U], xx[U];
                    bool ok[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        es.get(i + u * E < end, i + u * E, c[u], v[u]);
                        const uint32_t lc = (sizeof(CT) == 2) ? c[u] : c[u] - it.cbase;
                        ok[u] = i + u * E < end && lc < c1;
                        xx[u] = xs[ok[u] ? lc - c0 : 0u];
                        if (!ok[u]) {
                            v[u] = 0.0f;
                            xx[u] = 0.0f;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        a.add(v[u], xx[u]);
                        if (ok[u]) upto = i + u * E + 1;
                    }
                    if (!ok[U - 1]) break;  // sorted columns: the rest of the row is in the next half
                }
                for (uint32_t o = 1; o < E; o <<= 1) {
                    a.merge(__shfl_xor(a.s, (int)o, 64), __shfl_xor(a.c, (int)o, 64));
                    upto = max(upto, (uint32_t)__shfl_xor((int)upto, (int)o, 64));
                }
                if (e == 0) {
                    acc2 r;
                    r.s = s_sum[lq];
                    r.c = s_cor[lq];
                    r.merge(a.s, a.c);
                    s_sum[lq] = r.s;
                    s_cor[lq] = r.c;
                    s_pos[lq] = upto;
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < rows; q += kThreads) kp.Y[kp.rowid[it.r0 + q]] = s_sum[q] + s_cor[q];
}

// prof (profiling launches only, else null): prof[2 i] / prof[2 i + 1] get
// the device wall clock when the block starts item i and when every wave of
// the block has finished it.
template <int K, int ST, typename CT>
__global__ __launch_bounds__(kThreads) void k_batch_spmm(kparams kp)
{
    __shared__ __align__(16) unsigned char lds[kLdsBytes];
    __shared__ float s_part[2][kWaves][64];
    __shared__ uint32_t s_t;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = lane % K;  // right-hand side of this lane
    uint32_t prev = 0xffffffffu;
    for (;;) {
        // The barrier after the atomic also orders the previous item's LDS
        // reads before this item's staging writes.
        if (threadIdx.x == 0) s_t = atomicAdd(kp.counter, 1u);
        __syncthreads();
        const uint32_t i = s_t;
        __syncthreads();
        if (kp.prof && threadIdx.x == 0) {
            const uint64_t now = wall_clock64();
            if (prev != 0xffffffffu) kp.prof[2 * (uint64_t)prev + 1] = now;
            if (i < kp.n_items) kp.prof[2 * (uint64_t)i] = now;
        }
        prev = i;
        if (i >= kp.n_items) break;  // block-uniform: the whole block leaves
        const item_t it = kp.items[i];
        if constexpr (K == 1) {
            if (it.flags & kItemTwoPass) {
                run_item_two_pass<ST, CT>(kp, it, lds, lane, w);
                continue;
            }
        }
        const bool xs = it.xcols != 0, staged = (it.flags & kItemStaged) != 0;
        const uint32_t xbytes = xs ? xblock_bytes(it.xcols * K) : 0u;
        if (xs) stage_array<stage_depth<K>()>(lds, 0, kp.X, (uint64_t)it.cbase * K, it.xcols * K);
        if (staged) {
            const uint32_t rows = it.r1 - it.r0;
            const uint32_t nnz = kp.off[it.r1] - it.ebase;
            const lds_layout<ST, CT> L(xbytes, rows, nnz);
            uint32_t *s_off = (uint32_t *)(lds + L.off_o);
            uint32_t *s_rid = (uint32_t *)(lds + L.rid_o);
            // the entry copies first: the row arrays go through registers, and
            // waiting for their loads waits for every copy issued before them
            stage_array<stage_depth<K>()>(lds, L.col_o, (const CT *)kp.col, it.ebase, nnz);
            if constexpr (ST == SPIRA_VAL_F32) {
                stage_array<stage_depth<K>()>(lds, L.v1_o, kp.f32, it.ebase, nnz);
            } else {
                stage_array<stage_depth<K>()>(lds, L.v1_o, kp.hi, it.ebase, nnz);
                if constexpr (ST == SPIRA_VAL_SPLIT_BF16)
                    stage_array<stage_depth<K>()>(lds, L.v2_o, kp.lo, it.ebase, nnz);
            }
            for (uint32_t q = threadIdx.x; q <= rows; q += kThreads) s_off[q] = kp.off[it.r0 + q] - it.ebase;
            for (uint32_t q = threadIdx.x; q < rows; q += kThreads) s_rid[q] = kp.rowid[it.r0 + q];
        }
        if (xs || staged) {
            stage_wait<stage_depth<K>()>();
            __syncthreads();
        }
        if (xs) {
            if (staged) run_item<K, ST, CT, true, true>(kp, it, lds, s_part, lane, w, j);
            else run_item<K, ST, CT, true, false>(kp, it, lds, s_part, lane, w, j);
        } else {
            if (staged) run_item<K, ST, CT, false, true>(kp, it, lds, s_part, lane, w, j);
            else run_item<K, ST, CT, false, false>(kp, it, lds, s_part, lane, w, j);
        }
    }
}

// Copy matrix m's rows into the batch in stored order: stored row q (global)
// takes source row perm[q] - row_base; 16 lanes per row.
template <typename CT>
__global__ void k_gather_rows(const uint64_t *__restrict__ src_off, const uint32_t *__restrict__ src_col,
                              const double *__restrict__ src_val, uint32_t row_base, uint32_t n,
                              uint32_t col_base, const uint32_t *__restrict__ perm,
                              const uint32_t *__restrict__ dst_off, int storage, CT *__restrict__ ocol,
                              float *__restrict__ of32, uint16_t *__restrict__ ohi, uint16_t *__restrict__ olo)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = (uint32_t)(g / 16);
    const uint32_t l = (uint32_t)(g % 16);
    if (q >= n) return;
    const uint32_t r = perm[row_base + q] - row_base;
    const uint64_t s0 = src_off[r], s1 = src_off[r + 1];
    const uint32_t d0 = dst_off[row_base + q];
    for (uint64_t i = s0 + l; i < s1; i += 16) {
        const uint32_t d = d0 + (uint32_t)(i - s0);
        ocol[d] = (CT)(sizeof(CT) == 2 ? src_col[i] : src_col[i] + col_base);
        const float f = (float)src_val[i];
        if (storage == SPIRA_VAL_F32) {
            of32[d] = f;
        } else {
            const uint32_t u = __float_as_uint(f);
            const uint32_t ru = u + 0x7fffu + ((u >> 16) & 1u);
            const uint16_t h = (u & 0x7fffffffu) > 0x7f800000u ? (uint16_t)((u >> 16) | 0x40u) : (uint16_t)(ru >> 16);
            ohi[d] = h;
            if (storage == SPIRA_VAL_SPLIT_BF16) {
                const float rest = (float)(src_val[i] - (double)bf16f(h));
                const uint32_t v = __float_as_uint(rest);
                const uint32_t rv = v + 0x7fffu + ((v >> 16) & 1u);
                olo[d] = (v & 0x7fffffffu) > 0x7f800000u ? (uint16_t)((v >> 16) | 0x40u) : (uint16_t)(rv >> 16);
            }
        }
    }
}

__global__ void k_cat_off(const uint64_t *__restrict__ off, uint64_t n, uint64_t nnz_base, uint64_t row_base,
                          uint64_t *__restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x)
        out[row_base + i] = off[i] + nnz_base;
}

inline uint32_t len_class(uint64_t len)
{
    uint32_t c = 0;
    while (len) {
        ++c;
        len >>= 1;
    }
    return c;
}

}  // namespace

}  // namespace spira_amd

using namespace spira_amd;

struct spira_batch {
    spira_dev *dev = nullptr;
    uint32_t n_mats = 0;
    int storage = SPIRA_VAL_F32;
    bool col16 = true;
    uint64_t n_rows = 0, n_cols = 0, nnz = 0;
    std::vector<uint64_t> row_base, col_base, nnz_base;  // n_mats+1 each
    std::vector<uint32_t> h_off;                          // stored-order offsets (host copy, tile building)
    std::vector<uint32_t> class_start;                    // n_mats * (kClasses+1): first stored row of each class
    uint32_t *off32 = nullptr;
    uint32_t *rowid = nullptr;
    void *col = nullptr;
    float *f32 = nullptr;
    uint16_t *hi = nullptr, *lo = nullptr;
    // schedule
    uint32_t k = 0;
    std::vector<spira_batch_sched> sched;
    tile_t *tiles = nullptr;
    item_t *items = nullptr;
    size_t tiles_bytes = 0, items_bytes = 0;
    uint32_t n_tiles = 0, n_items = 0;
    std::vector<uint32_t> item_mat;  // matrix of each queued item
    uint32_t *counter = nullptr;
    // captured graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<void *, size_t>> owned;
    void *alloc(size_t bytes)
    {
        void *p = dev_alloc(dev, bytes);
        owned.push_back({p, bytes ? bytes : 8});
        return p;
    }
};

namespace spira_amd {

static uint32_t max_elog2(uint32_t K)
{
    uint32_t e = 0;
    while ((1u << (e + 1)) * K <= 64u) ++e;
    return e;
}

template <int ST, typename CT>
static uint32_t staged_bytes(uint32_t xbytes, uint32_t rows, uint32_t nnz)
{
    return lds_layout<ST, CT>(xbytes, rows, nnz).end_o;
}

static uint32_t staged_bytes_rt(int storage, bool col16, uint32_t xbytes, uint32_t rows, uint32_t nnz)
{
    if (col16) {
        if (storage == SPIRA_VAL_F32) return staged_bytes<SPIRA_VAL_F32, uint16_t>(xbytes, rows, nnz);
        if (storage == SPIRA_VAL_BF16) return staged_bytes<SPIRA_VAL_BF16, uint16_t>(xbytes, rows, nnz);
        return staged_bytes<SPIRA_VAL_SPLIT_BF16, uint16_t>(xbytes, rows, nnz);
    }
    if (storage == SPIRA_VAL_F32) return staged_bytes<SPIRA_VAL_F32, uint32_t>(xbytes, rows, nnz);
    if (storage == SPIRA_VAL_BF16) return staged_bytes<SPIRA_VAL_BF16, uint32_t>(xbytes, rows, nnz);
    return staged_bytes<SPIRA_VAL_SPLIT_BF16, uint32_t>(xbytes, rows, nnz);
}

static void batch_build_work(spira_batch *b)
{
    const uint32_t K = b->k;
    const uint32_t emax = max_elog2(K);
    // first length class (rows of 2^(c-1) nonzeros and more) worked on by the
    // whole block: 2^(c-1) >= 32 * (64 / k) nonzeros per wavefront lane slot
    const uint32_t long_class = emax + 6;
    std::vector<tile_t> tl;
    struct pending {
        item_t it;
        uint64_t work;
        uint32_t mat;
    };
    std::vector<pending> its;
    const std::vector<uint32_t> &off = b->h_off;
    for (uint32_t m = 0; m < b->n_mats; ++m) {
        const spira_batch_sched &s = b->sched[m];
        const uint64_t budget = std::max<uint32_t>(s.item_nnz, 1u);
        const uint32_t ncols = (uint32_t)(b->col_base[m + 1] - b->col_base[m]);
        // k = 1 X blocks too wide for one LDS load go in two halves (rows of
        // whole-block tiles excepted: those read X from global memory)
        const bool two_pass = s.xlds && K == 1 && ncols > SPIRA_BATCH_XLDS_FLOATS;
        const uint32_t xcols = (s.xlds && !two_pass) ? ncols : 0u;  // single-load X block
        const uint32_t xbytes = xblock_bytes(xcols * K);
        const bool stage = !two_pass && xbytes + kMinEntryLds <= kLdsBytes;
        const uint32_t *cs = &b->class_start[(size_t)m * (kClasses + 1)];
        // cost of a run of rows = nonzeros + rows (an empty row still writes k outputs)
        item_t cur{(uint32_t)tl.size(), 0, (uint32_t)b->col_base[m], xcols, 0, 0, 0, 0};
        uint32_t ir0 = 0;  // first stored row of the open item
        bool open = false, item_tp = false;
        uint64_t cur_cost = 0;
        // the open item's rows [ir0, r_end) fit in LDS beside the X block
        auto fits = [&](uint32_t r_end) {
            const uint32_t rows = r_end - ir0, nnz = off[r_end] - off[ir0];
            if (rows > kLdsBytes / 8 || nnz > kLdsBytes / 2) return false;
            return staged_bytes_rt(b->storage, b->col16, xbytes, rows, nnz) <= kLdsBytes;
        };
        auto close_item = [&](uint32_t r_end, bool may_stage) {
            cur.t1 = (uint32_t)tl.size();
            if (cur.t1 > cur.t0) {
                cur.r0 = ir0;
                cur.r1 = r_end;
                cur.ebase = off[ir0];
                cur.xcols = item_tp ? ncols : xcols;
                cur.flags = item_tp ? kItemTwoPass : ((may_stage && fits(r_end)) ? kItemStaged : 0u);
                its.push_back({cur, (cur_cost + cur.xcols) * K, m});
            }
            cur.t0 = (uint32_t)tl.size();
            cur_cost = 0;
            open = false;
        };
        for (uint32_t c = 0; c < kClasses; ++c) {
            const uint32_t a = cs[c], z = cs[c + 1];
            if (a == z) continue;
            // rows that would give a lane slot of one wavefront more than 32
            // nonzeros are whole-block tiles of one row each
            const bool lng = c >= long_class;
            const uint32_t meta = lng ? kTileLong : std::min<uint32_t>(c > s.plog2 ? c - s.plog2 : 0u, emax);
            if (open && item_tp && lng) close_item(a, false);  // two-pass items hold no whole-block rows
            uint32_t p = a;
            while (p < z) {
                if (!open) {
                    ir0 = p;
                    open = true;
                    item_tp = two_pass && !lng;
                }
                // extend the tile while the item stays within its budget and,
                // when staged, within the LDS (two-pass items: within the rows
                // their LDS records hold)
                uint32_t q = p;
                while (q < z && (!lng || q == p) && cur_cost + (off[q + 1] - off[p]) + (q + 1 - p) <= budget &&
                       (!stage || fits(q + 1)) && (!item_tp || q + 1 - ir0 <= kTwoPassRows))
                    ++q;
                if (q == p) {
                    if (p > ir0) {  // the item is full: close it and retry the row in a new one
                        close_item(p, stage);
                        continue;
                    }
                    q = p + 1;  // a lone row over budget
                }
                tl.push_back(tile_t{p, q, meta, 0});
                cur_cost += (off[q] - off[p]) + (q - p);
                p = q;
                if (cur_cost >= budget || (stage && p < z && !fits(p + 1)) || (item_tp && p - ir0 >= kTwoPassRows))
                    close_item(p, stage);
            }
        }
        if (open) close_item(cs[kClasses], stage);
    }
    // heaviest items first (longest-processing-time order for the queue)
    std::stable_sort(its.begin(), its.end(), [](const pending &x, const pending &y) { return x.work > y.work; });
    std::vector<item_t> items(its.size());
    b->item_mat.resize(its.size());
    for (size_t i = 0; i < its.size(); ++i) {
        items[i] = its[i].it;
        b->item_mat[i] = its[i].mat;
    }
    hipStream_t st = b->dev->stream;
    if (b->tiles || b->items) SPIRA_HIP(hipStreamSynchronize(st));
    if (b->tiles) dev_release(b->dev, b->tiles, b->tiles_bytes);
    if (b->items) dev_release(b->dev, b->items, b->items_bytes);
    b->tiles = nullptr;
    b->items = nullptr;
    b->tiles_bytes = std::max<size_t>(tl.size(), 1) * sizeof(tile_t);
    b->items_bytes = std::max<size_t>(items.size(), 1) * sizeof(item_t);
    b->tiles = (tile_t *)dev_alloc(b->dev, b->tiles_bytes);
    b->items = (item_t *)dev_alloc(b->dev, b->items_bytes);
    SPIRA_HIP(hipMemcpyAsync(b->tiles, tl.data(), tl.size() * sizeof(tile_t), hipMemcpyHostToDevice, st));
    SPIRA_HIP(hipMemcpyAsync(b->items, items.data(), items.size() * sizeof(item_t), hipMemcpyHostToDevice, st));
    dev_sync(b->dev);
    b->n_tiles = (uint32_t)tl.size();
    b->n_items = (uint32_t)items.size();
}

// One 1024-thread block per CU (the LDS area takes most of a CU's LDS); the
// grid is exactly the resident blocks, so every block is live from the start
// and a block-microsecond is a CU-microsecond.
template <int K, int ST, typename CT>
static void launch_k(spira_batch *b, const float *X, float *Y, uint64_t *prof)
{
    kparams kp{b->items, b->n_items, b->counter, b->tiles, b->off32, b->rowid, b->col, b->f32, b->hi, b->lo,
               b->nnz ? (uint32_t)(b->nnz - 1) : 0u, X, Y, prof};
    const unsigned grid = std::max(1u, std::min(b->n_items, (uint32_t)b->dev->n_cu));
    k_batch_spmm<K, ST, CT><<<grid, kThreads, 0, b->dev->stream>>>(kp);
}

template <int K, int ST>
static void launch_col(spira_batch *b, const float *X, float *Y, uint64_t *prof)
{
    if (b->col16) launch_k<K, ST, uint16_t>(b, X, Y, prof);
    else launch_k<K, ST, uint32_t>(b, X, Y, prof);
}

template <int K>
static void launch_st(spira_batch *b, const float *X, float *Y, uint64_t *prof)
{
    switch (b->storage) {
    case SPIRA_VAL_F32: launch_col<K, SPIRA_VAL_F32>(b, X, Y, prof); break;
    case SPIRA_VAL_BF16: launch_col<K, SPIRA_VAL_BF16>(b, X, Y, prof); break;
    default: launch_col<K, SPIRA_VAL_SPLIT_BF16>(b, X, Y, prof); break;
    }
}

static void batch_launch(spira_batch *b, const float *X, float *Y, uint64_t *prof = nullptr)
{
    k_reset_counter<<<1, 1, 0, b->dev->stream>>>(b->counter);
    switch (b->k) {
    case 1: launch_st<1>(b, X, Y, prof); break;
    case 4: launch_st<4>(b, X, Y, prof); break;
    case 16: launch_st<16>(b, X, Y, prof); break;
    default: launch_st<32>(b, X, Y, prof); break;
    }
    SPIRA_HIP(hipGetLastError());
}

static void batch_drop_graph(spira_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Built-in schedule of one matrix (the method without a selector).
static spira_batch_sched default_sched(uint64_t n_cols, uint32_t K)
{
    spira_batch_sched s{};
    s.plog2 = 2;
    s.xlds = (K == 1 && n_cols <= SPIRA_BATCH_XLDS_TWO_PASS_COLS) ? 1 : 0;
    s.item_nnz = std::max<uint32_t>(4096u, 131072u / K);
    return s;
}

}  // namespace spira_amd

namespace {
thread_local std::string g_batch_error;

template <class F>
int bguard(spira_dev *dev, F &&f)
{
    std::string *sink = dev ? &dev->last_error : &g_batch_error;
    try {
        if (dev) SPIRA_HIP(hipSetDevice(dev->device));
        f();
        return SPIRA_OK;
    } catch (const status_error &e) {
        *sink = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        *sink = "host allocation failed";
        return SPIRA_ERR_NOMEM;
    } catch (const std::exception &e) {
        *sink = e.what();
        return SPIRA_ERR_DEVICE;
    }
}
}  // namespace

extern "C" {

int spira_batch_create(spira_dev *dev, uint32_t n_mats, spira_dcsr *const *mats, int storage, spira_batch **out)
{
    if (!dev || !out || (n_mats && !mats)) return SPIRA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    spira_batch *b = new spira_batch();
    b->dev = dev;
    int rc = bguard(dev, [&] {
        if (storage != SPIRA_VAL_F32 && storage != SPIRA_VAL_BF16 && storage != SPIRA_VAL_SPLIT_BF16)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: unknown value storage");
        b->n_mats = n_mats;
        b->storage = storage;
        b->row_base.assign(n_mats + 1, 0);
        b->col_base.assign(n_mats + 1, 0);
        b->nnz_base.assign(n_mats + 1, 0);
        for (uint32_t m = 0; m < n_mats; ++m) {
            const spira_dcsr *A = mats[m];
            if (!A || !A->built) fail(SPIRA_ERR_LOGIC, "batch: every matrix must be locked");
            if (A->dev != dev) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: matrices must belong to this device handle");
            b->row_base[m + 1] = b->row_base[m] + A->n_rows;
            b->col_base[m + 1] = b->col_base[m] + A->n_cols;
            b->nnz_base[m + 1] = b->nnz_base[m] + A->nnz;
            if (A->n_cols > 65536) b->col16 = false;
        }
        b->n_rows = b->row_base[n_mats];
        b->n_cols = b->col_base[n_mats];
        b->nnz = b->nnz_base[n_mats];
        if (b->nnz >= (1ull << 32) || b->n_rows >= (1ull << 32) || b->n_cols >= (1ull << 32))
            fail(SPIRA_ERR_UNSUPPORTED, "batch: totals must stay below 2^32 rows, columns and nonzeros");
        hipStream_t st = dev->stream;
        // source offsets, concatenated, to the host
        const uint64_t nr = b->n_rows;
        uint64_t *cat = (uint64_t *)dev->ws.get(WS_MISC2, (nr + 1) * sizeof(uint64_t));
        for (uint32_t m = 0; m < n_mats; ++m) {
            const spira_dcsr *A = mats[m];
            k_cat_off<<<grid_for(A->n_rows + 1, 256, 1024), 256, 0, st>>>(A->off, A->n_rows, b->nnz_base[m],
                                                                         b->row_base[m], cat);
        }
        SPIRA_HIP(hipGetLastError());
        std::vector<uint64_t> src(nr + 1, 0);
        if (n_mats) SPIRA_HIP(hipMemcpyAsync(src.data(), cat, (nr + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        dev_sync(dev);
        // stored order: within each matrix, rows by length class (stable)
        std::vector<uint32_t> perm(nr);
        b->h_off.assign(nr + 1, 0);
        b->class_start.assign((size_t)n_mats * (kClasses + 1), 0);
        std::vector<uint32_t> cnt(kClasses + 1);
        for (uint32_t m = 0; m < n_mats; ++m) {
            const uint64_t r0 = b->row_base[m], r1 = b->row_base[m + 1];
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (uint64_t r = r0; r < r1; ++r) ++cnt[len_class(src[r + 1] - src[r]) + 1];
            for (uint32_t c = 0; c < kClasses; ++c) cnt[c + 1] += cnt[c];
            uint32_t *cs = &b->class_start[(size_t)m * (kClasses + 1)];
            for (uint32_t c = 0; c <= kClasses; ++c) cs[c] = (uint32_t)(r0 + cnt[c]);
            for (uint64_t r = r0; r < r1; ++r) perm[r0 + cnt[len_class(src[r + 1] - src[r])]++] = (uint32_t)r;
        }
        for (uint64_t q = 0; q < nr; ++q) b->h_off[q + 1] = b->h_off[q] + (uint32_t)(src[perm[q] + 1] - src[perm[q]]);
        src.clear();
        src.shrink_to_fit();
        // device arrays (entry arrays padded: items copy them in 16-byte spans)
        const uint64_t nz = b->nnz ? b->nnz : 1;
        b->off32 = (uint32_t *)b->alloc((nr + 1) * sizeof(uint32_t));
        b->rowid = (uint32_t *)b->alloc((nr ? nr : 1) * sizeof(uint32_t));
        b->col = b->alloc(nz * (b->col16 ? sizeof(uint16_t) : sizeof(uint32_t)) + kArrayPad);
        if (storage == SPIRA_VAL_F32) b->f32 = (float *)b->alloc(nz * sizeof(float) + kArrayPad);
        else b->hi = (uint16_t *)b->alloc(nz * sizeof(uint16_t) + kArrayPad);
        if (storage == SPIRA_VAL_SPLIT_BF16) b->lo = (uint16_t *)b->alloc(nz * sizeof(uint16_t) + kArrayPad);
        SPIRA_HIP(hipMemcpyAsync(b->off32, b->h_off.data(), (nr + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        if (nr) SPIRA_HIP(hipMemcpyAsync(b->rowid, perm.data(), nr * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        for (uint32_t m = 0; m < n_mats; ++m) {
            const spira_dcsr *A = mats[m];
            if (!A->n_rows || !A->nnz) continue;
            const unsigned grid = grid_for((uint64_t)A->n_rows * 16, 256);
            if (b->col16)
                k_gather_rows<uint16_t><<<grid, 256, 0, st>>>(A->off, A->col, A->val, (uint32_t)b->row_base[m],
                                                             (uint32_t)A->n_rows, (uint32_t)b->col_base[m],
                                                             b->rowid, b->off32, storage, (uint16_t *)b->col,
                                                             b->f32, b->hi, b->lo);
            else
                k_gather_rows<uint32_t><<<grid, 256, 0, st>>>(A->off, A->col, A->val, (uint32_t)b->row_base[m],
                                                             (uint32_t)A->n_rows, (uint32_t)b->col_base[m],
                                                             b->rowid, b->off32, storage, (uint32_t *)b->col,
                                                             b->f32, b->hi, b->lo);
        }
        SPIRA_HIP(hipGetLastError());
        b->counter = (uint32_t *)b->alloc(64);
        dev_sync(dev);  // perm (host) is read by the copies above until here
    });
    if (rc != SPIRA_OK) {
        for (auto &o : b->owned) dev_release(dev, o.first, o.second);
        delete b;
        return rc;
    }
    *out = b;
    return SPIRA_OK;
}

void spira_batch_destroy(spira_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->dev->device);
    (void)hipStreamSynchronize(b->dev->stream);
    batch_drop_graph(b);
    if (b->tiles) dev_release(b->dev, b->tiles, b->tiles_bytes);
    if (b->items) dev_release(b->dev, b->items, b->items_bytes);
    for (auto &o : b->owned) dev_release(b->dev, o.first, o.second);
    delete b;
}

int spira_batch_info(const spira_batch *b, uint64_t *n_rows, uint64_t *n_cols, uint64_t *nnz, uint32_t *n_tiles)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    if (n_rows) *n_rows = b->n_rows;
    if (n_cols) *n_cols = b->n_cols;
    if (nnz) *nnz = b->nnz;
    if (n_tiles) *n_tiles = b->n_tiles;
    return SPIRA_OK;
}

int spira_batch_bases(const spira_batch *b, uint64_t *row_base, uint64_t *col_base)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    if (row_base) std::memcpy(row_base, b->row_base.data(), (b->n_mats + 1) * sizeof(uint64_t));
    if (col_base) std::memcpy(col_base, b->col_base.data(), (b->n_mats + 1) * sizeof(uint64_t));
    return SPIRA_OK;
}

int spira_batch_set_schedule(spira_batch *b, uint32_t k, const spira_batch_sched *per_matrix)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (k != 1 && k != 4 && k != 16 && k != 32)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: k must be 1, 4, 16 or 32");
        std::vector<spira_batch_sched> s(b->n_mats);
        for (uint32_t m = 0; m < b->n_mats; ++m) {
            const uint64_t nc = b->col_base[m + 1] - b->col_base[m];
            s[m] = per_matrix ? per_matrix[m] : default_sched(nc, k);
            if (s[m].plog2 > SPIRA_BATCH_MAX_PLOG2)
                fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: plog2 must be at most SPIRA_BATCH_MAX_PLOG2");
            if (s[m].xlds > 1) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: xlds must be 0 or 1");
            if (s[m].xlds && nc * k > SPIRA_BATCH_XLDS_FLOATS && !(k == 1 && nc <= SPIRA_BATCH_XLDS_TWO_PASS_COLS))
                fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: X block of this matrix does not fit in LDS at this k");
            if (s[m].item_nnz == 0) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: item_nnz must be positive");
        }
        b->k = k;
        b->sched.swap(s);
        batch_drop_graph(b);
        batch_build_work(b);
    });
}

int spira_batch_spmm(spira_batch *b, const float *X, float *Y, uint32_t flags)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (!b->k) fail(SPIRA_ERR_LOGIC, "batch: set a schedule first");
        if ((b->n_cols && !X) || (b->n_rows && !Y)) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: X or Y is NULL");
        batch_launch(b, X, Y);
        if (!(flags & SPIRA_NO_SYNC)) dev_sync(b->dev);
    });
}

int spira_batch_profile(spira_batch *b, const float *X, float *Y, double *matrix_cost_us)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (!b->k) fail(SPIRA_ERR_LOGIC, "batch: set a schedule first");
        if ((b->n_cols && !X) || (b->n_rows && !Y)) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: X or Y is NULL");
        if (!matrix_cost_us && b->n_mats) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: matrix_cost_us is NULL");
        int rate_khz = 0;
        SPIRA_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, b->dev->device));
        if (rate_khz <= 0) fail(SPIRA_ERR_DEVICE, "batch: device wall clock rate unavailable");
        const size_t pbytes = std::max<size_t>(2 * (size_t)b->n_items, 2) * sizeof(uint64_t);
        uint64_t *prof = (uint64_t *)dev_alloc(b->dev, pbytes);
        std::vector<uint64_t> h(2 * (size_t)b->n_items);
        try {
            batch_launch(b, X, Y, prof);
            if (b->n_items)
                SPIRA_HIP(hipMemcpyAsync(h.data(), prof, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                         b->dev->stream));
            dev_sync(b->dev);
        } catch (...) {
            (void)hipStreamSynchronize(b->dev->stream);
            dev_release(b->dev, prof, pbytes);
            throw;
        }
        dev_release(b->dev, prof, pbytes);
        const double us_per_tick = 1e3 / (double)rate_khz;
        std::fill(matrix_cost_us, matrix_cost_us + b->n_mats, 0.0);
        for (uint32_t i = 0; i < b->n_items; ++i)
            matrix_cost_us[b->item_mat[i]] += (double)(h[2 * i + 1] - h[2 * i]) * us_per_tick;
    });
}

int spira_batch_graph_capture(spira_batch *b, const float *X, float *Y)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (!b->k) fail(SPIRA_ERR_LOGIC, "batch: set a schedule first");
        if ((b->n_cols && !X) || (b->n_rows && !Y)) fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: X or Y is NULL");
        batch_drop_graph(b);
        hipStream_t st = b->dev->stream;
        SPIRA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            batch_launch(b, X, Y);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SPIRA_HIP(hipStreamEndCapture(st, &b->graph));
        SPIRA_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    });
}

int spira_batch_graph_launch(spira_batch *b, uint32_t flags)
{
    if (!b) return SPIRA_ERR_INVALID_ARGUMENT;
    return bguard(b->dev, [&] {
        if (!b->exec) fail(SPIRA_ERR_LOGIC, "batch: no captured graph");
        SPIRA_HIP(hipGraphLaunch(b->exec, b->dev->stream));
        if (!(flags & SPIRA_NO_SYNC)) dev_sync(b->dev);
    });
}

}  // extern "C"
