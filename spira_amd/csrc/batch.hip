// batch.hip -- many matrices of uneven size, k right-hand sides (k = 1, 4, 16
// or 32), one persistent launch fed by device work queues; the repeated call
// can be captured in a hipGraph.
//
// Layout.  A batch is one block-diagonal CSR whose rows are stored, within
// each matrix, sorted by length class (class c holds rows of 2^(c-1) ..
// 2^c - 1 nonzeros; class 0 the empty rows), each class in ascending row
// order; rowid[p] maps a stored row back to its row of the block-diagonal
// system, so X (sum of n_cols x k, row major) and Y (sum of n_rows x k) keep
// the caller's order.  Every stored row is padded to whole groups of 4
// entries, so a lane reads a group's values with one 16-byte (FP32) or 8-byte
// (BF16) load and its columns with one 8-byte (u16) or 16-byte (u32) load.  A
// padding entry has value 0 and column n_cols of its matrix (0 for a matrix
// of 65536 columns): the X block staged in LDS carries a zero row there, and
// the paths that read X from global memory mask entries past the row's real
// length, which the top 2 bits of rowid hold (its padding count).  Columns
// are stored as u16 offsets inside their matrix when every matrix has at
// most 65536 columns (u32 batch columns otherwise); values as FP32, BF16 or
// split BF16.
//
// Work.  A schedule gives every matrix a lanes-per-row rule (2^plog2 target
// nonzeros per lane: lanes per row = 2^class / 2^plog2, at most 64 / k), a
// work-item size, whether its X block is staged in LDS and whether its
// entries are.  A tile is a run of stored rows of one class (one lanes-per-row
// value: rows per wavefront = 64 / (lanes * k)); rows long enough to give a
// wavefront's lane slot more than 32 nonzeros (from 2048 / k nonzeros) are
// tiles of their own, worked on by the whole block.  A work item is a run of
// consecutive stored rows of one matrix (its tiles) of about item_nnz
// nonzeros + rows and is the unit of the device queues.  There is one queue
// per XCD (workgroups are dealt to the 8 XCDs round robin, so block b lives
// on XCD b % 8): matrices are dealt whole to the queues, heaviest matrix to
// the least loaded queue, and a queue holds its matrices' items matrix after
// matrix, so the CUs of an XCD work on one or two matrices at a time and
// share their X block in that XCD's L2.  Every block of the persistent grid
// (one 1024-thread block per CU) takes the next item of its own queue with
// one atomic add; when that queue is empty it takes items from the next
// queues in turn, and it leaves when all 8 are empty.
//
// LDS.  A block first copies what its item reads into LDS with 16-byte
// loads, all in flight at once: the matrix's X block when the schedule
// stages it (n_cols * k floats, up to SPIRA_BATCH_XLDS_FLOATS, then the zero
// row), and the item's row bounds, row ids, columns and values when they fit
// beside it (not for SPIRA_BATCH_STREAM schedules).  The rows are then
// computed from LDS, so the only global reads left in the compute phase are
// the X gathers of unstaged X blocks.  Items whose entries are not staged
// (streamed schedules, long rows, or an X block that fills the LDS) read
// their entries from global memory with the next rows' bounds loaded a step
// ahead.
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
constexpr uint32_t kGroup = 4;          // entries per group; rows are padded to whole groups
constexpr uint32_t kLdsBytes = 155584;  // LDS for the X block and staged entries
constexpr uint32_t kMinEntryLds = 65536;  // stage entries only with at least this much LDS left
constexpr uint32_t kClasses = 34;         // bit length of a u32 row length, plus one
constexpr uint32_t kTileLong = 0x80u;
constexpr uint32_t kItemStaged = 1u;     // item flag: entries staged in LDS
constexpr uint32_t kItemTwoPass = 2u;    // item flag: k = 1, X block staged in two column halves
constexpr uint32_t kQueues = 8;          // work queues: one per XCD
constexpr uint32_t kCounterStride = 16;  // u32 words between queue counters (one 64-byte line each)
constexpr uint32_t kIdBits = 30;         // rowid word: row (30 bits) | padding entries << 30
constexpr uint32_t kIdMask = (1u << kIdBits) - 1;
constexpr size_t kArrayPad = 64;  // zeroed bytes after each entry array (16-byte staging reads)
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

// LDS bytes of an X block of n rows of k floats plus its zero row: it is
// copied as the 16-byte aligned span around its elements (see stage_array),
// so it may start up to 12 bytes into its area.
inline __host__ __device__ uint32_t xblock_bytes(uint32_t n, uint32_t k) { return n ? align16((n + 1) * k * 4) + 32 : 0u; }
static_assert((SPIRA_BATCH_XLDS_FLOATS + 32) * 4 + 48 <= kLdsBytes, "X block must fit the LDS area");

// Two-pass items: each half of the X block (at most kHalfFloats floats) sits
// at the start of the LDS area; each row's (sum, correction) record between
// the passes is kept in global memory (2 words per stored row).
constexpr uint32_t kHalfFloats = 32768;
constexpr uint32_t kHalfBytes = kHalfFloats * 4 + 32;
static_assert(kHalfBytes <= kLdsBytes, "an X half must fit the LDS area");

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
    uint32_t gbase;   // first group of row r0
    uint32_t flags;   // kItemStaged, kItemTwoPass; bits 2-31: the matrix's zero group
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
    __host__ __device__ lds_layout(uint32_t xbytes, uint32_t rows, uint32_t ents)
    {
        off_o = align16(xbytes);
        rid_o = off_o + 4 * (rows + 1);
        col_o = align16(rid_o + 4 * rows);
        v1_o = col_o + span(ents, (uint32_t)sizeof(CT));
        v2_o = v1_o + span(ents, val_size);
        end_o = n_val_arrays == 2 ? v2_o + span(ents, val_size) : v2_o;
    }
};

// The error-free transformations below must run as written: contraction off,
// or the compiler fuses s + a * x into one FMA and the correction counts the
// product's rounding error twice.
struct acc2 {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float a, float x)
    {
#pragma clang fp contract(off)
        const float p = a * x;
        const float e = fmaf(a, x, -p);
        const float t = s + p;
        const float bp = t - s;
        c += ((s - (t - bp)) + (p - bp)) + e;
        s = t;
    }
    __device__ __forceinline__ void merge(float s2, float c2)
    {
#pragma clang fp contract(off)
        const float t = s + s2;
        const float bp = t - s;
        c += c2 + ((s - (t - bp)) + (s2 - bp));
        s = t;
    }
};

__global__ void k_reset_counters(uint32_t *counter)
{
    if (threadIdx.x < kQueues) counter[threadIdx.x * kCounterStride] = 0u;
}

// Groups of entries (4 columns, 4 values) from the global arrays or the
// item's LDS copy, one storage format each.  Global reads are ordinary loads:
// with X in LDS nothing of the batch's is worth keeping in L2, and measured
// on MI355X the non-temporal form is slower here (k = 1, 4096 matrices: 3.68
// against 3.20 ms).
template <int ST, typename CT, bool LDS>
struct gsrc {
    const CT *col;
    const float *f32;
    const uint16_t *hi, *lo;
    template <typename V>
    static __device__ __forceinline__ V ld(const V *p)
    {
        return *p;
    }
    static __device__ __forceinline__ void bf16x4(const u32x2 w, float (&v)[kGroup])
    {
        v[0] = __uint_as_float(w.x << 16);
        v[1] = __uint_as_float(w.x & 0xffff0000u);
        v[2] = __uint_as_float(w.y << 16);
        v[3] = __uint_as_float(w.y & 0xffff0000u);
    }
    __device__ __forceinline__ void get(uint32_t g, uint32_t (&c)[kGroup], float (&v)[kGroup]) const
    {
        if constexpr (sizeof(CT) == 2) {
            const u32x2 w = ld((const u32x2 *)col + g);
            c[0] = w.x & 0xffffu;
            c[1] = w.x >> 16;
            c[2] = w.y & 0xffffu;
            c[3] = w.y >> 16;
        } else {
            const u32x4 w = ld((const u32x4 *)col + g);
            c[0] = w.x;
            c[1] = w.y;
            c[2] = w.z;
            c[3] = w.w;
        }
        if constexpr (ST == SPIRA_VAL_F32) {
            const u32x4 w = ld((const u32x4 *)f32 + g);
            v[0] = __uint_as_float(w.x);
            v[1] = __uint_as_float(w.y);
            v[2] = __uint_as_float(w.z);
            v[3] = __uint_as_float(w.w);
        } else {
            bf16x4(ld((const u32x2 *)hi + g), v);
            if constexpr (ST == SPIRA_VAL_SPLIT_BF16) {
                float l[kGroup];
                bf16x4(ld((const u32x2 *)lo + g), l);
#pragma unroll
                for (uint32_t u = 0; u < kGroup; ++u) v[u] += l[u];
            }
        }
    }
};

// X element of stored column c for right-hand side j.  XS: the matrix's X
// block in LDS, whose zero row at column n_cols serves the padding entries,
// read unmasked; otherwise global X, read only where ok (a masked lane reads
// element j of X and gets 0).
template <int K, typename CT, bool XS>
struct xsrc {
    const float *X;  // global X, or the LDS block
    uint32_t cbase;
    __device__ __forceinline__ float operator()(uint32_t c, bool ok, int j) const
    {
        const uint32_t lc = (sizeof(CT) == 2) ? c : c - cbase;  // column inside the matrix
        if constexpr (XS) {
            return X[lc * K + j];
        } else {
            const uint64_t gi = ((uint64_t)cbase + lc) * K + j;
            const float v = X[ok ? gi : (uint64_t)j];
            return ok ? v : 0.0f;
        }
    }
};

// X element of a two-pass item's half in LDS: xs holds columns [c0, c1) of
// the matrix; read only where ok (a masked lane reads element 0 and gets 0).
template <typename CT>
struct xsrc_half {
    const float *xs;
    uint32_t cbase, c0;
    __device__ __forceinline__ float operator()(uint32_t c, bool ok, int) const
    {
        const uint32_t lc = ((sizeof(CT) == 2) ? c : c - cbase) - c0;
        const float v = xs[ok ? lc : 0u];
        return ok ? v : 0.0f;
    }
};

// One row as a lane group sees it: groups [g0, g1) hold its entries, which
// are entries [lo, lo + len) (entry e of group g is entry 4 g + e); id is its
// output slot.
struct row_rng {
    uint32_t g0, g1;
    uint32_t lo, len;
    uint32_t id;
};

// The 4 entries of group g of row r (MASK: zeroed where outside the row's
// entries) times their X elements, added to a.
template <bool MASK, typename XV>
__device__ __forceinline__ void group_dot(acc2 &a, const row_rng &r, uint32_t g, const uint32_t (&c)[kGroup],
                                          float (&v)[kGroup], const XV &xv, int j)
{
    float x[kGroup];
    if constexpr (MASK) {
        const uint32_t q = kGroup * g - r.lo;  // compared unsigned: wraps below lo
#pragma unroll
        for (uint32_t u = 0; u < kGroup; ++u) {
            const bool m = q + u < r.len;
            x[u] = xv(c[u], m, j);
            v[u] = m ? v[u] : 0.0f;
        }
    } else {
#pragma unroll
        for (uint32_t u = 0; u < kGroup; ++u) x[u] = xv(c[u], true, j);
    }
#pragma unroll
    for (uint32_t u = 0; u < kGroup; ++u) a.add(v[u], x[u]);
}

// Two rows of one lane group: lane e of the group takes groups e, e + E, ...
// of each; per step one group of each row is loaded.  An exhausted row's
// lanes add exact zeros, so a step is straight-line code: without MASK they
// load group gz, the zero group (value 0, column of the X block's zero row);
// with MASK they load group gz (a group of the item) and mask all of it.
template <int K, bool MASK, typename GS, typename XV>
__device__ __forceinline__ void rows2_groups(acc2 &a0, acc2 &a1, const row_rng &r0, const row_rng &r1, uint32_t e,
                                             uint32_t E, uint32_t gz, const GS &gs, const XV &xv, int j)
{
    for (uint32_t g0 = r0.g0 + e, g1 = r1.g0 + e; g0 < r0.g1 || g1 < r1.g1; g0 += E, g1 += E) {
        uint32_t c0[kGroup], c1[kGroup];
        float v0[kGroup], v1[kGroup];
        gs.get(g0 < r0.g1 ? g0 : gz, c0, v0);
        gs.get(g1 < r1.g1 ? g1 : gz, c1, v1);
        group_dot<MASK>(a0, r0, g0, c0, v0, xv, j);
        group_dot<MASK>(a1, r1, g1, c1, v1, xv, j);
    }
}

// Lane i's value of lane i + O inside its row of 16 lanes (DPP row_shl; the
// last O lanes of a row read 0, which no merge below uses).
template <uint32_t O>
__device__ __forceinline__ float lane_up(float v)
{
    static_assert(O >= 1 && O < 16, "DPP row shifts stay inside a row of 16 lanes");
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + O, 0xf, 0xf, false));
}

template <uint32_t O>
__device__ __forceinline__ void merge_up(acc2 &a, uint32_t span)
{
    if constexpr (O < 16) {
        if (span > O) a.merge(lane_up<O>(a.s), lane_up<O>(a.c));
    }
}

// Merge the E lanes of a row (K apart) into the row's first lane.  Round o
// adds lane i + o to lane i: for the first lane this is the xor butterfly's
// order, so the result does not depend on how a round moves the values.
// Offsets inside a row of 16 lanes use DPP, the rest (E * K > 16) ds_bpermute.
template <int K>
__device__ __forceinline__ void merge_lanes(acc2 &a, uint32_t E)
{
    const uint32_t span = (uint32_t)K * E;
    merge_up<(uint32_t)K>(a, span);
    merge_up<2u * K>(a, span);
    merge_up<4u * K>(a, span);
    merge_up<8u * K>(a, span);
    for (uint32_t o = K < 16 ? 16u : (uint32_t)K; o < span; o <<= 1)
        a.merge(__shfl_xor(a.s, (int)o, 64), __shfl_xor(a.c, (int)o, 64));
}

// Rows p, p + S, p + 2 S, ... (< r1) of one lane group, two per step, the
// next step's bounds loaded ahead.  bnd(q) gives row q's range; out(q, rng,
// acc) takes each finished row (lanes e == 0 only).
template <int K, bool MASK, typename GS, typename XV, typename BND, typename OUT>
__device__ __forceinline__ void rows_loop(uint32_t p, uint32_t r1, uint32_t S, uint32_t e, uint32_t E,
                                          uint32_t gz, const GS &gs, const XV &xv, int j, const BND &bnd,
                                          const OUT &out)
{
    const row_rng none{0u, 0u, 0u, 0u, 0u};
    auto ref = [&](uint32_t q) { return q < r1 ? bnd(q) : none; };
    row_rng a0 = ref(p), a1 = ref(p + S);
    while (p < r1) {
        const row_rng n0 = ref(p + 2 * S), n1 = ref(p + 3 * S);
        acc2 s0, s1;
        rows2_groups<K, MASK>(s0, s1, a0, a1, e, E, gz, gs, xv, j);
        merge_lanes<K>(s0, E);
        merge_lanes<K>(s1, E);
        if (e == 0) {
            out(p, a0, s0);
            if (p + S < r1) out(p + S, a1, s1);
        }
        p += 2 * S;
        a0 = n0;
        a1 = n1;
    }
}

// Range of stored row q from its group bounds g0 < g1 and rowid word.
__device__ __forceinline__ row_rng row_of(uint32_t g0, uint32_t g1, uint32_t rid)
{
    return row_rng{g0, g1, kGroup * g0, kGroup * (g1 - g0) - (rid >> kIdBits), rid & kIdMask};
}

// Copy nbytes (a multiple of 16) from 16-byte aligned global src to LDS dst
// with D 16-byte loads in flight per lane, or (D = 0) straight into LDS: one
// wavefront instruction moves 64 consecutive chunks, lane l's to dst + 16 l
// (lanes past the end masked off); the caller waits with stage_wait<0>().
// Ordinary (not non-temporal) loads, as for the entries.
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
                                                 (__attribute__((address_space(3))) void *)(d + c0), 16, 0, 0);
        return;
    }
    uint32_t i = threadIdx.x;
    if constexpr (D > 1) {
        for (; i + (D - 1) * kThreads < n; i += D * kThreads) {
            u32x4 v[D];
#pragma unroll
            for (uint32_t u = 0; u < D; ++u) v[u] = s[i + u * kThreads];
#pragma unroll
            for (uint32_t u = 0; u < D; ++u) d[i + u * kThreads] = v[u];
        }
    }
    for (; i < n; i += kThreads) d[i] = s[i];
}

// Stage one array: the 16-byte aligned span holding elements [base, base + n)
// of arr, D loads in flight per lane; returns the LDS pointer to element base.
template <uint32_t D, typename T>
__device__ __forceinline__ T *stage_array(unsigned char *lds, uint32_t o, const T *arr, uint64_t base, uint32_t n)
{
    const uintptr_t a = (uintptr_t)(arr + base);
    const uintptr_t a0 = a & ~(uintptr_t)15;
    const uintptr_t a1 = ((uintptr_t)(arr + base + n) + 15) & ~(uintptr_t)15;
    stage16<D>(lds + o, (const unsigned char *)a0, (uint32_t)(a1 - a0));
    return (T *)(lds + o + (uint32_t)(a - a0));
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
    uint32_t *counter;     // kQueues counters, kCounterStride words apart
    const uint32_t *qbeg;  // queue q holds items [qbeg[q], qbeg[q + 1])
    const uint32_t *mid;   // two-pass rows: first entry of the second column half (null without two-pass items)
    uint32_t *rec;         // two-pass records: 2 words per stored row (null without two-pass items)
    const tile_t *tiles;
    const uint32_t *goff;  // group offsets of the stored rows
    const uint32_t *rowid;
    const void *col;
    const float *f32;
    const uint16_t *hi, *lo;
    const float *X;
    float *Y;
    uint64_t *prof;
};

// All tiles of one item; XS: X block in LDS, ESTAGED: entries in LDS.
template <int K, int ST, typename CT, bool XS, bool ESTAGED>
__device__ __forceinline__ void run_item(const kparams &kp, const item_t &it, unsigned char *lds,
                                         float (*s_part)[kWaves][64], int lane, int w, int j)
{
    constexpr bool MASK = !XS;
    const xsrc<K, CT, XS> xv{XS ? staged_at(lds, 0, kp.X, (uint64_t)it.cbase * K) : kp.X, it.cbase};
    const uint32_t ng = kp.goff[it.r1] - it.gbase;  // groups of the item
    const uint32_t *off = kp.goff, *rowid = kp.rowid;
    uint32_t rbase = 0, gb = 0;  // row and group numbering of off / rowid and the entry source
    gsrc<ST, CT, ESTAGED> gs{};
    if constexpr (ESTAGED) {
        const lds_layout<ST, CT> L(XS ? xblock_bytes(it.xcols, K) : 0u, it.r1 - it.r0, ng * kGroup);
        off = (const uint32_t *)(lds + L.off_o);
        rowid = (const uint32_t *)(lds + L.rid_o);
        rbase = it.r0;
        const uint64_t e0 = (uint64_t)it.gbase * kGroup;
        gs.col = staged_at(lds, L.col_o, (const CT *)kp.col, e0);
        if constexpr (ST == SPIRA_VAL_F32) {
            gs.f32 = staged_at(lds, L.v1_o, kp.f32, e0);
        } else {
            gs.hi = staged_at(lds, L.v1_o, kp.hi, e0);
            if constexpr (ST == SPIRA_VAL_SPLIT_BF16) gs.lo = staged_at(lds, L.v2_o, kp.lo, e0);
        }
    } else {
        gs.col = (const CT *)kp.col;
        gs.f32 = kp.f32;
        gs.hi = kp.hi;
        gs.lo = kp.lo;
        gb = it.gbase;
    }
    // group loaded by lanes whose row is exhausted: the zero group (in LDS
    // right after the staged groups, else the matrix's in the global arrays)
    // or, masked, a group of the item
    const uint32_t gz = XS ? (ESTAGED ? ng : it.flags >> 2) : gb + (ng ? ng - 1 : 0u);
    auto bnd = [&](uint32_t q) { return row_of(off[q - rbase], off[q - rbase + 1], rowid[q - rbase]); };
    float *Y = kp.Y;
    for (uint32_t tt = it.t0; tt < it.t1; ++tt) {
        const tile_t tl = kp.tiles[tt];
        if (tl.meta & kTileLong) {
            // one long row: every wave strides over it, then the waves merge
            const row_rng r = bnd(tl.r0);
            constexpr int slots = 64 / K;
            const uint32_t E = (uint32_t)(kWaves * slots);
            acc2 a;
            for (uint32_t g = r.g0 + (uint32_t)(lane / K) + (uint32_t)w * slots; g < r.g1; g += E) {
                uint32_t c[kGroup];
                float v[kGroup];
                gs.get(g, c, v);
                group_dot<MASK>(a, r, g, c, v, xv, j);
            }
            for (int o = K; o < 64; o <<= 1) a.merge(__shfl_xor(a.s, o, 64), __shfl_xor(a.c, o, 64));
            if (lane < K) {
                s_part[0][w][j] = a.s;
                s_part[1][w][j] = a.c;
            }
            __syncthreads();
            if (w == 0 && lane < K) {
                acc2 b;
                for (int ww = 0; ww < kWaves; ++ww) b.merge(s_part[0][ww][j], s_part[1][ww][j]);
                Y[(uint64_t)r.id * K + j] = b.s + b.c;
            }
            __syncthreads();
            continue;
        }
        // lanes = (row slot rs, entry slot e, rhs j): E lanes per row, 64 / (E
        // K) rows per wavefront step; the group takes rows p and p + S per step
        // (S = rows per block sweep).  The lanes of a row share its row and so
        // its activity: the merge pairs only them.
        const uint32_t E = 1u << (tl.meta & 0x7fu);
        const uint32_t e = (uint32_t)(lane / K) & (E - 1);
        const uint32_t rs = (uint32_t)lane / (K * E);
        const uint32_t rpw = 64u / (K * E);
        rows_loop<K, MASK>(tl.r0 + (uint32_t)w * rpw + rs, tl.r1, kWaves * rpw, e, E, gz, gs, xv, j, bnd,
                           [&](uint32_t, const row_rng &r, const acc2 &a) { Y[(uint64_t)r.id * K + j] = a.s + a.c; });
    }
}

// k = 1 item whose X block is staged in two column halves.  Pass h stages
// columns [h W, min(n, (h + 1) W)) of the matrix (W = ceil(n / 2)); a row's
// entries of half 0 are [4 goff[p], mid[p]) and those of half 1 [mid[p], its
// end) (columns are sorted; mid is computed when the batch first gets a
// two-pass schedule).  Pass 0 leaves the row's (sum, correction) in global
// memory (rec[2 p ..]); pass 1 merges its own into it and writes the output.
// Both passes give a row to the same lanes of the same wavefront.  Entries
// are read from global memory; the rows carry no whole-block tiles.
template <int ST, typename CT>
__device__ __forceinline__ void run_item_two_pass(const kparams &kp, const item_t &it, unsigned char *lds, int lane,
                                                  int w)
{
    const uint32_t n = it.xcols, W = (n + 1) / 2;
    const gsrc<ST, CT, false> gs{(const CT *)kp.col, kp.f32, kp.hi, kp.lo};
    const uint32_t ng = kp.goff[it.r1] - it.gbase;
    const uint32_t glast = it.gbase + (ng ? ng - 1 : 0u);
    const uint32_t *off = kp.goff, *mid = kp.mid, *rowid = kp.rowid;
    uint32_t *rec = kp.rec;
    float *Y = kp.Y;
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t c0 = h * W, c1 = min(n, c0 + W);
        __syncthreads();  // the previous item or pass is done with the LDS
        const xsrc_half<CT> xh{stage_array<stage_depth<1>()>(lds, 0, kp.X, (uint64_t)it.cbase + c0, c1 - c0),
                               it.cbase, c0};
        stage_wait<stage_depth<1>()>();
        __syncthreads();
        for (uint32_t tt = it.t0; tt < it.t1; ++tt) {
            const tile_t tl = kp.tiles[tt];
            const uint32_t E = 1u << (tl.meta & 0x7fu);
            const uint32_t e = (uint32_t)lane & (E - 1);
            const uint32_t rs = (uint32_t)lane / E;
            const uint32_t rpw = 64u / E;
            const uint32_t p = tl.r0 + (uint32_t)w * rpw + rs, S = kWaves * rpw;
            if (h == 0)
                rows_loop<1, true>(
                    p, tl.r1, S, e, E, glast, gs, xh, 0,
                    [&](uint32_t q) {
                        const uint32_t g0 = off[q], m = mid[q];
                        return row_rng{g0, (m + kGroup - 1) / kGroup, kGroup * g0, m - kGroup * g0, q};
                    },
                    [&](uint32_t, const row_rng &r, const acc2 &a) {
                        rec[2 * (uint64_t)r.id] = __float_as_uint(a.s);
                        rec[2 * (uint64_t)r.id + 1] = __float_as_uint(a.c);
                    });
            else
                rows_loop<1, true>(
                    p, tl.r1, S, e, E, glast, gs, xh, 0,
                    [&](uint32_t q) {
                        const uint32_t g1 = off[q + 1], m = mid[q], rid = rowid[q];
                        return row_rng{m / kGroup, g1, m, kGroup * g1 - (rid >> kIdBits) - m, q};
                    },
                    [&](uint32_t, const row_rng &r, const acc2 &a) {
                        acc2 t;
                        t.s = __uint_as_float(rec[2 * (uint64_t)r.id]);
                        t.c = __uint_as_float(rec[2 * (uint64_t)r.id + 1]);
                        t.merge(a.s, a.c);
                        Y[kp.rowid[r.id] & kIdMask] = t.s + t.c;
                    });
        }
    }
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
    // queues tried and found empty so far (used by thread 0): the block's own
    // queue first, then the next ones in turn; an empty queue stays empty
    uint32_t tried = 0;
    const uint32_t home = blockIdx.x % kQueues;
    for (;;) {
        // The barrier after the atomic also orders the previous item's LDS
        // reads before this item's staging writes.
        if (threadIdx.x == 0) {
            uint32_t t = kp.n_items;
            for (; tried < kQueues; ++tried) {
                const uint32_t q = (home + tried) % kQueues;
                const uint32_t i = kp.qbeg[q] + atomicAdd(kp.counter + q * kCounterStride, 1u);
                if (i < kp.qbeg[q + 1]) {
                    t = i;
                    break;
                }
            }
            s_t = t;
        }
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
        const uint32_t xbytes = xs ? xblock_bytes(it.xcols, K) : 0u;
        float *xblk = nullptr;
        if (xs) xblk = stage_array<stage_depth<K>()>(lds, 0, kp.X, (uint64_t)it.cbase * K, it.xcols * K);
        if (staged) {
            const uint32_t rows = it.r1 - it.r0;
            const uint32_t ng = kp.goff[it.r1] - it.gbase;
            const lds_layout<ST, CT> L(xbytes, rows, ng * kGroup);
            uint32_t *s_off = (uint32_t *)(lds + L.off_o);
            uint32_t *s_rid = (uint32_t *)(lds + L.rid_o);
            const uint64_t e0 = (uint64_t)it.gbase * kGroup;
            // the entry copies first: the row arrays go through registers, and
            // waiting for their loads waits for every copy issued before them
            stage_array<stage_depth<K>()>(lds, L.col_o, (const CT *)kp.col, e0, ng * kGroup);
            if constexpr (ST == SPIRA_VAL_F32) {
                stage_array<stage_depth<K>()>(lds, L.v1_o, kp.f32, e0, ng * kGroup);
            } else {
                stage_array<stage_depth<K>()>(lds, L.v1_o, kp.hi, e0, ng * kGroup);
                if constexpr (ST == SPIRA_VAL_SPLIT_BF16)
                    stage_array<stage_depth<K>()>(lds, L.v2_o, kp.lo, e0, ng * kGroup);
            }
            for (uint32_t q = threadIdx.x; q <= rows; q += kThreads) s_off[q] = kp.goff[it.r0 + q] - it.gbase;
            for (uint32_t q = threadIdx.x; q < rows; q += kThreads) s_rid[q] = kp.rowid[it.r0 + q];
        }
        if (xs || staged) {
            stage_wait<stage_depth<K>()>();
            __syncthreads();
        }
        if (xs) {
            // the zero row at column n_cols, read by padding entries, and for
            // staged entries the zero group after them (written once every copy
            // has landed: the copied spans may cover them)
            if (threadIdx.x < K) xblk[it.xcols * K + threadIdx.x] = 0.0f;
            if (staged && threadIdx.x < kGroup) {
                const uint32_t ng = kp.goff[it.r1] - it.gbase;
                const lds_layout<ST, CT> L(xbytes, it.r1 - it.r0, ng * kGroup);
                const uint64_t e0 = (uint64_t)it.gbase * kGroup, z = (uint64_t)ng * kGroup + threadIdx.x;
                ((CT *)staged_at(lds, L.col_o, (const CT *)kp.col, e0))[z] =
                    (CT)((sizeof(CT) == 2) ? it.xcols : it.cbase + it.xcols);
                if constexpr (ST == SPIRA_VAL_F32) {
                    ((float *)staged_at(lds, L.v1_o, kp.f32, e0))[z] = 0.0f;
                } else {
                    ((uint16_t *)staged_at(lds, L.v1_o, kp.hi, e0))[z] = 0;
                    if constexpr (ST == SPIRA_VAL_SPLIT_BF16) ((uint16_t *)staged_at(lds, L.v2_o, kp.lo, e0))[z] = 0;
                }
            }
            __syncthreads();
            if (staged) run_item<K, ST, CT, true, true>(kp, it, lds, s_part, lane, w, j);
            else run_item<K, ST, CT, true, false>(kp, it, lds, s_part, lane, w, j);
        } else {
            if (staged) run_item<K, ST, CT, false, true>(kp, it, lds, s_part, lane, w, j);
            else run_item<K, ST, CT, false, false>(kp, it, lds, s_part, lane, w, j);
        }
    }
}

// Copy matrix m's rows into the batch in stored order: stored row q (global)
// takes source row perm[q] - row_base, its entries at 4 dst_goff[q], then
// padding entries (column padc, value 0) up to the row's next group; 16 lanes
// per row.
template <typename CT>
__global__ void k_gather_rows(const uint64_t *__restrict__ src_off, const uint32_t *__restrict__ src_col,
                              const double *__restrict__ src_val, uint32_t row_base, uint32_t n,
                              uint32_t col_base, CT padc, const uint32_t *__restrict__ perm,
                              const uint32_t *__restrict__ dst_goff, int storage, CT *__restrict__ ocol,
                              float *__restrict__ of32, uint16_t *__restrict__ ohi, uint16_t *__restrict__ olo)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = (uint32_t)(g / 16);
    const uint32_t l = (uint32_t)(g % 16);
    if (q >= n) return;
    const uint32_t r = (perm[row_base + q] & kIdMask) - row_base;
    const uint64_t s0 = src_off[r], s1 = src_off[r + 1];
    const uint32_t len = (uint32_t)(s1 - s0);
    const uint32_t d0 = kGroup * dst_goff[row_base + q];
    const uint32_t plen = (len + kGroup - 1) & ~(kGroup - 1);
    for (uint32_t i = l; i < plen; i += 16) {
        const uint32_t d = d0 + i;
        if (i >= len) {
            ocol[d] = padc;
            if (storage == SPIRA_VAL_F32) {
                of32[d] = 0.0f;
            } else {
                ohi[d] = 0;
                if (storage == SPIRA_VAL_SPLIT_BF16) olo[d] = 0;
            }
            continue;
        }
        ocol[d] = (CT)(sizeof(CT) == 2 ? src_col[s0 + i] : src_col[s0 + i] + col_base);
        const double sv = src_val[s0 + i];
        const float f = (float)sv;
        if (storage == SPIRA_VAL_F32) {
            of32[d] = f;
        } else {
            const uint32_t u = __float_as_uint(f);
            const uint32_t ru = u + 0x7fffu + ((u >> 16) & 1u);
            const uint16_t h = (u & 0x7fffffffu) > 0x7f800000u ? (uint16_t)((u >> 16) | 0x40u) : (uint16_t)(ru >> 16);
            ohi[d] = h;
            if (storage == SPIRA_VAL_SPLIT_BF16) {
                const float rest = (float)(sv - (double)__uint_as_float(((uint32_t)h) << 16));
                const uint32_t v = __float_as_uint(rest);
                const uint32_t rv = v + 0x7fffu + ((v >> 16) & 1u);
                olo[d] = (v & 0x7fffffffu) > 0x7f800000u ? (uint16_t)((v >> 16) | 0x40u) : (uint16_t)(rv >> 16);
            }
        }
    }
}

// mid[q] for every stored row q of a matrix with wsplit[m] > 0: the first
// entry of the row whose column inside its matrix is at least wsplit[m]
// (columns are sorted within a row; the search covers the real entries, not
// the padding).  Rows of other matrices are left as they are.  row_base /
// col_base: first stored row / batch column of each matrix (n_mats + 1
// entries).
template <typename CT>
__global__ void k_split_mid(const uint32_t *__restrict__ goff, const uint32_t *__restrict__ rowid,
                            const CT *__restrict__ col, const uint32_t *__restrict__ row_base,
                            const uint32_t *__restrict__ col_base, const uint32_t *__restrict__ wsplit,
                            uint32_t n_mats, uint32_t n_rows, uint32_t *__restrict__ mid)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_rows) return;
    uint32_t lo = 0, hi = n_mats;  // row_base[lo] <= q < row_base[hi]
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) / 2;
        if (row_base[m] <= q) lo = m;
        else hi = m;
    }
    const uint32_t W = wsplit[lo];
    if (!W) return;
    const uint32_t cb = sizeof(CT) == 2 ? 0u : col_base[lo];
    uint32_t a = kGroup * goff[q], b = kGroup * goff[q + 1] - (rowid[q] >> kIdBits);
    while (a < b) {
        const uint32_t m = (a + b) / 2;
        if ((uint32_t)col[m] - cb < W) a = m + 1;
        else b = m;
    }
    mid[q] = a;
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
    uint64_t ents = 0;  // stored entries, padding included
    std::vector<uint64_t> row_base, col_base, nnz_base;  // n_mats+1 each
    std::vector<uint32_t> h_off;                          // stored-order padded entry offsets (host, tile building)
    std::vector<uint32_t> class_start;                    // n_mats * (kClasses+1): first stored row of each class
    uint32_t *goff = nullptr;   // stored-order group offsets (h_off / 4)
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
    uint32_t *counter = nullptr;     // kQueues queue counters
    uint32_t *qbeg = nullptr;        // kQueues + 1 queue bounds
    uint32_t *mid = nullptr;         // two-pass split of every stored row of a wide matrix (first need)
    uint32_t *rec = nullptr;         // two-pass records, 2 words per stored row (first need)
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

// Two-pass support, built when a schedule first gives some matrix two-pass
// items: every stored row of every matrix too wide for a one-load X block at
// k = 1 gets its split point between the column halves, and every stored row a
// (sum, correction) record slot.
static void batch_split_rows(spira_batch *b)
{
    const uint32_t nm = b->n_mats;
    const uint64_t nr = std::max<uint64_t>(b->n_rows, 1);
    std::vector<uint32_t> h(3 * (size_t)(nm + 1), 0u);  // row_base, col_base, wsplit
    for (uint32_t m = 0; m <= nm; ++m) {
        h[m] = (uint32_t)b->row_base[m];
        h[nm + 1 + m] = (uint32_t)b->col_base[m];
        if (m < nm) {
            const uint64_t nc = b->col_base[m + 1] - b->col_base[m];
            h[2 * (nm + 1) + m] = nc > SPIRA_BATCH_XLDS_FLOATS ? (uint32_t)((nc + 1) / 2) : 0u;
        }
    }
    uint32_t *mid = (uint32_t *)b->alloc(nr * sizeof(uint32_t));
    uint32_t *rec = (uint32_t *)b->alloc(nr * 2 * sizeof(uint32_t));
    const size_t tb = h.size() * sizeof(uint32_t);
    uint32_t *tab = (uint32_t *)dev_alloc(b->dev, tb);
    hipStream_t st = b->dev->stream;
    try {
        SPIRA_HIP(hipMemcpyAsync(tab, h.data(), tb, hipMemcpyHostToDevice, st));
        if (b->n_rows && nm) {
            const unsigned grid = grid_for(b->n_rows, 256);
            if (b->col16)
                k_split_mid<uint16_t><<<grid, 256, 0, st>>>(b->goff, b->rowid, (const uint16_t *)b->col, tab, tab + nm + 1,
                                                           tab + 2 * (nm + 1), nm, (uint32_t)b->n_rows, mid);
            else
                k_split_mid<uint32_t><<<grid, 256, 0, st>>>(b->goff, b->rowid, (const uint32_t *)b->col, tab, tab + nm + 1,
                                                           tab + 2 * (nm + 1), nm, (uint32_t)b->n_rows, mid);
            SPIRA_HIP(hipGetLastError());
        }
        dev_sync(b->dev);
    } catch (...) {
        (void)hipStreamSynchronize(st);
        dev_release(b->dev, tab, tb);
        throw;
    }
    dev_release(b->dev, tab, tb);
    b->mid = mid;
    b->rec = rec;
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
        const uint32_t xbytes = xblock_bytes(xcols, K);
        const bool stage = !two_pass && !(s.flags & SPIRA_BATCH_STREAM) && xbytes + kMinEntryLds <= kLdsBytes;
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
                cur.gbase = off[ir0] / kGroup;
                cur.xcols = item_tp ? ncols : xcols;
                cur.flags = (item_tp ? kItemTwoPass : ((may_stage && fits(r_end)) ? kItemStaged : 0u)) |
                            (uint32_t)((b->ents / kGroup + m) << 2);
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
                // when staged, within the LDS
                uint32_t q = p;
                while (q < z && (!lng || q == p) && cur_cost + (off[q + 1] - off[p]) + (q + 1 - p) <= budget &&
                       (!stage || fits(q + 1)))
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
                if (cur_cost >= budget || (stage && p < z && !fits(p + 1))) close_item(p, stage);
            }
        }
        if (open) close_item(cs[kClasses], stage);
    }
    // Queues: matrices dealt whole, heaviest first, each to the queue with the
    // least work so far (longest-processing-time order); a queue runs its
    // matrices in the order dealt, each matrix's items in stored-row order.
    std::vector<uint64_t> mwork(b->n_mats, 0);
    std::vector<uint32_t> mfirst(b->n_mats + 1, 0);  // items of matrix m: [mfirst[m], mfirst[m + 1])
    bool any_two_pass = false;
    for (const pending &x : its) {
        mwork[x.mat] += x.work;
        ++mfirst[x.mat + 1];
        any_two_pass |= (x.it.flags & kItemTwoPass) != 0;
    }
    for (uint32_t m = 0; m < b->n_mats; ++m) mfirst[m + 1] += mfirst[m];
    std::vector<uint32_t> morder(b->n_mats);
    std::iota(morder.begin(), morder.end(), 0u);
    std::stable_sort(morder.begin(), morder.end(), [&](uint32_t x, uint32_t y) { return mwork[x] > mwork[y]; });
    std::vector<std::vector<uint32_t>> qmats(kQueues);
    std::vector<uint64_t> qwork(kQueues, 0);
    for (uint32_t m : morder) {
        if (mfirst[m + 1] == mfirst[m]) continue;
        const uint32_t q = (uint32_t)(std::min_element(qwork.begin(), qwork.end()) - qwork.begin());
        qmats[q].push_back(m);
        qwork[q] += mwork[m];
    }
    std::vector<item_t> items;
    items.reserve(its.size());
    b->item_mat.clear();
    b->item_mat.reserve(its.size());
    uint32_t hq[kQueues + 1];
    for (uint32_t q = 0; q < kQueues; ++q) {
        hq[q] = (uint32_t)items.size();
        for (uint32_t m : qmats[q])
            for (uint32_t i = mfirst[m]; i < mfirst[m + 1]; ++i) {
                items.push_back(its[i].it);
                b->item_mat.push_back(m);
            }
    }
    hq[kQueues] = (uint32_t)items.size();
    hipStream_t st = b->dev->stream;
    if (any_two_pass && !b->mid) batch_split_rows(b);
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
    SPIRA_HIP(hipMemcpyAsync(b->qbeg, hq, sizeof(hq), hipMemcpyHostToDevice, st));
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
    kparams kp{b->items, b->n_items, b->counter, b->qbeg, b->mid, b->rec, b->tiles, b->goff, b->rowid, b->col, b->f32,
               b->hi, b->lo, X, Y, prof};
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
    k_reset_counters<<<1, kWave, 0, b->dev->stream>>>(b->counter);
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
        if (b->nnz >= (1ull << 32) || b->n_rows >= (1ull << kIdBits) || b->n_cols >= (1ull << 32))
            fail(SPIRA_ERR_UNSUPPORTED, "batch: totals must stay below 2^30 rows and 2^32 columns and nonzeros");
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
        // rows padded to whole groups; the padding count rides in rowid's top bits
        uint64_t ents = 0;
        std::vector<uint32_t> goff(nr + 1, 0);
        for (uint64_t q = 0; q < nr; ++q) {
            const uint64_t len = src[perm[q] + 1] - src[perm[q]];
            const uint64_t plen = (len + kGroup - 1) / kGroup * kGroup;
            ents += plen;
            if (ents >= (1ull << 32)) fail(SPIRA_ERR_UNSUPPORTED, "batch: padded nonzeros must stay below 2^32");
            b->h_off[q + 1] = (uint32_t)ents;
            goff[q + 1] = (uint32_t)(ents / kGroup);
            perm[q] |= (uint32_t)(plen - len) << kIdBits;
        }
        src.clear();
        src.shrink_to_fit();
        b->ents = ents;
        if (ents / kGroup + n_mats >= (1ull << kIdBits))
            fail(SPIRA_ERR_UNSUPPORTED, "batch: too many nonzeros and matrices for the group numbering");
        // device arrays: the rows' groups, then one zero group per matrix
        // (value 0, the matrix's padding column), then a zeroed tail (items
        // copy the arrays in 16-byte spans)
        const uint64_t nz = ents + (uint64_t)kGroup * n_mats + kGroup;
        b->goff = (uint32_t *)b->alloc((nr + 1) * sizeof(uint32_t));
        b->rowid = (uint32_t *)b->alloc((nr ? nr : 1) * sizeof(uint32_t));
        // every array is zeroed (so the spans read past the last group hold column 0, value 0)
        auto zeroed = [&](size_t bytes) {
            void *p = b->alloc(bytes);
            SPIRA_HIP(hipMemsetAsync(p, 0, bytes, st));
            return p;
        };
        b->col = zeroed(nz * (b->col16 ? sizeof(uint16_t) : sizeof(uint32_t)) + kArrayPad);
        if (storage == SPIRA_VAL_F32) b->f32 = (float *)zeroed(nz * sizeof(float) + kArrayPad);
        else b->hi = (uint16_t *)zeroed(nz * sizeof(uint16_t) + kArrayPad);
        if (storage == SPIRA_VAL_SPLIT_BF16) b->lo = (uint16_t *)zeroed(nz * sizeof(uint16_t) + kArrayPad);
        SPIRA_HIP(hipMemcpyAsync(b->goff, goff.data(), (nr + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        {
            std::vector<uint32_t> zc32;
            std::vector<uint16_t> zc16;
            for (uint32_t m = 0; m < n_mats; ++m) {
                const uint64_t nc = b->col_base[m + 1] - b->col_base[m];
                for (uint32_t u = 0; u < kGroup; ++u) {
                    if (b->col16) zc16.push_back((uint16_t)(nc <= 0xffffu ? nc : 0u));
                    else zc32.push_back((uint32_t)(b->col_base[m] + nc));
                }
            }
            if (n_mats) {
                if (b->col16)
                    SPIRA_HIP(hipMemcpyAsync((uint16_t *)b->col + ents, zc16.data(), zc16.size() * 2,
                                             hipMemcpyHostToDevice, st));
                else
                    SPIRA_HIP(hipMemcpyAsync((uint32_t *)b->col + ents, zc32.data(), zc32.size() * 4,
                                             hipMemcpyHostToDevice, st));
            }
            dev_sync(dev);  // zc16 / zc32 are read until here
        }
        if (nr) SPIRA_HIP(hipMemcpyAsync(b->rowid, perm.data(), nr * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        for (uint32_t m = 0; m < n_mats; ++m) {
            const spira_dcsr *A = mats[m];
            if (!A->n_rows || !A->nnz) continue;
            const unsigned grid = grid_for((uint64_t)A->n_rows * 16, 256);
            // padding column: the zero row after the matrix's X block (column
            // 0 for a matrix of 65536 columns, whose X block never sits in LDS
            // in one piece: its padding is masked)
            if (b->col16)
                k_gather_rows<uint16_t><<<grid, 256, 0, st>>>(
                    A->off, A->col, A->val, (uint32_t)b->row_base[m], (uint32_t)A->n_rows, (uint32_t)b->col_base[m],
                    (uint16_t)(A->n_cols <= 0xffffu ? A->n_cols : 0u), b->rowid, b->goff, storage, (uint16_t *)b->col,
                    b->f32, b->hi, b->lo);
            else
                k_gather_rows<uint32_t><<<grid, 256, 0, st>>>(
                    A->off, A->col, A->val, (uint32_t)b->row_base[m], (uint32_t)A->n_rows, (uint32_t)b->col_base[m],
                    (uint32_t)(b->col_base[m] + A->n_cols), b->rowid, b->goff, storage, (uint32_t *)b->col, b->f32,
                    b->hi, b->lo);
        }
        SPIRA_HIP(hipGetLastError());
        b->counter = (uint32_t *)b->alloc(kQueues * kCounterStride * sizeof(uint32_t));
        b->qbeg = (uint32_t *)b->alloc((kQueues + 1) * sizeof(uint32_t));
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
            if (s[m].flags & ~(uint16_t)SPIRA_BATCH_STREAM)
                fail(SPIRA_ERR_INVALID_ARGUMENT, "batch: unknown schedule flags");
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
