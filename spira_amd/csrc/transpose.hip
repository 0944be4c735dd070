// transpose.hip -- A^T of a locked device CSR, bit-identical to the
// reference's serial transpose (serial/transpose.hpp:35-157: column histogram,
// offsets, then a fill in ascending input-row order, so every output row lists
// its columns ascending; values copied as stored, explicit zeros included) or,
// with SPIRA_TRANSPOSE_DROP_ZEROS, to the parallel one
// (parallel/algorithms/transpose.hpp:37-69: the same entries inserted into a
// new matrix whose lock() removes explicit zeros).
//
//   1. k_key_rows: one thread per row marks the row owning the first entry of
//      every key block (a row table), so no block searches for its rows.
//   2. k_transpose_keys: per entry a 64-bit key, (column << 32) | input row,
//      with the sentinel column n_cols for a dropped zero.
//   3. A stable LSD radix sort (rocPRIM onesweep, 8-bit digits; wider digits
//      were measured slower per pass than the passes they save) of
//      (key, value) over the column bits only:
//      entries come out grouped by output row and, inside a row, in
//      input-entry order, which is input-row order -- the reference's fill
//      order.  The values travel with the keys, so nothing is gathered later.
//   4. k_transpose_out: one streaming pass over the sorted keys writes the
//      output columns (the low halves) and, where the column changes, the
//      output offsets of every row in between.
#include "internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

namespace spira_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;         // entries per thread in the key pass
constexpr int kRowCache = 2048;   // row offsets a key block keeps in LDS
constexpr uint64_t kKeyBlock = (uint64_t)kThreads * kItems;  // entries per key block


// tab[b] = row owning entry b * kKeyBlock (b = 0 .. ceil(nnz / kKeyBlock));
// rows are non-decreasing in the entry index, so a block's rows are
// [tab[b], tab[b + 1]].
__global__ void k_key_rows(uint64_t n_rows, uint64_t nnz, const uint64_t *__restrict__ off, uint32_t *__restrict__ tab)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    const uint64_t e0 = off[r], e1 = off[r + 1];
    if (e0 == e1) return;
    for (uint64_t b = (e0 + kKeyBlock - 1) / kKeyBlock; b * kKeyBlock < e1; ++b) tab[b] = (uint32_t)r;
    if (e1 == nnz) tab[(nnz + kKeyBlock - 1) / kKeyBlock] = (uint32_t)r;  // the row owning the last entry
}

__global__ __launch_bounds__(kThreads) void k_transpose_keys(uint64_t n_rows, uint64_t nnz, uint64_t n_cols,
                                                              const uint64_t *__restrict__ off,
                                                              const uint32_t *__restrict__ col,
                                                              const double *__restrict__ val, int drop_zeros,
                                                              const uint32_t *__restrict__ tab,
                                                              uint64_t *__restrict__ key)
{
    __shared__ uint64_t s_off[kRowCache + 1];
    const uint64_t first = (uint64_t)blockIdx.x * kKeyBlock;
    if (first >= nnz) return;
    const uint64_t lo = tab[blockIdx.x], hi = tab[blockIdx.x + 1];
    const uint64_t nr = hi - lo + 1;  // rows the block's entries lie in
    if (nr <= kRowCache) {
        for (uint64_t i = threadIdx.x; i <= nr; i += kThreads) s_off[i] = off[lo + i];  // off[lo .. hi + 1]
        __syncthreads();
        uint32_t r = 0;  // a thread's entries ascend, so each search starts at the previous row
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const uint64_t k = first + (uint64_t)it * kThreads + threadIdx.x;
            if (k >= nnz) break;
            uint32_t b = (uint32_t)nr;
            while (b - r > 1) {
                const uint32_t mid = r + (b - r) / 2;
                if (s_off[mid] <= k) r = mid;
                else b = mid;
            }
            const uint64_t c = (drop_zeros && is_zero_f64(val[k])) ? n_cols : (uint64_t)col[k];
            key[k] = (c << 32) | (lo + r);
        }
    } else {  // many empty rows inside the block's span: search the offsets in memory
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const uint64_t k = first + (uint64_t)it * kThreads + threadIdx.x;
            if (k >= nnz) break;
            const uint64_t r = row_of(off, lo, hi + 1, k);
            const uint64_t c = (drop_zeros && is_zero_f64(val[k])) ? n_cols : (uint64_t)col[k];
            key[k] = (c << 32) | r;
        }
    }
}

// Sorted keys -> output columns (the low halves) and offsets: t_off[c] =
// number of keys whose column is below c, c = 0 .. n_cols (dropped zeros carry
// the column n_cols and sort last, so t_off[n_cols] is the output nnz).  The
// entry where the column steps up from c0 to c1 writes t_off[c0 + 1 .. c1];
// the first entry writes t_off[0 .. c1], the last t_off[c1 + 1 .. n_cols].
__global__ void k_transpose_out(const uint64_t *__restrict__ skey, uint64_t nnz, uint64_t n_cols,
                                uint64_t *__restrict__ t_off, uint32_t *__restrict__ t_col)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nnz == 0) {
        for (uint64_t c = t0; c <= n_cols; c += stride) t_off[c] = 0;
        return;
    }
    for (uint64_t p = t0; p < nnz; p += stride) {
        const uint64_t k = skey[p];
        const uint64_t c = k >> 32;
        t_col[p] = (uint32_t)k;
        const uint64_t from = p == 0 ? 0 : (skey[p - 1] >> 32) + 1;
        for (uint64_t q = from; q <= c; ++q) t_off[q] = p;
        if (p == nnz - 1)
            for (uint64_t q = c + 1; q <= n_cols; ++q) t_off[q] = nnz;
    }
}

unsigned key_bits(uint64_t max_key)
{
    unsigned b = 1;
    while (b < 32 && (1ull << b) <= max_key) ++b;
    return b;
}

}  // namespace

}  // namespace spira_amd

using namespace spira_amd;

extern "C" int spira_transpose_f64(spira_dcsr *A, spira_dcsr *out, uint32_t flags)
{
    if (!A || !out) return SPIRA_ERR_INVALID_ARGUMENT;
    spira_dev *dev = A->dev;
    try {
        SPIRA_HIP(hipSetDevice(dev->device));
        if (out->dev != dev) fail(SPIRA_ERR_INVALID_ARGUMENT, "transpose: A and out belong to different devices");
        if (!A->built) fail(SPIRA_ERR_LOGIC, "transpose: matrix must be locked");
        if (out->n_rows != A->n_cols || out->n_cols != A->n_rows)
            fail(SPIRA_ERR_INVALID_ARGUMENT, "transpose: out must have A's shape with rows and columns swapped");
        if (flags & ~(uint32_t)SPIRA_TRANSPOSE_DROP_ZEROS) fail(SPIRA_ERR_INVALID_ARGUMENT, "transpose: unknown flags");
        if (A->nnz >= (1ull << 32) || A->n_rows >= (1ull << 32) || A->n_cols >= (1ull << 32) - 1)
            fail(SPIRA_ERR_UNSUPPORTED, "transpose: needs nnz, rows and columns below 2^32");
        const bool drop = (flags & SPIRA_TRANSPOSE_DROP_ZEROS) != 0;
        hipStream_t st = dev->stream;
        const uint64_t n = A->n_rows, nc = A->n_cols, nnz = A->nnz;
        const size_t ob = (nc + 1) * sizeof(uint64_t), cb = (nnz ? nnz : 1) * sizeof(uint32_t),
                     vb = (nnz ? nnz : 1) * sizeof(double), kb = (nnz ? nnz : 1) * sizeof(uint64_t);
        uint64_t *t_off = (uint64_t *)dev_alloc(dev, ob);
        uint32_t *t_col = nullptr;
        uint64_t *key = nullptr, *skey = nullptr;
        double *t_val = nullptr;
        auto release_tmp = [&] {
            if (key) dev_release(dev, key, kb);
            if (skey) dev_release(dev, skey, kb);
            key = skey = nullptr;
        };
        try {
            t_col = (uint32_t *)dev_alloc(dev, cb);
            t_val = (double *)dev_alloc(dev, vb);
            key = (uint64_t *)dev_alloc(dev, kb);
            skey = (uint64_t *)dev_alloc(dev, kb);
            if (nnz > 0) {
                const uint64_t nkb = (nnz + kKeyBlock - 1) / kKeyBlock;
                uint32_t *tab = (uint32_t *)dev->ws.get(WS_MISC2, (nkb + 1) * sizeof(uint32_t));
                k_key_rows<<<grid_for(n, kThreads), kThreads, 0, st>>>(n, nnz, A->off, tab);
                SPIRA_HIP(hipGetLastError());
                k_transpose_keys<<<(unsigned)nkb, kThreads, 0, st>>>(n, nnz, nc, A->off, A->col, A->val, drop ? 1 : 0,
                                                                     tab, key);
                SPIRA_HIP(hipGetLastError());
                // columns 0 .. n_cols - 1, and the sentinel n_cols when zeros
                // are dropped, in bits 32 and up
                const unsigned bits = key_bits(drop ? nc : nc - 1);
                size_t tmp_bytes = 0;
                SPIRA_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, key, skey, A->val, t_val,
                                                                 (unsigned)nnz, 32, 32 + bits, st));
                void *tmp = dev->ws.get(WS_SORT, tmp_bytes ? tmp_bytes : 8);
                SPIRA_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, key, skey, A->val, t_val,
                                                                 (unsigned)nnz, 32, 32 + bits, st));
            }
            k_transpose_out<<<grid_for(std::max<uint64_t>(nnz, 1), kThreads, 16384), kThreads, 0, st>>>(
                skey, nnz, nc, t_off, t_col);
            SPIRA_HIP(hipGetLastError());
            const uint64_t nnz_out = read_scalar(dev, t_off + nc);
            release_tmp();
            // install (out may be A itself for a square matrix: everything above
            // has been read already, and stream order protects the old arrays)
            const bool zeros = drop ? false : A->has_zeros;
            free_dcsr_arrays(out);
            out->off = t_off;
            out->col = t_col;
            out->val = t_val;
            out->nnz = nnz_out;
            out->capacity = nnz;
            out->built = true;
            out->has_zeros = zeros;
        } catch (...) {
            (void)hipStreamSynchronize(st);
            release_tmp();
            dev_release(dev, t_off, ob);
            if (t_col) dev_release(dev, t_col, cb);
            if (t_val) dev_release(dev, t_val, vb);
            throw;
        }
        return SPIRA_OK;
    } catch (const status_error &e) {
        dev->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        dev->last_error = "host allocation failed";
        return SPIRA_ERR_NOMEM;
    } catch (const std::exception &e) {
        dev->last_error = e.what();
        return SPIRA_ERR_DEVICE;
    }
}
