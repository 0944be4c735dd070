/*
 * spira_oracle.c -- CPU restatement of Spira's locked-mode hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see spira_oracle.h).  Compiled with
 * -ffp-contract=off so that `acc += x*v` is a rounded multiply followed by a
 * rounded add, exactly the C++ expression in serial/spmv.hpp:52.
 */
#define _GNU_SOURCE
#include "spira_oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

int spira_or_is_zero(double v) { return fabs(v) <= 0.0; }
static inline int is_zero_f(double v) { return fabs(v) <= 0.0; }

/* ------------------------------------------------------------------------ */
/* a2: sort + dedup of one row buffer (aos_array_buffer.hpp:76-101)          */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t col;
    double val;
} entry_t;

/* Stable merge sort by column (the std::stable_sort of aos_array_buffer.hpp:82). */
static void stable_sort_by_col(entry_t *a, entry_t *tmp, uint64_t n)
{
    if (n <= 24) {
        for (uint64_t i = 1; i < n; ++i) {
            entry_t e = a[i];
            uint64_t j = i;
            while (j > 0 && a[j - 1].col > e.col) {
                a[j] = a[j - 1];
                --j;
            }
            a[j] = e;
        }
        return;
    }
    uint64_t h = n / 2;
    stable_sort_by_col(a, tmp, h);
    stable_sort_by_col(a + h, tmp, n - h);
    uint64_t i = 0, j = h, k = 0;
    while (i < h && j < n) {
        if (a[j].col < a[i].col) tmp[k++] = a[j++];   /* ties take the left run */
        else tmp[k++] = a[i++];
    }
    while (i < h) tmp[k++] = a[i++];
    while (j < n) tmp[k++] = a[j++];
    memcpy(a, tmp, n * sizeof(entry_t));
}

/* Dedup one segment of k staged entries; returns the kept count.  scratch must
 * hold 2k entries. */
static uint64_t dedup_segment(const uint32_t *scol, const double *sval,
                              uint64_t k, entry_t *scratch, uint32_t *ocol,
                              double *oval)
{
    if (k == 0) return 0;
    entry_t *buf = scratch, *tmp = scratch + k;
    /* std::reverse (aos_array_buffer.hpp:81): last write first */
    for (uint64_t i = 0; i < k; ++i) {
        buf[i].col = scol[k - 1 - i];
        buf[i].val = sval[k - 1 - i];
    }
    stable_sort_by_col(buf, tmp, k);
    /* keep the first of each column run (:85-96); zeros are kept */
    uint64_t w = 0;
    for (uint64_t i = 0; i < k; ++i) {
        if (i > 0 && buf[i].col == buf[i - 1].col) continue;
        ocol[w] = buf[i].col;
        oval[w] = buf[i].val;
        ++w;
    }
    return w;
}

static uint64_t max_segment(uint64_t r0, uint64_t r1, const uint64_t *ptr)
{
    uint64_t m = 0;
    for (uint64_t r = r0; r < r1; ++r)
        if (ptr[r + 1] - ptr[r] > m) m = ptr[r + 1] - ptr[r];
    return m;
}

uint64_t spira_or_dedup(uint64_t n_rows, const uint64_t *stage_ptr,
                        const uint32_t *stage_col, const double *stage_val,
                        uint64_t *dptr, uint32_t *dcol, double *dval)
{
    uint64_t ms = max_segment(0, n_rows, stage_ptr);
    entry_t *scratch = (entry_t *)malloc((2 * ms + 1) * sizeof(entry_t));
    uint64_t w = 0;
    dptr[0] = 0;
    for (uint64_t r = 0; r < n_rows; ++r) {
        uint64_t b = stage_ptr[r], k = stage_ptr[r + 1] - b;
        w += dedup_segment(stage_col + b, stage_val + b, k, scratch, dcol + w,
                           dval + w);
        dptr[r + 1] = w;
    }
    free(scratch);
    return w;
}

/* ------------------------------------------------------------------------ */
/* a3: build_csr (csr_build.hpp:32-98)                                       */
/* ------------------------------------------------------------------------ */

uint64_t spira_or_build_csr(uint64_t n_rows, const uint64_t *dptr,
                            const uint32_t *dcol, const double *dval,
                            uint64_t *off, uint32_t *col, double *val)
{
    /* pass 1: count non-zeros and scan (:45-59) */
    off[0] = 0;
    for (uint64_t r = 0; r < n_rows; ++r) {
        uint64_t c = 0;
        for (uint64_t k = dptr[r]; k < dptr[r + 1]; ++k)
            if (!is_zero_f(dval[k])) ++c;
        off[r + 1] = off[r] + c;
    }
    /* pass 2: copy non-zeros (:62-79) */
    for (uint64_t r = 0; r < n_rows; ++r) {
        uint64_t w = off[r];
        for (uint64_t k = dptr[r]; k < dptr[r + 1]; ++k)
            if (!is_zero_f(dval[k])) {
                col[w] = dcol[k];
                val[w] = dval[k];
                ++w;
            }
    }
    return off[n_rows];
}

/* ------------------------------------------------------------------------ */
/* a4: merge_csr (csr_build.hpp:126-488)                                     */
/* ------------------------------------------------------------------------ */

/* Merge one row; returns the count written at (col, val). */
static uint64_t merge_row(const uint32_t *oc, const double *ov, uint64_t on,
                          const uint32_t *bc, const double *bv, uint64_t bn,
                          int dirty, uint32_t *col, double *val)
{
    uint64_t w = 0;
    if (!dirty) { /* clean row: verbatim copy, zeros included (:362-377) */
        memcpy(col, oc, on * sizeof(uint32_t));
        memcpy(val, ov, on * sizeof(double));
        return on;
    }
    uint64_t i = 0, j = 0;
    /* two-pointer merge, buffer wins, zeros dropped (:414-449) */
    while (i < on && j < bn) {
        if (oc[i] < bc[j]) {
            if (!is_zero_f(ov[i])) { col[w] = oc[i]; val[w] = ov[i]; ++w; }
            ++i;
        } else if (bc[j] < oc[i]) {
            if (!is_zero_f(bv[j])) { col[w] = bc[j]; val[w] = bv[j]; ++w; }
            ++j;
        } else {
            if (!is_zero_f(bv[j])) { col[w] = bc[j]; val[w] = bv[j]; ++w; }
            ++i;
            ++j;
        }
    }
    for (; i < on; ++i)
        if (!is_zero_f(ov[i])) { col[w] = oc[i]; val[w] = ov[i]; ++w; }
    for (; j < bn; ++j)
        if (!is_zero_f(bv[j])) { col[w] = bc[j]; val[w] = bv[j]; ++w; }
    return w;
}

uint64_t spira_or_merge_csr(uint64_t n_rows, const uint64_t *old_off,
                            const uint32_t *old_col, const double *old_val,
                            const uint64_t *dptr, const uint32_t *dcol,
                            const double *dval, const uint8_t *dirty,
                            uint64_t *off, uint32_t *col, double *val)
{
    off[0] = 0;
    for (uint64_t r = 0; r < n_rows; ++r) {
        uint64_t ob = old_off[r], on = old_off[r + 1] - ob;
        uint64_t bb = dptr[r], bn = dptr[r + 1] - bb;
        int d = dirty ? (dirty[r] != 0) : (bn != 0);
        uint64_t w = merge_row(old_col + ob, old_val + ob, on, dcol + bb,
                               dval + bb, bn, d, col + off[r], val + off[r]);
        off[r + 1] = off[r] + w;
    }
    return off[n_rows];
}

/* ------------------------------------------------------------------------ */
/* a6: SpMV                                                                  */
/* ------------------------------------------------------------------------ */

void spira_or_spmv_f64(uint64_t n_rows, const uint64_t *off,
                       const uint32_t *col, const double *val,
                       const double *x, double *y)
{
    for (uint64_t i = 0; i < n_rows; ++i) {
        double acc = 0.0;
        for (uint64_t k = off[i]; k < off[i + 1]; ++k)
            acc += x[col[k]] * val[k];
        y[i] = acc;
    }
}

void spira_or_spmv_f32(uint64_t n_rows, const uint64_t *off,
                       const uint32_t *col, const float *val,
                       const float *x, float *y)
{
    for (uint64_t i = 0; i < n_rows; ++i) {
        float acc = 0.0f;
        for (uint64_t k = off[i]; k < off[i + 1]; ++k)
            acc += x[col[k]] * val[k];
        y[i] = acc;
    }
}

/* TwoSum (Knuth): s + e == a + b exactly. */
static inline void two_sum(double a, double b, double *s, double *e)
{
    double t = a + b;
    double bb = t - a;
    *e = (a - (t - bb)) + (b - bb);
    *s = t;
}

void spira_or_spmv_dd(uint64_t n_rows, const uint64_t *off,
                      const uint32_t *col, const double *val,
                      const double *x, double *y_hi, double *y_lo)
{
    for (uint64_t i = 0; i < n_rows; ++i) {
        double hi = 0.0, lo = 0.0;
        for (uint64_t k = off[i]; k < off[i + 1]; ++k) {
            double a = val[k], b = x[col[k]];
            double p = a * b;
            double pe = fma(a, b, -p); /* TwoProd: p + pe == a*b exactly */
            double s, e;
            two_sum(hi, p, &s, &e);
            e += lo + pe;
            two_sum(s, e, &hi, &lo); /* renormalise */
        }
        y_hi[i] = hi;
        y_lo[i] = lo;
    }
}

void spira_or_spmm_f64(uint64_t n_rows, const uint64_t *off,
                       const uint32_t *col, const double *val, uint32_t k,
                       const double *X, double *Y)
{
    for (uint64_t i = 0; i < n_rows; ++i) {
        for (uint32_t j = 0; j < k; ++j) Y[i * k + j] = 0.0;
        for (uint64_t e = off[i]; e < off[i + 1]; ++e) {
            const double *xr = X + (uint64_t)col[e] * k;
            for (uint32_t j = 0; j < k; ++j) Y[i * k + j] += xr[j] * val[e];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Partition boundaries (partition.hpp:78-126)                               */
/* ------------------------------------------------------------------------ */

void spira_or_partition_boundaries(uint64_t n_rows, const uint64_t *row_nnz,
                                   uint64_t n_threads, uint64_t *b)
{
    for (uint64_t t = 0; t <= n_threads; ++t) b[t] = n_rows;
    b[0] = 0;
    if (n_threads <= 1 || n_rows == 0) return;
    uint64_t *prefix = (uint64_t *)malloc((n_rows + 1) * sizeof(uint64_t));
    prefix[0] = 0;
    for (uint64_t i = 0; i < n_rows; ++i) prefix[i + 1] = prefix[i] + row_nnz[i];
    uint64_t total = prefix[n_rows];
    if (total == 0) {
        for (uint64_t t = 1; t < n_threads; ++t) b[t] = t * n_rows / n_threads;
        free(prefix);
        return;
    }
    for (uint64_t t = 1; t < n_threads; ++t) {
        uint64_t target = t * total / n_threads;
        if (target < 1) target = 1;
        /* lower_bound over prefix[b[t-1] .. n_rows] */
        uint64_t lo = b[t - 1], hi = n_rows + 1;
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo) / 2;
            if (prefix[mid] < target) lo = mid + 1;
            else hi = mid;
        }
        b[t] = lo < n_rows ? lo : n_rows;
    }
    free(prefix);
}

uint64_t spira_or_transpose(uint64_t n_rows, uint64_t n_cols, const uint64_t *off,
                            const uint32_t *col, const double *val, int drop_zeros,
                            uint64_t *t_off, uint32_t *t_col, double *t_val)
{
    /* pass 1: histogram of the column indices */
    for (uint64_t j = 0; j <= n_cols; ++j) t_off[j] = 0;
    const uint64_t nnz = off[n_rows];
    for (uint64_t k = 0; k < nnz; ++k)
        if (!drop_zeros || !spira_or_is_zero(val[k])) t_off[(uint64_t)col[k] + 1] += 1;
    for (uint64_t j = 0; j < n_cols; ++j) t_off[j + 1] += t_off[j];
    /* pass 2: fill row by row through a cursor per output row */
    uint64_t *cursor = (uint64_t *)malloc((n_cols ? n_cols : 1) * sizeof(uint64_t));
    for (uint64_t j = 0; j < n_cols; ++j) cursor[j] = t_off[j];
    for (uint64_t i = 0; i < n_rows; ++i) {
        for (uint64_t k = off[i]; k < off[i + 1]; ++k) {
            if (drop_zeros && spira_or_is_zero(val[k])) continue;
            const uint64_t pos = cursor[col[k]]++;
            t_col[pos] = (uint32_t)i;
            t_val[pos] = val[k];
        }
    }
    free(cursor);
    return t_off[n_cols];
}

static int cmp_u32(const void *x, const void *y)
{
    const uint32_t a = *(const uint32_t *)x, b = *(const uint32_t *)y;
    return (a > b) - (a < b);
}

uint64_t spira_or_add(uint64_t n_rows, const uint64_t *a_off, const uint32_t *a_col, const double *a_val,
                      const uint64_t *b_off, const uint32_t *b_col, const double *b_val, int subtract,
                      uint64_t *c_off, uint32_t *c_col, double *c_val)
{
    /* serial/matrix_addition.hpp:34-86 addRows: a two-pointer merge per row;
     * a column in both rows inserts a + b unless that sum is zero, a column in
     * one row inserts its value; the output's first lock() then drops every
     * zero (csr_build.hpp:45-95).  A - B adds B scaled by -1
     * (matrix_operators.hpp:32-41, scalars.hpp:23-35: the scaled copy's lock
     * drops B's zeros, which the sum drops as well). */
    uint64_t nz = 0;
    c_off[0] = 0;
    for (uint64_t i = 0; i < n_rows; ++i) {
        uint64_t ka = a_off[i], kb = b_off[i];
        const uint64_t ea = a_off[i + 1], eb = b_off[i + 1];
        while (ka < ea || kb < eb) {
            uint32_t c;
            double v;
            if (kb >= eb || (ka < ea && a_col[ka] < b_col[kb])) {
                c = a_col[ka];
                v = a_val[ka++];
            } else {
                const double bv = subtract ? b_val[kb] * -1.0 : b_val[kb];
                if (ka >= ea || b_col[kb] < a_col[ka]) {
                    c = b_col[kb++];
                    v = bv;
                } else {
                    c = a_col[ka];
                    v = a_val[ka++] + bv;
                    ++kb;
                }
            }
            if (spira_or_is_zero(v)) continue;
            c_col[nz] = c;
            c_val[nz++] = v;
        }
        c_off[i + 1] = nz;
    }
    return nz;
}

int64_t spira_or_spgemm(uint64_t n_rows, uint64_t n_cols, const uint64_t *a_off, const uint32_t *a_col,
                        const double *a_val, const uint64_t *b_off, const uint32_t *b_col, const double *b_val,
                        uint64_t *c_off, uint32_t *c_col, double *c_val, uint64_t capacity)
{
    /* serial/spgemm.hpp:12-50 (Gustavson): per row of A, for its entries in
     * column order and then B's row entries in column order, acc[j] += a * b
     * with acc[j] starting at 0.0 (product rounded, then the sum); entries
     * with acc[j] != 0 are inserted and the output's lock() sorts the row.
     * The reference's hash map only fixes which j are present, not the order
     * of the additions into one j, so a dense accumulator with a touched list
     * gives the same bits. */
    double *acc = (double *)malloc((n_cols ? n_cols : 1) * sizeof(double));
    uint64_t *mark = (uint64_t *)calloc(n_cols ? n_cols : 1, sizeof(uint64_t));
    uint32_t *touched = (uint32_t *)malloc((n_cols ? n_cols : 1) * sizeof(uint32_t));
    uint64_t nz = 0;
    c_off[0] = 0;
    for (uint64_t i = 0; i < n_rows; ++i) {
        uint64_t nt = 0;
        for (uint64_t ka = a_off[i]; ka < a_off[i + 1]; ++ka) {
            const uint32_t k = a_col[ka];
            const double a = a_val[ka];
            for (uint64_t kb = b_off[k]; kb < b_off[k + 1]; ++kb) {
                const uint32_t j = b_col[kb];
                if (mark[j] != i + 1) {
                    mark[j] = i + 1;
                    acc[j] = 0.0;
                    touched[nt++] = j;
                }
                const double p = a * b_val[kb];
                acc[j] = acc[j] + p;
            }
        }
        qsort(touched, nt, sizeof(uint32_t), cmp_u32);
        for (uint64_t t = 0; t < nt; ++t) {
            const double v = acc[touched[t]];
            if (!(v != 0.0)) continue;
            if (nz >= capacity) {
                free(acc);
                free(mark);
                free(touched);
                return -1;
            }
            c_col[nz] = touched[t];
            c_val[nz++] = v;
        }
        c_off[i + 1] = nz;
    }
    free(acc);
    free(mark);
    free(touched);
    return (int64_t)nz;
}

/* ------------------------------------------------------------------------ */
/* Threaded baseline                                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t row_start, row_end;
    int built;
    uint64_t nnz;
    uint64_t *off; /* local, rebased */
    uint32_t *col;
    double *val;
} part_t;

typedef enum { JOB_NONE, JOB_LOAD, JOB_SPMV, JOB_LOCK, JOB_FLUSH, JOB_EXIT } job_t;

struct spira_or_pool {
    uint32_t n;
    uint64_t n_rows, n_cols;
    part_t *parts;
    pthread_t *threads;
    pthread_mutex_t mu;
    pthread_cond_t cv_start, cv_done;
    uint64_t generation;
    uint32_t remaining;
    job_t job;
    /* job arguments */
    const uint64_t *a_off;
    const uint32_t *a_col;
    const double *a_val;
    const double *a_x;
    double *a_y;
    int failed;
    /* row dot product of JOB_SPMV: the scalar loop above (reference order) or
     * the per-row SIMD gather dot that dispatch.cpp:34-67 selects */
    double (*dot)(const double *vals, const uint32_t *cols, const double *x, uint64_t n);
    const int *cpus;     /* JOB_PIN: CPU of each worker */
    uint64_t flush_bytes; /* JOB_FLUSH: bytes each worker reads */
    char **arenas;       /* JOB_FLUSH: one arena per worker */
    uint64_t sink;       /* keeps the flush reads live */
};

typedef struct {
    spira_or_pool *p;
    uint32_t id;
} worker_arg_t;

static void job_load(spira_or_pool *p, part_t *q)
{
    uint64_t r0 = q->row_start, r1 = q->row_end, n = r1 - r0;
    uint64_t base = p->a_off[r0], nnz = p->a_off[r1] - base;
    free(q->off); free(q->col); free(q->val);
    q->off = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    q->col = (uint32_t *)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    q->val = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
    for (uint64_t i = 0; i <= n; ++i) q->off[i] = p->a_off[r0 + i] - base;
    memcpy(q->col, p->a_col + base, nnz * sizeof(uint32_t));
    memcpy(q->val, p->a_val + base, nnz * sizeof(double));
    q->nnz = nnz;
    q->built = 1;
}

static void job_spmv(spira_or_pool *p, part_t *q)
{
    const double *x = p->a_x;
    double *y = p->a_y + q->row_start;
    uint64_t n = q->row_end - q->row_start;
    if (p->dot) {
        for (uint64_t i = 0; i < n; ++i)
            y[i] = p->dot(q->val + q->off[i], q->col + q->off[i], x, q->off[i + 1] - q->off[i]);
        return;
    }
    for (uint64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (uint64_t k = q->off[i]; k < q->off[i + 1]; ++k)
            acc += x[q->col[k]] * q->val[k];
        y[i] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* Per-row SIMD gather dots of the CPU baseline (timing only; the oracle's  */
/* own SpMV above keeps the reference's scalar order).  Four accumulators    */
/* over blocks of four vectors, then single vectors, a fixed pairwise        */
/* reduction and a scalar tail: the shape of dot_avx.cpp:6-56 (4 doubles per */
/* gather) and dot_avx512.cpp:7-56 (8 per gather), which the reference picks */
/* at start-up (dispatch.cpp:34-47: AVX-512F first, then AVX2 + FMA).        */
/* ------------------------------------------------------------------------ */
__attribute__((target("avx2,fma"))) static double dot_avx2(const double *v, const uint32_t *c,
                                                          const double *x, uint64_t n)
{
    __m256d a[4] = {_mm256_setzero_pd(), _mm256_setzero_pd(), _mm256_setzero_pd(), _mm256_setzero_pd()};
    uint64_t i = 0;
    for (; i + 16 <= n; i += 16)
        for (int j = 0; j < 4; ++j) {
            const __m128i ix = _mm_loadu_si128((const __m128i *)(c + i + 4 * j));
            a[j] = _mm256_fmadd_pd(_mm256_loadu_pd(v + i + 4 * j), _mm256_i32gather_pd(x, ix, 8), a[j]);
        }
    for (; i + 4 <= n; i += 4) {
        const __m128i ix = _mm_loadu_si128((const __m128i *)(c + i));
        a[0] = _mm256_fmadd_pd(_mm256_loadu_pd(v + i), _mm256_i32gather_pd(x, ix, 8), a[0]);
    }
    const __m256d s = _mm256_add_pd(_mm256_add_pd(a[0], a[1]), _mm256_add_pd(a[2], a[3]));
    const __m128d h = _mm_add_pd(_mm256_castpd256_pd128(s), _mm256_extractf128_pd(s, 1));
    double acc = _mm_cvtsd_f64(_mm_add_sd(h, _mm_unpackhi_pd(h, h)));
    for (; i < n; ++i) acc += v[i] * x[c[i]];
    return acc;
}

__attribute__((target("avx512f"))) static double dot_avx512(const double *v, const uint32_t *c,
                                                          const double *x, uint64_t n)
{
    __m512d a[4] = {_mm512_setzero_pd(), _mm512_setzero_pd(), _mm512_setzero_pd(), _mm512_setzero_pd()};
    uint64_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int j = 0; j < 4; ++j) {
            const __m256i ix = _mm256_loadu_si256((const __m256i *)(c + i + 8 * j));
            a[j] = _mm512_fmadd_pd(_mm512_loadu_pd(v + i + 8 * j), _mm512_i32gather_pd(ix, x, 8), a[j]);
        }
    for (; i + 8 <= n; i += 8) {
        const __m256i ix = _mm256_loadu_si256((const __m256i *)(c + i));
        a[0] = _mm512_fmadd_pd(_mm512_loadu_pd(v + i), _mm512_i32gather_pd(ix, x, 8), a[0]);
    }
    double acc = _mm512_reduce_add_pd(_mm512_add_pd(_mm512_add_pd(a[0], a[1]), _mm512_add_pd(a[2], a[3])));
    for (; i < n; ++i) acc += x[c[i]] * v[i];
    return acc;
}

static double dot_scalar(const double *v, const uint32_t *c, const double *x, uint64_t n)
{
    double acc = 0.0;
    for (uint64_t i = 0; i < n; ++i) acc += v[i] * x[c[i]];
    return acc;
}

/* Cold-cache mode of bench/spira_bench.cpp:43-67: read one byte per 64-byte
 * line of an arena of the last-level cache's size.  Each worker reads its
 * own arena, so every core's share of the LLC is evicted. */
static void job_flush(spira_or_pool *p, uint32_t id)
{
    char *a = p->arenas[id];
    uint64_t s = 0;
    for (uint64_t i = 0; i < p->flush_bytes; i += 64) s += (unsigned char)a[i];
    __atomic_fetch_add(&p->sink, s, __ATOMIC_RELAXED);
}

/* lock_partition (parallel_matrix.hpp:321-386) over flat staged segments. */
static int job_lock(spira_or_pool *p, part_t *q)
{
    uint64_t r0 = q->row_start, r1 = q->row_end, n = r1 - r0;
    const uint64_t *sp = p->a_off;
    uint64_t sbase = sp[r0], m = sp[r1] - sbase;
    uint64_t ms = max_segment(r0, r1, sp);
    entry_t *scratch = (entry_t *)malloc((2 * ms + 1) * sizeof(entry_t));
    uint64_t *dptr = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    uint32_t *dcol = (uint32_t *)malloc((m ? m : 1) * sizeof(uint32_t));
    double *dval = (double *)malloc((m ? m : 1) * sizeof(double));
    if (!scratch || !dptr || !dcol || !dval) return -1;
    /* lock_for_compact per row (:329-334) */
    uint64_t w = 0;
    dptr[0] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t b = sp[r0 + i] - sbase, k = sp[r0 + i + 1] - sp[r0 + i];
        w += dedup_segment(p->a_col + sbase + b, p->a_val + sbase + b, k,
                           scratch, dcol + w, dval + w);
        dptr[i + 1] = w;
    }
    free(scratch);
    if (!q->built) { /* build_csr (:358) */
        uint64_t *off = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
        uint32_t *col = (uint32_t *)malloc((w ? w : 1) * sizeof(uint32_t));
        double *val = (double *)malloc((w ? w : 1) * sizeof(double));
        q->nnz = spira_or_build_csr(n, dptr, dcol, dval, off, col, val);
        q->off = off; q->col = col; q->val = val;
        q->built = 1;
    } else { /* merge_csr with dirty = buffer non-empty (:346-356) */
        uint64_t cap = q->nnz + w;
        uint64_t *off = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
        uint32_t *col = (uint32_t *)malloc((cap ? cap : 1) * sizeof(uint32_t));
        double *val = (double *)malloc((cap ? cap : 1) * sizeof(double));
        q->nnz = spira_or_merge_csr(n, q->off, q->col, q->val, dptr, dcol, dval,
                                    NULL, off, col, val);
        free(q->off); free(q->col); free(q->val);
        q->off = off; q->col = col; q->val = val;
    }
    free(dptr); free(dcol); free(dval);
    return 0;
}

static void *worker_main(void *arg)
{
    worker_arg_t *wa = (worker_arg_t *)arg;
    spira_or_pool *p = wa->p;
    uint32_t id = wa->id;
    free(wa);
    uint64_t seen = 0;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        while (p->generation == seen) pthread_cond_wait(&p->cv_start, &p->mu);
        seen = p->generation;
        job_t job = p->job;
        pthread_mutex_unlock(&p->mu);
        if (job == JOB_EXIT) return NULL;
        part_t *q = &p->parts[id];
        int rc = 0;
        if (job == JOB_LOAD) job_load(p, q);
        else if (job == JOB_SPMV) job_spmv(p, q);
        else if (job == JOB_LOCK) rc = job_lock(p, q);
        else if (job == JOB_FLUSH) job_flush(p, id);
        pthread_mutex_lock(&p->mu);
        if (rc) p->failed = 1;
        if (--p->remaining == 0) pthread_cond_signal(&p->cv_done);
        pthread_mutex_unlock(&p->mu);
    }
}

/* thread_pool::execute (thread_pool.hpp:102-112): publish, wake all, wait. */
static int pool_run(spira_or_pool *p, job_t job)
{
    pthread_mutex_lock(&p->mu);
    p->job = job;
    p->remaining = p->n;
    p->failed = 0;
    p->generation++;
    pthread_cond_broadcast(&p->cv_start);
    if (job != JOB_EXIT)
        while (p->remaining != 0) pthread_cond_wait(&p->cv_done, &p->mu);
    int f = p->failed;
    pthread_mutex_unlock(&p->mu);
    return f ? -1 : 0;
}

spira_or_pool *spira_or_pool_create(uint32_t n_threads, uint64_t n_rows,
                                    uint64_t n_cols)
{
    if (n_threads < 1) return NULL; /* parallel_matrix.hpp:293-294 */
    spira_or_pool *p = (spira_or_pool *)calloc(1, sizeof(spira_or_pool));
    p->n = n_threads;
    p->n_rows = n_rows;
    p->n_cols = n_cols;
    p->parts = (part_t *)calloc(n_threads, sizeof(part_t));
    for (uint32_t t = 0; t < n_threads; ++t) {
        p->parts[t].row_start = (uint64_t)t * n_rows / n_threads;
        p->parts[t].row_end = (uint64_t)(t + 1) * n_rows / n_threads;
    }
    pthread_mutex_init(&p->mu, NULL);
    pthread_cond_init(&p->cv_start, NULL);
    pthread_cond_init(&p->cv_done, NULL);
    p->threads = (pthread_t *)calloc(n_threads, sizeof(pthread_t));
    for (uint32_t t = 0; t < n_threads; ++t) {
        worker_arg_t *wa = (worker_arg_t *)malloc(sizeof(worker_arg_t));
        wa->p = p;
        wa->id = t;
        pthread_create(&p->threads[t], NULL, worker_main, wa);
    }
    return p;
}

void spira_or_pool_destroy(spira_or_pool *p)
{
    if (!p) return;
    pool_run(p, JOB_EXIT);
    for (uint32_t t = 0; t < p->n; ++t) pthread_join(p->threads[t], NULL);
    for (uint32_t t = 0; t < p->n; ++t) {
        free(p->parts[t].off); free(p->parts[t].col); free(p->parts[t].val);
    }
    if (p->arenas) {
        for (uint32_t t = 0; t < p->n; ++t) free(p->arenas[t]);
        free(p->arenas);
    }
    free(p->parts);
    free(p->threads);
    pthread_mutex_destroy(&p->mu);
    pthread_cond_destroy(&p->cv_start);
    pthread_cond_destroy(&p->cv_done);
    free(p);
}

uint32_t spira_or_pool_threads(const spira_or_pool *p) { return p->n; }

int spira_or_pool_set_kernel(spira_or_pool *p, int simd)
{
    if (!simd) {
        p->dot = NULL;
        return 0;
    }
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f")) {
        p->dot = dot_avx512;
        return 3;
    }
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) {
        p->dot = dot_avx2;
        return 2;
    }
    p->dot = dot_scalar;
    return 1;
}

int spira_or_pool_force_kernel(spira_or_pool *p, int which)
{
    __builtin_cpu_init();
    if (which == 3 && !__builtin_cpu_supports("avx512f")) return -1;
    if (which == 2 && !(__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))) return -1;
    p->dot = which == 3 ? dot_avx512 : which == 2 ? dot_avx2 : which == 1 ? dot_scalar : NULL;
    return 0;
}

int spira_or_pool_pin(spira_or_pool *p, const int *cpus)
{
    int failed = 0;
    for (uint32_t t = 0; t < p->n; ++t) {
        cpu_set_t set;
        CPU_ZERO(&set);
        CPU_SET(cpus[t], &set);
        if (pthread_setaffinity_np(p->threads[t], sizeof(set), &set) != 0) failed = 1;
    }
    return failed ? -1 : 0;
}

int spira_or_pool_flush(spira_or_pool *p, uint64_t bytes)
{
    if (!p->arenas || p->flush_bytes != bytes) {
        if (p->arenas) {
            for (uint32_t t = 0; t < p->n; ++t) free(p->arenas[t]);
            free(p->arenas);
        }
        p->arenas = (char **)calloc(p->n, sizeof(char *));
        for (uint32_t t = 0; t < p->n; ++t) {
            p->arenas[t] = (char *)malloc(bytes ? bytes : 1);
            if (!p->arenas[t]) return -1;
            memset(p->arenas[t], 1, bytes);
        }
        p->flush_bytes = bytes;
    }
    return pool_run(p, JOB_FLUSH);
}

void spira_or_pool_partition(const spira_or_pool *p, uint32_t t,
                             uint64_t *row_start, uint64_t *row_end)
{
    *row_start = p->parts[t].row_start;
    *row_end = p->parts[t].row_end;
}

int spira_or_pool_load_csr(spira_or_pool *p, const uint64_t *off,
                           const uint32_t *col, const double *val)
{
    p->a_off = off; p->a_col = col; p->a_val = val;
    return pool_run(p, JOB_LOAD);
}

int spira_or_pool_spmv(spira_or_pool *p, const double *x, double *y)
{
    for (uint32_t t = 0; t < p->n; ++t)
        if (!p->parts[t].built) return -2; /* spmv on an unlocked matrix */
    p->a_x = x; p->a_y = y;
    return pool_run(p, JOB_SPMV);
}

int spira_or_pool_lock(spira_or_pool *p, const uint64_t *stage_ptr,
                       const uint32_t *stage_col, const double *stage_val)
{
    p->a_off = stage_ptr; p->a_col = stage_col; p->a_val = stage_val;
    return pool_run(p, JOB_LOCK);
}

uint64_t spira_or_pool_nnz(const spira_or_pool *p)
{
    uint64_t s = 0;
    for (uint32_t t = 0; t < p->n; ++t) s += p->parts[t].nnz;
    return s;
}

int spira_or_pool_export_csr(const spira_or_pool *p, uint64_t *off,
                             uint32_t *col, double *val)
{
    uint64_t base = 0;
    off[0] = 0;
    for (uint32_t t = 0; t < p->n; ++t) {
        const part_t *q = &p->parts[t];
        uint64_t n = q->row_end - q->row_start;
        if (!q->built) return -2;
        for (uint64_t i = 0; i < n; ++i) off[q->row_start + i + 1] = base + q->off[i + 1];
        memcpy(col + base, q->col, q->nnz * sizeof(uint32_t));
        memcpy(val + base, q->val, q->nnz * sizeof(double));
        base += q->nnz;
    }
    return 0;
}
