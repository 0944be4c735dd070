/*
 * spira_oracle.h -- CPU restatement of Spira's locked-mode hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against, and the CPU baseline that bench.py times beside it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product (libspira_amd.so) never links or calls it.
 *
 * Conventions (all follow the reference's SoA / uint32_t / double instantiation):
 *   offsets  uint64_t[n_rows+1]   (std::size_t in csr_storage_soa.hpp:23)
 *   cols     uint32_t[nnz]        (concepts.hpp:10-15, soa_tag + uint32_t)
 *   vals     double[nnz]
 * Staged (open-mode) inserts are handed over as per-row segments in insertion
 * order: stage_ptr[n_rows+1], stage_col[m], stage_val[m].
 *
 * Pinning status of every function is stated in oracle/README.md and checked by
 * tests/test_oracle_*.py.
 */
#ifndef SPIRA_ORACLE_H
#define SPIRA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* is_zero(v) <=> |v| <= 0 (traits.hpp:73-88): true for +-0.0, false for NaN. */
int spira_or_is_zero(double v);

/* a2: per-row sort + dedup, last write wins, zeros kept
 * (aos_array_buffer.hpp:76-101: reverse, stable_sort by column, keep first).
 * Output segments are compacted: dptr[n_rows+1] is the new per-row pointer,
 * dcol/dval must hold stage_ptr[n_rows] entries.  Returns total kept. */
uint64_t spira_or_dedup(uint64_t n_rows, const uint64_t *stage_ptr,
                        const uint32_t *stage_col, const double *stage_val,
                        uint64_t *dptr, uint32_t *dcol, double *dval);

/* a3: first-lock build (csr_build.hpp:32-98): count non-zeros per row,
 * exclusive scan into off, copy non-zeros.  Input must be deduped + sorted.
 * col/val must hold dptr[n_rows] entries.  Returns nnz. */
uint64_t spira_or_build_csr(uint64_t n_rows, const uint64_t *dptr,
                            const uint32_t *dcol, const double *dval,
                            uint64_t *off, uint32_t *col, double *val);

/* a4: re-lock merge (csr_build.hpp:126-488, grow/forward path; the reuse path
 * produces identical arrays).  Clean rows (dirty[i]==0) are copied verbatim;
 * dirty rows are merged with buffer-wins and every zero dropped.  dirty may be
 * NULL, meaning dirty[i] = (deduped buffer of row i non-empty), which is the
 * parallel_matrix rule (parallel_matrix.hpp:351-355).  col/val must hold
 * old_off[n_rows] + dptr[n_rows] entries.  Returns the new nnz. */
uint64_t spira_or_merge_csr(uint64_t n_rows, const uint64_t *old_off,
                            const uint32_t *old_col, const double *old_val,
                            const uint64_t *dptr, const uint32_t *dcol,
                            const double *dval, const uint8_t *dirty,
                            uint64_t *off, uint32_t *col, double *val);

/* a6: FP64 SpMV in the reference's scalar order (serial/spmv.hpp:48-54):
 * acc = 0; acc += x[col]*val (no fused multiply-add); y overwritten. */
void spira_or_spmv_f64(uint64_t n_rows, const uint64_t *off,
                       const uint32_t *col, const double *val,
                       const double *x, double *y);

/* Same loop with FP32 values and vectors and an FP32 accumulator
 * (dot_scalar.cpp:13, the float instantiation). */
void spira_or_spmv_f32(uint64_t n_rows, const uint64_t *off,
                       const uint32_t *col, const float *val,
                       const float *x, float *y);

/* Error reference: double-double accumulation (TwoProd by fma + TwoSum,
 * Dekker 1971 / Knuth TAOCP 4.2.2).  y = y_hi + y_lo to ~106 bits. */
void spira_or_spmv_dd(uint64_t n_rows, const uint64_t *off,
                      const uint32_t *col, const double *val,
                      const double *x, double *y_hi, double *y_lo);

/* Multi-RHS FP64 reference: X is n_cols x k row-major, Y n_rows x k row-major. */
void spira_or_spmm_f64(uint64_t n_rows, const uint64_t *off,
                       const uint32_t *col, const double *val, uint32_t k,
                       const double *X, double *Y);

/* ---------------------------------------------------------------------------
 * Threaded CPU baseline: restatement of parallel_matrix (uniform row split,
 * parallel_matrix.hpp:296-306), one CSR per partition first-touched by its own
 * worker (partition.hpp:44-47), a persistent worker pool released by a
 * generation counter (thread_pool.hpp:102-134), SpMV per partition
 * (parallel/algorithms/spmv.hpp:140-151) and lock per partition
 * (parallel_matrix.hpp:321-386).
 * ------------------------------------------------------------------------- */
typedef struct spira_or_pool spira_or_pool;

spira_or_pool *spira_or_pool_create(uint32_t n_threads, uint64_t n_rows,
                                    uint64_t n_cols);
void spira_or_pool_destroy(spira_or_pool *p);
uint32_t spira_or_pool_threads(const spira_or_pool *p);
/* Row dot product of pool SpMV: 0 = the scalar reference-order loop (the
 * default); nonzero = the SIMD gather dot the reference's dispatch.cpp:34-67
 * would pick on this CPU.  Returns 0 scalar loop, 1 scalar dot, 2 AVX2, 3
 * AVX-512. */
int spira_or_pool_set_kernel(spira_or_pool *p, int simd);
/* Force one dot: 0 loop, 1 scalar dot, 2 AVX2, 3 AVX-512 (-1 if unsupported). */
int spira_or_pool_force_kernel(spira_or_pool *p, int which);
/* Pin worker t to CPU cpus[t] (pthread_setaffinity_np).  0 on success. */
int spira_or_pool_pin(spira_or_pool *p, const int *cpus);
/* Every worker reads one byte per 64-byte line of its own arena of `bytes`
 * (allocated and touched on first use): the cold-LLC mode. */
int spira_or_pool_flush(spira_or_pool *p, uint64_t bytes);
/* Row range of partition t (uniform split). */
void spira_or_pool_partition(const spira_or_pool *p, uint32_t t,
                             uint64_t *row_start, uint64_t *row_end);
/* Install a global CSR: each worker copies its own slice (rebased offsets). */
int spira_or_pool_load_csr(spira_or_pool *p, const uint64_t *off,
                           const uint32_t *col, const double *val);
/* y = A*x over all partitions. */
int spira_or_pool_spmv(spira_or_pool *p, const double *x, double *y);
/* lock(): each worker dedups its rows' staged segments (global row indexing)
 * and builds (first lock) or merges its partition CSR.  Returns 0 on success. */
int spira_or_pool_lock(spira_or_pool *p, const uint64_t *stage_ptr,
                       const uint32_t *stage_col, const double *stage_val);
/* Total nnz and a gather of the partition CSRs back into one global CSR. */
uint64_t spira_or_pool_nnz(const spira_or_pool *p);
int spira_or_pool_export_csr(const spira_or_pool *p, uint64_t *off,
                             uint32_t *col, double *val);

/* Nnz-balanced partition boundaries (partition.hpp:78-126). */
void spira_or_partition_boundaries(uint64_t n_rows, const uint64_t *row_nnz,
                                   uint64_t n_threads, uint64_t *boundaries);

/* Transpose of a CSR (serial/transpose.hpp:35-157: column histogram, offsets,
 * then a fill in ascending input-row order, so output rows are sorted).  With
 * drop_zeros, parallel/algorithms/transpose.hpp:37-69: the same entries
 * inserted into a new matrix whose lock() removes explicit zeros.  t_off holds
 * n_cols + 1 entries, t_col / t_val the input's nnz; returns the output nnz. */
uint64_t spira_or_transpose(uint64_t n_rows, uint64_t n_cols, const uint64_t *off,
                            const uint32_t *col, const double *val, int drop_zeros,
                            uint64_t *t_off, uint32_t *t_col, double *t_val);

/* C = A + B, or A - B with subtract (serial/matrix_addition.hpp:34-112,
 * matrix_operators.hpp:21-41): rows merged by column, sums and single values
 * kept unless zero.  c_* hold at least nnz(A) + nnz(B) entries.  Returns
 * nnz(C). */
uint64_t spira_or_add(uint64_t n_rows, const uint64_t *a_off, const uint32_t *a_col, const double *a_val,
                      const uint64_t *b_off, const uint32_t *b_col, const double *b_val, int subtract,
                      uint64_t *c_off, uint32_t *c_col, double *c_val);

/* C = A * B (serial/spgemm.hpp:12-50): Gustavson rows, acc[j] += a * b in
 * A's then B's column order, entries != 0 kept, rows sorted.  n_cols = B's
 * columns.  Returns nnz(C), or -1 when it exceeds capacity. */
int64_t spira_or_spgemm(uint64_t n_rows, uint64_t n_cols, const uint64_t *a_off, const uint32_t *a_col,
                        const double *a_val, const uint64_t *b_off, const uint32_t *b_col, const double *b_val,
                        uint64_t *c_off, uint32_t *c_col, double *c_val, uint64_t capacity);

#ifdef __cplusplus
}
#endif

#endif
