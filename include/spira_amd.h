/*
 * spira_amd.h -- C-ABI of libspira_amd: Spira's locked-mode hot path
 * (lock() and SpMV) on one MI355X (gfx950).
 *
 * Plain C types only: no torch, no HIP types in the signatures.  The stream is
 * passed as an opaque `void *` (a hipStream_t).  Every function returns a
 * status code (SPIRA_OK == 0); the message of the last failure on a handle is
 * spira_dev_last_error().  Status codes map 1:1 onto the reference's
 * exceptions so a C++ wrapper can rethrow exactly what Spira throws.
 *
 * Threading: calls on one spira_dev are serialised on its stream and must not
 * overlap from several host threads; separate handles may run concurrently.
 * Host arrays are borrowed for the duration of a call; device objects
 * (spira_dcsr) are owned by the caller until spira_dcsr_destroy().
 */
#ifndef SPIRA_AMD_H
#define SPIRA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spira_dev spira_dev;   /* device ordinal, stream, workspace */
typedef struct spira_dcsr spira_dcsr; /* device CSR: u64 offsets, u32 cols, f64 vals */

/* ---- status codes ------------------------------------------------------- */
#define SPIRA_OK 0
#define SPIRA_ERR_INVALID_ARGUMENT 1 /* std::invalid_argument */
#define SPIRA_ERR_LOGIC 2            /* std::logic_error */
#define SPIRA_ERR_OUT_OF_RANGE 3     /* std::out_of_range */
#define SPIRA_ERR_NOMEM 4            /* std::bad_alloc */
#define SPIRA_ERR_DEVICE 5           /* HIP runtime failure */
#define SPIRA_ERR_UNSUPPORTED 6      /* input outside this build's limits */

/* ---- pointer-location and behaviour flags ------------------------------ */
#define SPIRA_PTR_HOST 0u
#define SPIRA_X_DEVICE 1u   /* x (or X) is a device pointer */
#define SPIRA_Y_DEVICE 2u   /* y (or Y) is a device pointer */
#define SPIRA_IN_DEVICE 4u  /* upload / staged input arrays are device pointers */
#define SPIRA_OUT_DEVICE 8u /* download destinations are device pointers */
#define SPIRA_NO_SYNC 16u   /* return without synchronising the stream
                               (device pointers only) */
#define SPIRA_REPRODUCIBLE 32u /* SpMV: bit-identical y on every call with the
                                  same A and x (AUTO keeps to variants whose
                                  summation order is fixed) */
#define SPIRA_COUNT_GIVEN 64u  /* lock with SPIRA_IN_DEVICE: *out_nnz holds the
                                  staged count m (the length of stage_col and
                                  stage_val) on entry, so stage_ptr[n_rows] is
                                  not read back first; m != stage_ptr[n_rows]
                                  gives SPIRA_ERR_INVALID_ARGUMENT */

/* ---- SpMV kernel variants ------------------------------------------------ */
#define SPIRA_SPMV_AUTO (-1)      /* cached choice, else the built-in heuristic */
#define SPIRA_SPMV_VEC1 0         /* one lane per row */
#define SPIRA_SPMV_VEC2 1         /* G lanes per row, shuffle reduction ... */
#define SPIRA_SPMV_VEC4 2
#define SPIRA_SPMV_VEC8 3
#define SPIRA_SPMV_VEC16 4
#define SPIRA_SPMV_VEC32 5
#define SPIRA_SPMV_VEC64 6        /* one wavefront per row */
#define SPIRA_SPMV_STREAM1K 7     /* LDS-staged row blocks of <= 1024 rows+nnz */
#define SPIRA_SPMV_STREAM2K 8     /* LDS-staged row blocks of <= 2048 rows+nnz */
#define SPIRA_SPMV_XSLICE1 9      /* column slice per XCD, 1 lane per row, then a sum of 8 partials */
#define SPIRA_SPMV_XSLICE2 10     /* ... 2 lanes per row */
#define SPIRA_SPMV_XSLICE4 11     /* ... 4 lanes per row */
#define SPIRA_SPMV_XSTREAM 12     /* column slice per XCD, LDS row blocks per slice */
#define SPIRA_SPMV_XPHASE 13      /* row blocks with y in LDS, x slices in L2-sized phases per XCD */
#define SPIRA_SPMV_XPHASE_PACED 14 /* ... 20 entries per thread per step, an XCD's blocks paced to within 3 slices */
#define SPIRA_SPMV_N_VARIANTS 15

/* ---- device handle --------------------------------------------------------
 * Replaces the reference's CPU dispatch role (dispatch.cpp:34-67,
 * runtime_config.hpp:7-31): one handle per GPU and stream. */
int spira_dev_create(spira_dev **out, int device_ordinal, void *stream);
void spira_dev_destroy(spira_dev *dev);
const char *spira_dev_last_error(const spira_dev *dev);
void *spira_dev_stream(const spira_dev *dev);
int spira_dev_synchronize(spira_dev *dev);

/* ---- device CSR ------------------------------------------------------------
 * csr_storage<soa_tag, uint32_t, double> (csr_storage_soa.hpp:17-76) on the
 * device.  A fresh handle is "not built", like matrix::csr_ before the first
 * lock (matrix.hpp:178, :306). */
int spira_dcsr_create(spira_dev *dev, uint64_t n_rows, uint64_t n_cols, spira_dcsr **out);
void spira_dcsr_destroy(spira_dcsr *A);
/* n_rows, n_cols, nnz and whether a CSR has been built (any output may be NULL). */
int spira_dcsr_info(const spira_dcsr *A, uint64_t *n_rows, uint64_t *n_cols, uint64_t *nnz, int *built);

/* matrix::load_csr (matrix.hpp:520-605): validate and install an external CSR.
 * Validation order and messages follow matrix.hpp:524-575 (offsets[0] == 0,
 * offsets[n] == nnz, non-decreasing offsets, in-bounds and strictly increasing
 * columns); every failure is SPIRA_ERR_INVALID_ARGUMENT.  Explicit zero values
 * are accepted, as in the reference.  flags: SPIRA_IN_DEVICE. */
int spira_dcsr_upload_f64(spira_dcsr *A, uint64_t nnz, const uint64_t *offsets,
                          const uint32_t *cols, const double *vals, uint32_t flags);

/* matrix::csr() read-out (matrix.hpp:88-91): copy offsets[n_rows+1],
 * cols[nnz], vals[nnz] out (any pointer may be NULL to skip that array).
 * flags: SPIRA_OUT_DEVICE.  SPIRA_ERR_LOGIC if no CSR has been built. */
int spira_dcsr_download_f64(const spira_dcsr *A, uint64_t *offsets, uint32_t *cols,
                            double *vals, uint32_t flags);

/* ---- lock() ---------------------------------------------------------------
 * matrix::lock() / parallel_matrix::lock_partition (matrix.hpp:277-342,
 * parallel_matrix.hpp:321-386) with build_csr / merge_csr
 * (csr_build.hpp:32-98, :126-488) on the device.
 *   stage_ptr[n_rows+1], stage_col[m], stage_val[m]: the open-mode inserts of
 *   every row, in insertion order (row r owns [stage_ptr[r], stage_ptr[r+1])).
 *   dirty: optional u8[n_rows]; NULL means dirty[r] = (row r has inserts), the
 *   parallel_matrix rule (parallel_matrix.hpp:351-355).  matrix::dirty_ is
 *   passed explicitly when it differs (row_at_mut without an insert).
 * First call on an unbuilt A = build_csr; later calls = merge_csr.  Result is
 * bit-identical to the reference's offsets/cols/vals.  *out_nnz (may be NULL)
 * receives the new nnz.  Column indices >= n_cols give SPIRA_ERR_OUT_OF_RANGE
 * (matrix.hpp:167-171).  Rows may hold any number of staged inserts
 * (fewer than 2^32 in one call, as for the whole call).
 * flags: SPIRA_IN_DEVICE, SPIRA_COUNT_GIVEN (with SPIRA_IN_DEVICE; out_nnz must
 * then be given). */
int spira_lock_f64(spira_dcsr *A, const uint64_t *stage_ptr, const uint32_t *stage_col,
                   const double *stage_val, const uint8_t *dirty, uint32_t flags,
                   uint64_t *out_nnz);

/* ---- open-mode staging on the device --------------------------------------
 * The reference's open-mode inserts (matrix::insert, matrix.hpp:453-462;
 * parallel_matrix::insert and its staging arrays, parallel_matrix.hpp:581-602,
 * insert_staging.hpp:58-96) as one device COO append buffer kept in insertion
 * order.  spira_stage_append copies n (row, col, value) triplets (host arrays,
 * or device arrays with SPIRA_IN_DEVICE) to its end; if any row >= n_rows or
 * col >= n_cols the whole append is rejected with SPIRA_ERR_OUT_OF_RANGE
 * (matrix.hpp:161-171 messages).  spira_lock_staged_f64 is spira_lock_f64 on
 * the buffer's contents: a stable radix sort by row on the device forms the
 * per-row segments (each row's inserts stay in insertion order), then the same
 * dedup + build/merge runs; the buffer is emptied on success.  dirty as in
 * spira_lock_f64 (host array, or device array with SPIRA_IN_DEVICE).  Fewer
 * than 2^32 staged inserts; n_rows < 2^32. */
typedef struct spira_stage spira_stage;
int spira_stage_create(spira_dev *dev, uint64_t n_rows, uint64_t n_cols, spira_stage **out);
void spira_stage_destroy(spira_stage *s);
int spira_stage_append(spira_stage *s, const uint32_t *rows, const uint32_t *cols, const double *vals, uint64_t n,
                       uint32_t flags);
int spira_stage_size(const spira_stage *s, uint64_t *n);
int spira_stage_clear(spira_stage *s);
int spira_lock_staged_f64(spira_dcsr *A, spira_stage *s, const uint8_t *dirty, uint32_t flags, uint64_t *out_nnz);

/* ---- SpMV -----------------------------------------------------------------
 * serial::algorithms::spmv / parallel::algorithms::spmv for
 * <soa_tag, uint32_t, double> (serial/spmv.hpp:142-186,
 * parallel/algorithms/spmv.hpp:127-161): y = A*x, y overwritten, empty rows
 * give +0.0.  Checks in the reference's order: x_len != n_cols and
 * y_len != n_rows give SPIRA_ERR_INVALID_ARGUMENT, an unbuilt A gives
 * SPIRA_ERR_LOGIC.  FP64 throughout; normwise relative error vs. the
 * sequential FP64 loop <= 1e-12.  variant: SPIRA_SPMV_*; SPIRA_SPMV_AUTO is
 * the pinned variant (spira_spmv_set_variant, the autotuner) or else the
 * learned selector's choice for A (computed on the device once per locked
 * matrix).  When that choice copies A into a layout of its own (XSLICE,
 * XSTREAM, XPHASE: several passes over A, worth some 5-10 calls), the copy is
 * made by the second AUTO call after a lock or upload; the first AUTO call
 * runs the fastest variant that needs no copy (VEC, STREAM), so a matrix
 * multiplied once per lock never pays for it (with SPIRA_REPRODUCIBLE the
 * first call builds it, so that every call runs one variant).  Every variant except XPHASE and XPHASE_PACED adds each row's
 * products in a fixed order, so y is the same bits on every call; the XPHASE
 * variants add into LDS with atomics, whose order (and so the last bits of
 * y) can change between calls.  With SPIRA_REPRODUCIBLE, AUTO keeps to the
 * fixed-order variants (the fastest of them by the autotuner's times, else
 * by the selector) and an explicit XPHASE variant is refused
 * (SPIRA_ERR_INVALID_ARGUMENT).
 * flags: SPIRA_X_DEVICE | SPIRA_Y_DEVICE | SPIRA_NO_SYNC | SPIRA_REPRODUCIBLE. */
int spira_spmv_f64(spira_dcsr *A, const double *x, uint64_t x_len, double *y,
                   uint64_t y_len, int variant, uint32_t flags);

/* ---- SpMV of the float instantiation ----------------------------------------
 * serial::algorithms::spmv / parallel::algorithms::spmv for
 * <soa_tag, uint32_t, float> (serial/spmv.hpp:90-135,
 * parallel/algorithms/spmv.hpp:85-119): y = A*x with x and y FP32 and A's
 * values used in FP32 (a float matrix's values are exact in the handle's
 * FP64 storage).  Checks and flags as spira_spmv_f64.  Each row is summed
 * more accurately than the reference's FP32 loop (compensated FP32 in the VEC
 * and STREAM kernels, FP64 in XPHASE) and rounded once to FP32: normwise
 * relative error <= 1e-6 against the FP64 sum of the same FP32 data.
 * variant: SPIRA_SPMV_AUTO (the FP64 AUTO choice mapped onto these kernels,
 * XPHASE_PACED run as XPHASE),
 * VEC4..VEC32, STREAM1K, STREAM2K, XPHASE, XPHASE_PACED. */
int spira_spmv_f32(spira_dcsr *A, const float *x, uint64_t x_len, float *y, uint64_t y_len, int variant,
                   uint32_t flags);

/* ---- allocation-free SpMV ----------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  spira_spmv_f64 with the same checks, arguments and
 * results, for device x and y only (flags must hold SPIRA_X_DEVICE |
 * SPIRA_Y_DEVICE, and may add SPIRA_NO_SYNC), calling no allocator: A's SpMV
 * plan for the variant must already exist (one spira_spmv_f64 call with the
 * same variant after the last lock or upload builds it; SPIRA_ERR_LOGIC
 * otherwise).  AUTO runs AUTO's choice once its plan exists, else the variant
 * the last AUTO call of spira_spmv_f64 ran (the plan-free one of the first
 * call after a lock).  The scratch its kernels use -- the XPHASE_PACED pacing
 * counters (512 bytes) or the XSLICE/XSTREAM partial sums (8 * n_rows
 * doubles) -- comes from one fixed array of SPIRA_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends.  A variant needing more scratch gives
 * SPIRA_ERR_UNSUPPORTED (XSLICE/XSTREAM above 524,288 rows).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (from several
 * handles or streams) race and must be serialised by the caller. */
#define SPIRA_NOALLOC_SCRATCH_BYTES 33554432u /* 32 MiB */
int spira_spmv_f64_noalloc(spira_dcsr *A, const double *x, uint64_t x_len, double *y, uint64_t y_len, int variant,
                           uint32_t flags);

/* Variant AUTO settles on for A: the one it runs from the second AUTO call
 * after a lock or upload (the first may run a plan-free variant instead). */
int spira_spmv_choose(spira_dcsr *A, int *variant);
/* Pin the AUTO choice for A (autotuner / selector output). */
int spira_spmv_set_variant(spira_dcsr *A, int variant);

/* Exhaustive autotuner (replaces the reference's one-time CPUID kernel pick,
 * dispatch.cpp:34-67, with a per-matrix timed pick): runs every variant `reps`
 * times on device vectors x (n_cols) and y (n_rows), writes the mean time per
 * call in milliseconds to ms[SPIRA_SPMV_N_VARIANTS] (-1 for a variant whose
 * size limits A exceeds), pins the fastest as A's AUTO choice and returns it
 * in *best (either output may be NULL). */
int spira_spmv_autotune_f64(spira_dcsr *A, const double *x_dev, double *y_dev, int reps, float *ms, int *best);

/* The device scan every pass of this library builds on, exposed for callers and
 * tests (no counterpart in the reference): out[0..n) = exclusive prefix sums of
 * in[0..n), out[n] = the total (modulo 2^64); device arrays, in may equal out
 * (and may be NULL when n == 0).  Synchronous. */
int spira_scan_u64(spira_dev *dev, const uint64_t *in_dev, uint64_t *out_dev, uint64_t n);

/* Measured ceilings of this GPU for the SpMV roofline (no counterpart in the
 * reference; a measurement aid for benchmarks).  *gather_req_per_s: random
 * 8-byte loads that hit L2 (every CU gathering inside its XCD's own 2 MiB
 * slice of a 32 MiB array, 20 loads in flight per thread, each load its own
 * L1-to-L2 request), per second.  *stream_bytes_per_s: a read-only sweep of
 * 1 GiB with 16-byte non-temporal loads.  Allocates about 1.03 GiB from the
 * device's cache for the duration of the call; either output may be NULL. */
int spira_probe_ceilings(spira_dev *dev, double *gather_req_per_s, double *stream_bytes_per_s);

/* ---- partitions -------------------------------------------------------------
 * compute_partition_boundaries (partition.hpp:78-126) as used by
 * parallel_matrix::rebalance() (parallel_matrix.hpp:434-497), on the device
 * from A's offsets (the per-row nnz prefix): boundaries[n_threads+1] (host
 * array) receives nnz-balanced row boundaries, equal to the reference's.
 * SPIRA_ERR_INVALID_ARGUMENT for n_threads == 0, SPIRA_ERR_LOGIC if A is not
 * built ("rebalance: matrix must be locked"). */
int spira_partition_boundaries(spira_dcsr *A, uint64_t n_threads, uint64_t *boundaries);

/* ---- transpose ----------------------------------------------------------------
 * Replaces serial::algorithms::transpose (serial/transpose.hpp:35-157): out
 * (a handle of A's shape with rows and columns swapped, in any state; A itself
 * when A is square) receives A^T, built: every row lists its columns ascending,
 * values copied as stored (explicit zeros kept).  With
 * SPIRA_TRANSPOSE_DROP_ZEROS it replaces parallel::algorithms::transpose
 * (parallel/algorithms/transpose.hpp:37-69), whose final lock() removes
 * explicit zeros.  SPIRA_ERR_LOGIC if A is not built ("transpose: matrix must
 * be locked"), SPIRA_ERR_INVALID_ARGUMENT for a wrong shape or device;
 * nnz, rows and columns below 2^32. */
#define SPIRA_TRANSPOSE_DROP_ZEROS 1u
int spira_transpose_f64(spira_dcsr *A, spira_dcsr *out, uint32_t flags);

/* ---- matrix addition -------------------------------------------------------
 * Replaces serial::algorithms::MatrixAddition (serial/matrix_addition.hpp:
 * 88-112, with addRows :34-86), parallel::algorithms::MatrixAddition
 * (parallel/algorithms/matrix_addition.hpp:24-99) and matrix::operator+ /
 * operator+= (matrix_operators.hpp:21-50).  C (a handle of A's shape; may be
 * A or B) receives A + B, built: each row is the column-ordered union of A's
 * and B's row, a + b where both hold the column, the lone value elsewhere,
 * every zero dropped (the output's first lock()).  With SPIRA_ADD_SUBTRACT,
 * A + B * -1: operator- / operator-= (matrix_operators.hpp:29-62 with
 * scalars.hpp:23-35).  Checks in the reference's order: shapes
 * (SPIRA_ERR_INVALID_ARGUMENT, "Matrices aren't the same size"), then A and B
 * built (SPIRA_ERR_LOGIC, "MatrixAddition: A must be locked", "... B ..."). */
#define SPIRA_ADD_SUBTRACT 1u
int spira_add_f64(spira_dcsr *A, spira_dcsr *B, spira_dcsr *C, uint32_t flags);

/* ---- SpGEMM ------------------------------------------------------------------
 * Replaces serial::algorithms::spgemm (serial/spgemm.hpp:12-50),
 * parallel::algorithms::spgemm (parallel/algorithms/spgemm.hpp:24-69) and
 * matrix::operator* / operator*= of two matrices (matrix_operators.hpp:
 * 67-83).  C (A's rows, B's columns; may be A or B) receives A * B, built:
 * for each output entry the products a_ik * b_kj are rounded, then added from
 * 0.0 in the order of k and, inside B's row, of j -- the reference's hash
 * accumulator order -- and entries whose sum is 0 are dropped.  Checks:
 * A.n_cols != B.n_rows (SPIRA_ERR_INVALID_ARGUMENT, "A.cols must equal
 * B.rows"), then A and B built (SPIRA_ERR_LOGIC); fewer than 2^32 - 1 rows,
 * at most 2^32 products in all (a scratch of 12 bytes per product) and fewer
 * than 2^32 in the rows with more than 1024 products or 256 A entries
 * (SPIRA_ERR_UNSUPPORTED). */
int spira_spgemm_f64(spira_dcsr *A, spira_dcsr *B, spira_dcsr *C);

/* ---- mixed precision ------------------------------------------------------
 * No counterpart in the reference (its float instantiation,
 * serial/spmv.hpp:90-135, stores and accumulates in FP32).  y = A*x with A's
 * values read from a reduced-precision copy (built once from the FP64 values,
 * dropped by the next lock): */
#define SPIRA_VAL_F32 1        /* 4 B per value */
#define SPIRA_VAL_BF16 2       /* 2 B per value */
#define SPIRA_VAL_SPLIT_BF16 3 /* hi = bf16(v), lo = bf16(v - hi): 4 B per value */
/* x is rounded to FP32; each row is accumulated in FP32 with compensated
 * (TwoSum / FMA TwoProd, "Dot2") summation and written to y in FP64.  Rows
 * whose cancellation ratio sum|a*x| / |sum a*x| exceeds recompute_ratio (or
 * whose result is not finite) are recomputed in FP64 from the FP64 values and
 * x; recompute_ratio <= 0 disables that pass.  variant: SPIRA_SPMV_AUTO,
 * VEC4..VEC32, STREAM1K, STREAM2K, or (without the recompute) XPHASE and
 * XPHASE_PACED, which read the reduced values in the XPHASE plan's order and
 * sum each row in FP64 (an FP32 value times the FP32 x is exact in FP64; the
 * products are added in LDS).
 * AUTO follows the FP64 AUTO choice, with XPHASE_PACED run as XPHASE (the
 * unpaced kernel is the faster one for reduced values), and row-major kernels
 * when the recompute is on.  *n_recomputed (may be NULL; reading it synchronises) receives the
 * number of recomputed rows.  Checks and flags as spira_spmv_f64. */
int spira_spmv_mixed_f64(spira_dcsr *A, int storage, const double *x, uint64_t x_len, double *y, uint64_t y_len,
                         int variant, float recompute_ratio, uint32_t flags, uint64_t *n_recomputed);

/* ---- batches: many matrices, k right-hand sides, one launch ---------------
 * No counterpart in the reference (it runs one matrix and one vector at a
 * time, serial/spmv.hpp:142-186).  A batch holds locked matrices as one
 * block-diagonal CSR with reduced-precision values (SPIRA_VAL_*): X is the
 * column-wise concatenation of every matrix's n_cols x k block (row major,
 * FP32), Y likewise with n_rows (FP32).  spira_batch_bases gives each matrix's
 * first row and column.  Inside the batch each matrix's rows are stored
 * sorted by length class (rows of 2^(c-1) .. 2^c - 1 nonzeros form class c),
 * columns as u16 offsets inside their matrix when every matrix has at most
 * 65536 columns.  A schedule fixes k (1, 4, 16 or 32) and, per matrix
 * (spira_batch_sched; NULL = built-in choices): lanes per row of class c =
 * 2^c >> plog2 (at least 1, at most 64 / k), the nonzeros per work item (the
 * device queue's unit), and whether the matrix's X block (n_cols * k floats,
 * at most SPIRA_BATCH_XLDS_FLOATS; at k = 1 also up to
 * SPIRA_BATCH_XLDS_TWO_PASS_COLS columns, staged in two halves) is staged in
 * LDS for its gathers, and flags: SPIRA_BATCH_STREAM reads the matrix's
 * entries straight from global memory (items of item_nnz nonzeros, the X block
 * staged once per item) instead of copying each item's entries into LDS beside
 * the X block (items then also end where the LDS is full).
 * spira_batch_spmm runs every item in one persistent launch fed by device
 * work queues, one per XCD: matrices are dealt whole to the queues, so the
 * CUs of an XCD share a matrix's X block in their L2, and a block whose queue
 * is empty takes items from the others.  Each output is accumulated in FP32
 * with compensated (Dot2) summation.  The call can be captured once into a hipGraph (bound to the X
 * and Y given) and replayed.  X and Y are device pointers; flags:
 * SPIRA_NO_SYNC.  Destroy batches before their device. */
#define SPIRA_BATCH_XLDS_FLOATS 36864
#define SPIRA_BATCH_XLDS_TWO_PASS_COLS 65536
#define SPIRA_BATCH_MAX_PLOG2 5
#define SPIRA_BATCH_STREAM 1u
typedef struct spira_batch_sched {
    uint8_t plog2;     /* target nonzeros per lane: 1 << plog2 */
    uint8_t xlds;      /* 1: stage the X block in LDS */
    uint16_t flags;    /* SPIRA_BATCH_STREAM or 0 */
    uint32_t item_nnz; /* work-item size (nonzeros + rows) */
} spira_batch_sched;
typedef struct spira_batch spira_batch;
int spira_batch_create(spira_dev *dev, uint32_t n_mats, spira_dcsr *const *mats, int storage, spira_batch **out);
void spira_batch_destroy(spira_batch *b);
int spira_batch_info(const spira_batch *b, uint64_t *n_rows, uint64_t *n_cols, uint64_t *nnz, uint32_t *n_tiles);
int spira_batch_bases(const spira_batch *b, uint64_t *row_base, uint64_t *col_base); /* n_mats+1 each */
int spira_batch_set_schedule(spira_batch *b, uint32_t k, const spira_batch_sched *per_matrix);
int spira_batch_spmm(spira_batch *b, const float *X, float *Y, uint32_t flags);
/* One spira_batch_spmm launch that also timestamps every work item on the
 * device clock; matrix_cost_us[m] (n_mats entries) receives the sum of the
 * durations of matrix m's items, scaled to CU-microseconds by the blocks
 * resident per CU.  This is the exhaustive autotuner's measurement: one
 * profiled launch per schedule gives every matrix's cost inside its batch. */
int spira_batch_profile(spira_batch *b, const float *X, float *Y, double *matrix_cost_us);
int spira_batch_graph_capture(spira_batch *b, const float *X, float *Y);
int spira_batch_graph_launch(spira_batch *b, uint32_t flags);

/* ---- learned schedule selector --------------------------------------------
 * No counterpart in the reference: it replaces the reference's one-time CPUID
 * pick of a kernel (dispatch.cpp:34-67) and the exhaustive per-matrix search
 * of spira_spmv_autotune_f64 / batch schedules with a small MLP.
 * spira_matrix_features writes SPIRA_FEATURES floats describing a locked
 * matrix (row-length moments and histogram, bandwidth, column spread), computed
 * on the device.  spira_mlp_* is an n_in -> hidden -> hidden -> n_out ReLU
 * network in FP32 on the device: forward, masked mean-squared-error gradient
 * (all parameters, flat: W1, b1, W2, b2, W3, b3; W row-major [out][in]) and an
 * Adam step.  X, T, M, Y are device arrays (batch x n_in / batch x n_out,
 * row major); mask_count is the number of M entries that are > 0. */
#define SPIRA_FEATURES 18
int spira_matrix_features(spira_dcsr *A, float *features_host);
typedef struct spira_mlp spira_mlp;
int spira_mlp_create(spira_dev *dev, uint32_t n_in, uint32_t hidden, uint32_t n_out, uint64_t seed, spira_mlp **out);
void spira_mlp_destroy(spira_mlp *net);
int spira_mlp_params(spira_mlp *net, float *get_host, const float *set_host, uint64_t *n_params);
int spira_mlp_forward(spira_mlp *net, const float *X, uint32_t batch, float *Y);
int spira_mlp_loss_grad(spira_mlp *net, const float *X, const float *T, const float *M, uint32_t batch,
                        float mask_count, float *loss_host, float *grad_host);
int spira_mlp_adam_step(spira_mlp *net, float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif

#endif /* SPIRA_AMD_H */
