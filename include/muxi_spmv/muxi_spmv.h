/*
 * muxi_spmv.h — C-ABI of the MI355X (gfx950) SpMV/SpMM library.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes and
 * returns an `spmv_status_t`. Device pointers are HIP device addresses;
 * streams are `hipStream_t` passed as `void*` (NULL = the null stream).
 *
 * The structs below keep the field names and order of the reference's
 * containers so that a caller holding one can hand it over unchanged:
 *   muxi_csr_t    <-> CSRMatrix<T>      src/formats/sparse_formats.h:24-103
 *   muxi_csr5_t   <-> CSR5Matrix<T>     src/formats/sparse_formats.h:284-386
 *   muxi_ell_t    <-> ELLPACKMatrix<T>  src/spmv/ellpack/spmv_ellpack.cuh:26-39
 *   spmv_opts_t   <-> spmv_opts_t       src/utils/common.h:73-94
 * Value arrays are `void*`; the element type is carried by the entry point's
 * suffix (_f32 / _f64) or, in the plan API, by a `spmv_type_t` argument.
 */
#ifndef MUXI_SPMV_H_
#define MUXI_SPMV_H_

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums: same names and values as src/utils/common.h:32-71 ---------- */
typedef enum {
    SPMV_SUCCESS = 0,
    SPMV_ERROR_INVALID_HANDLE = -1,
    SPMV_ERROR_INVALID_MATRIX = -2,
    SPMV_ERROR_INVALID_VECTOR = -3,
    SPMV_ERROR_MEMORY_ALLOC = -4,
    SPMV_ERROR_MEMORY_COPY = -5,
    SPMV_ERROR_DEVICE = -6,
    SPMV_ERROR_UNSUPPORTED_FORMAT = -7,
    SPMV_ERROR_UNSUPPORTED_TYPE = -8,
    SPMV_ERROR_INTERNAL = -9
} spmv_status_t;

typedef enum {
    SPMV_FORMAT_CSR = 0,
    SPMV_FORMAT_COO = 1,
    SPMV_FORMAT_CSR2 = 2,
    SPMV_FORMAT_BSR = 3,
    SPMV_FORMAT_CSR5 = 4,
    SPMV_FORMAT_UNKNOWN = -1
} spmv_format_t;

typedef enum {
    SPMV_TYPE_FLOAT = 0,
    SPMV_TYPE_DOUBLE = 1,
    SPMV_TYPE_HALF = 2,
    SPMV_TYPE_BF16 = 3,
    SPMV_TYPE_TF32 = 4,
    SPMV_TYPE_UNKNOWN = -1
} spmv_type_t;

typedef enum {
    SPMV_OP_NORMAL = 0,
    SPMV_OP_TRANSPOSE = 1,
    SPMV_OP_CONJUGATE = 2
} spmv_operation_t;

/* Same field order as src/utils/common.h:73-81. `stream` is the hipStream_t
 * itself (NULL = null stream). */
typedef struct {
    spmv_operation_t operation;
    int sync;
    int check_inputs;
    int use_tensor_core;
    int use_streaming;
    int stream_id;
    void* stream;
} spmv_opts_t;

/* ---- containers --------------------------------------------------------- */
typedef struct {
    int numRows;
    int numCols;
    int nnz;
    int* rowPtr;
    int* colIdx;
    void* values;
    int* d_rowPtr;
    int* d_colIdx;
    void* d_values;
    bool ownsHostMemory;
    bool ownsDeviceMemory;
} muxi_csr_t;

typedef struct {
    int numRows;
    int numCols;
    int nnz;
    int* rowPtr;
    int* colIdx;
    void* values;
    int sigma;
    int numTiles;
    int* tile_row_ptr;
    int* tile_nnz_offset;
    int* d_rowPtr;
    int* d_colIdx;
    void* d_values;
    int* d_tile_row_ptr;
    int* d_tile_nnz_offset;
    bool ownsCsrData;
    bool ownsTileData;
    bool ownsDeviceMemory;
    double conversionTimeMs;
} muxi_csr5_t;

typedef struct {
    int numRows;
    int numCols;
    int numElements;
    int nnz;
    int* d_colIdx;
    void* d_values;
} muxi_ell_t;

/* Handle: same fields as spmv_handle_t (src/api/spmv_api.h:18-28). */
typedef struct {
    int deviceId;
    void* stream;
    spmv_opts_t opts;
    int warpSize;
    int smCount;
    int maxThreadsPerBlock;
    size_t sharedMemPerSM;
} muxi_spmv_handle_t;

/* ---- options / handle (src/utils/common.h:84-94, src/api/spmv_api.cu:11-56) */
spmv_opts_t muxi_spmv_default_opts(void);
spmv_status_t muxi_spmv_create_handle(muxi_spmv_handle_t** handle);
spmv_status_t muxi_spmv_destroy_handle(muxi_spmv_handle_t* handle);
spmv_status_t muxi_spmv_set_stream(muxi_spmv_handle_t* handle, void* stream);
/* Last error text of this thread ("" when none). */
const char* muxi_spmv_last_error(void);
/* Library version string. */
const char* muxi_spmv_version(void);

/* ---- dispatching SpMV: spmv_csr<T> (src/spmv/csr/spmv_csr.cu:52-249) ------
 * y = alpha*A*x + beta*y (the reference ignores alpha/beta and overwrites y;
 * with alpha=1, beta=0 both agree). opts may be NULL (= default options).
 * opts->operation == SPMV_OP_TRANSPOSE computes y = alpha*A^T*x + beta*y. */
spmv_status_t muxi_spmv_csr_f32(const muxi_csr_t* A, const float* x, float* y,
                                float alpha, float beta, const spmv_opts_t* opts);
spmv_status_t muxi_spmv_csr_f64(const muxi_csr_t* A, const double* x, double* y,
                                double alpha, double beta, const spmv_opts_t* opts);
/* spmv_exec / spmv_exec_transpose (src/api/spmv_api.h:57-78; declared only
 * in the reference). matrix is a muxi_csr_t* (CSR) or muxi_csr5_t* (CSR5). */
spmv_status_t muxi_spmv_exec_f32(muxi_spmv_handle_t* h, spmv_format_t format, const void* matrix,
                                 const float* x, float* y, float alpha, float beta);
spmv_status_t muxi_spmv_exec_f64(muxi_spmv_handle_t* h, spmv_format_t format, const void* matrix,
                                 const double* x, double* y, double alpha, double beta);
spmv_status_t muxi_spmv_exec_transpose_f32(muxi_spmv_handle_t* h, spmv_format_t format,
                                           const void* matrix, const float* x, float* y,
                                           float alpha, float beta);
spmv_status_t muxi_spmv_exec_transpose_f64(muxi_spmv_handle_t* h, spmv_format_t format,
                                           const void* matrix, const double* x, double* y,
                                           double alpha, double beta);

/* ---- merge-path: spmv_merge_based / _v2 (src/spmv/csr5/spmv_csr5.cu:930-1072)
 * y = A*x, y overwritten. The partition count is chosen for this device. */
spmv_status_t muxi_spmv_merge_based_f32(const muxi_csr_t* A, const float* x, float* y, void* stream);
spmv_status_t muxi_spmv_merge_based_f64(const muxi_csr_t* A, const double* x, double* y, void* stream);
spmv_status_t muxi_spmv_merge_based_v2_f32(const muxi_csr_t* A, const float* x, float* y,
                                           void* stream, bool useOptimized);
spmv_status_t muxi_spmv_merge_based_v2_f64(const muxi_csr_t* A, const double* x, double* y,
                                           void* stream, bool useOptimized);
/* No counterpart in the reference: muxi_spmv_merge_based_f32 for callers that cannot allocate (code under
 * stream capture, serving loops, harnesses that forbid hipMalloc). Same kernels and result bit for bit; its
 * scratch (partition diagonals, row coordinates, carries) is a static 2 MiB device array per GPU, so
 * numRows + nnz may be at most about 357 million (SPMV_ERROR_INVALID_MATRIX beyond, with the limit in
 * muxi_spmv_last_error). All calls on one GPU share that array: calls that overlap on a GPU (two streams
 * or two threads) race; calls on one stream are ordered. */
spmv_status_t muxi_spmv_merge_based_noalloc_f32(const muxi_csr_t* A, const float* x, float* y, void* stream);
/* spmv_sparse_optimized (src/spmv/csr/spmv_sparse_optimized.cu:80-121): zeroes y when nnz==0. */
spmv_status_t muxi_spmv_sparse_optimized_f32(const muxi_csr_t* A, const float* x, float* y, void* stream);
spmv_status_t muxi_spmv_sparse_optimized_f64(const muxi_csr_t* A, const double* x, double* y, void* stream);

/* Merge-path plan pieces, bit-exact with the reference:
 *  partitions: pos[t] = min(t*ceil(L/P), L), t = 0..P, L = numRows+nnz
 *              (compute_merge_partitions_kernel, spmv_csr5.cu:912-925)
 *  coords:     for each diagonal k = pos[t], the unique r with
 *              rowPtr[r]+r <= k <= rowPtr[r+1]+r (r = numRows at k = L);
 *              rowIdx[t] = r, nnzIdx[t] = k - r (merge_path_search, :497-533)
 *  count:      P = max(1, min((R+N)/ePP, nSM*(256/64)*mult)) (:942-972, :1031-1038). */
int muxi_merge_partition_count(int numRows, int nnz, int elementsPerPartition,
                               int maxPartitionMultiplier, int numSMs);
spmv_status_t muxi_merge_partitions(int numRows, int nnz, int numPartitions, int* d_pos, void* stream);
spmv_status_t muxi_merge_path_coords(const int* d_rowPtr, int numRows, int nnz, const int* d_pos,
                                     int count, int* d_rowIdx, int* d_nnzIdx, void* stream);

/* ---- CSR5 (src/spmv/csr5/spmv_csr5.cu:223-362) ---------------------------
 * preprocess: sigma <= 0 selects 256; tile_row_ptr[t] = last r in [0,R) with
 * rowPtr[r] <= t*sigma, tile_nnz_offset[t] = t*sigma - rowPtr[row]. The CSR
 * arrays are aliased (ownsCsrData = false). Host tile arrays are filled too. */
spmv_status_t muxi_csr5_preprocess(const muxi_csr_t* csr, muxi_csr5_t* csr5, int sigma, void* stream);
spmv_status_t muxi_csr5_destroy(muxi_csr5_t* csr5);
spmv_status_t muxi_spmv_csr5_f32(const muxi_csr5_t* A, const float* x, float* y,
                                 float alpha, float beta, const spmv_opts_t* opts);
spmv_status_t muxi_spmv_csr5_f64(const muxi_csr5_t* A, const double* x, double* y,
                                 double alpha, double beta, const spmv_opts_t* opts);

/* ---- ELLPACK (src/spmv/ellpack/spmv_ellpack.cu:11-91) --------------------
 * Built on the device from the CSR device arrays: width = max row length when
 * maxElements <= 0; slot (row, i) at row + i*numRows; padding col=-1, val=0;
 * rows longer than the width are truncated. Fails with
 * SPMV_ERROR_INVALID_MATRIX when numRows*width exceeds INT32_MAX (the
 * reference overflows an int there). */
spmv_status_t muxi_csr_to_ellpack_f32(const muxi_csr_t* csr, muxi_ell_t* ell, int maxElements);
spmv_status_t muxi_csr_to_ellpack_f64(const muxi_csr_t* csr, muxi_ell_t* ell, int maxElements);
spmv_status_t muxi_spmv_ellpack_f32(const muxi_ell_t* ell, const float* d_x, float* d_y, void* stream);
spmv_status_t muxi_spmv_ellpack_f64(const muxi_ell_t* ell, const double* d_x, double* d_y, void* stream);
spmv_status_t muxi_ellpack_destroy(muxi_ell_t* ell);

/* ---- explicit kernels and plans ------------------------------------------
 * A plan binds one CSR matrix (device arrays), one kernel and its tile
 * parameters, and owns every scratch buffer the kernel needs, so that
 * muxi_plan_exec allocates nothing and may be captured in a hipGraph. */
typedef enum {
    MUXI_KERNEL_AUTO = 0,      /* dispatcher's choice */
    MUXI_KERNEL_SCALAR = 1,    /* one lane per row */
    MUXI_KERNEL_VECTOR = 2,    /* one wavefront per row */
    MUXI_KERNEL_VWAVE = 3,     /* param0 lanes per row (1..64), rowPtr staged in LDS */
    MUXI_KERNEL_ADAPTIVE = 4,  /* rows binned by length, each bin with its own lanes per row */
    MUXI_KERNEL_MERGE = 5,     /* merge-path, param0 items per thread (4,8,16) */
    MUXI_KERNEL_CSR5 = 6,      /* nnz tiles of param0 = sigma, one wavefront per tile */
    MUXI_KERNEL_ELL = 7,       /* ELLPACK built at plan time */
    MUXI_KERNEL_XSLAB = 8,     /* column slabs, slab = workgroup mod slabs (one x slice per XCD L2);
                                  param0 = (slabs << 8) | items per lane; slabs 0 = the fewest of
                                  1, 2, 4, 8 that keep a slab of x within 2 MiB; items 0 = 4 */
    MUXI_KERNEL_LSLAB = 9,     /* column slabs with the slab of x in LDS, pair sums combined per row
                                  block in LDS; param0 = nonzeros per lane (4 or 8; 0 = 4) | 0x100
                                  (slabs of 16384 columns); otherwise slabs of 32768 columns with 4
                                  nonzeros a lane and plain sums, else the widest of 28672, 24576 and
                                  16384 that fits the LDS (8 a lane, compensated or checked sums);
                                  FP32 x and y; at most 512 slabs (16,777,216 columns with 32768) */
    MUXI_KERNEL_COUNT = 10
} muxi_kernel_t;

/* Accumulation: 0 = plain, 1 = compensated (two-sum) accumulation. */
typedef struct muxi_plan muxi_plan;

/* SpMM plan: Y = alpha*A*X + beta*Y with X (numCols x nrhs) and Y (numRows x
 * nrhs) row-major, nrhs in [1, 64], on merge-path partitions; ipt = items
 * per lane group (4, 8, 16 or 32; <= 0 takes it from the schedule table the
 * project's autotuner fitted, by nrhs and numRows + nnz;
 * muxi_spmv_amd/sched_tune.py). Executed by muxi_plan_exec
 * with the same nrhs. Value types FLOAT/BF16/DOUBLE with x FLOAT, or
 * DOUBLE with x DOUBLE; plain accumulation. */
spmv_status_t muxi_plan_create_spmm(const muxi_csr_t* A, spmv_type_t vtype, spmv_type_t xtype, int nrhs, int ipt,
                                    void* stream, muxi_plan** plan);

/* Creates a plan. vtype is the element type of A.d_values (FLOAT, DOUBLE or
 * BF16), xtype the type of x and y (FLOAT or DOUBLE). With xtype FLOAT the
 * accumulator is FP32 (plain or compensated); with xtype DOUBLE it is FP64. */
spmv_status_t muxi_plan_create(const muxi_csr_t* A, spmv_type_t vtype, spmv_type_t xtype,
                               muxi_kernel_t kernel, int param0, int compensated,
                               void* stream, muxi_plan** plan);
/* muxi_plan_create for a row block: A holds rows [row_base, row_base +
 * A->numRows) of a matrix of rows_full rows, numbered from 0, with that
 * matrix's columns (one GPU's row shard of a multi-GPU split, or a chunk of
 * the host path). AUTO then measures the block's band against the whole
 * matrix's diagonal, and an ELL plan takes its 16-bit column offsets from the
 * whole-matrix row, as they would for the whole matrix. row_base = 0 and
 * rows_full = A->numRows is muxi_plan_create. No counterpart in the
 * reference, which has no row shards; for callers that split a matrix by
 * rows. */
spmv_status_t muxi_plan_create_rows(const muxi_csr_t* A, spmv_type_t vtype, spmv_type_t xtype,
                                    muxi_kernel_t kernel, int param0, int compensated, void* stream,
                                    long long row_base, long long rows_full, muxi_plan** plan);
/* Checked plan (no counterpart in the reference, whose FP64 path is the whole
 * spmv_csr<double> dispatch, src/spmv/csr/spmv_csr.cu:164-249): FP32 x and y,
 * values FLOAT/BF16/DOUBLE, FP32 (compensated = 0) or Dot2 (1) accumulation on
 * the LDS column-slab kernels (at most 512 slabs: 14,680,064 columns, 8,388,608
 * with Dot2). Each
 * exec also sums |a_rj x_j| per row beside y and tests the row's a-posteriori
 * error bound -- (n_r+2)u S_r for FP32 sums, 2u|y_r| + 2((n_r+2)u)^2 S_r for
 * Dot2, u = 2^-24, S_r the sum of |products|, slightly inflated for the
 * rounding of S_r itself -- against rtol * |y_r|; rows that fail are
 * recomputed in FP64 (Dot2 in double) by one more kernel of the same exec
 * (one workgroup per failing row) and their y written from that. rtol in
 * (0, 1). For callers that need each row of y within a stated relative
 * error. */
spmv_status_t muxi_plan_create_checked(const muxi_csr_t* A, spmv_type_t vtype, int compensated, double rtol,
                                       void* stream, muxi_plan** plan);
/* The rows a checked plan recomputed in FP64 in its last exec on `stream`
 * (waits for it): *count of them and up to cap row ids, in no fixed order. */
spmv_status_t muxi_plan_checked_rows(const muxi_plan* plan, int* count, int* rows, int cap, void* stream);
spmv_status_t muxi_plan_destroy(muxi_plan* plan);
/* y = alpha*A*x + beta*y with nrhs right-hand sides: for nrhs > 1, X is
 * numCols x nrhs and Y numRows x nrhs, both row-major. */
spmv_status_t muxi_plan_exec(const muxi_plan* plan, const void* x, void* y, double alpha,
                             double beta, int nrhs, void* stream);
/* muxi_plan_exec that also records ev_before on `stream` just before the
 * plan's main kernel and ev_after just after it (before any carry fix-up);
 * either event may be NULL. Events come from muxi_event_create. */
spmv_status_t muxi_plan_exec_ev(const muxi_plan* plan, const void* x, void* y, double alpha, double beta,
                                int nrhs, void* stream, void* ev_before, void* ev_after);
/* HIP event helpers for the timing paths above. */
spmv_status_t muxi_event_create(void** ev);
spmv_status_t muxi_event_destroy(void* ev);
spmv_status_t muxi_event_record(void* ev, void* stream);
/* Waits for `end` and returns the time between the two events in ms. */
spmv_status_t muxi_event_elapsed_ms(void* start, void* end, float* ms);
/* The kernel the plan runs (resolved from AUTO) and its parameter. */
spmv_status_t muxi_plan_info(const muxi_plan* plan, int* kernel, int* param0, size_t* scratch_bytes);

/* FP64 oracle kernel on the device: y64 = A*x with x given as FP32 or FP64
 * (xtype) and values as FLOAT/DOUBLE/BF16 (vtype); one lane per row, FP64
 * accumulation with compensated summation. */
spmv_status_t muxi_spmv_oracle_f64(const muxi_csr_t* A, spmv_type_t vtype, spmv_type_t xtype,
                                   const void* x, double* y64, void* stream);

/* Value conversion on the device: FP32 -> BF16 (round to nearest even),
 * FP32/FP64 <-> FP64/FP32. */
spmv_status_t muxi_convert_values(const void* src, spmv_type_t stype, void* dst, spmv_type_t dtype,
                                  size_t n, void* stream);

/* Device-to-device copy of `bytes` bytes on `stream`, then a stream sync
 * (used to read back library-owned arrays such as the ELLPACK slots). */
spmv_status_t muxi_device_copy(void* dst, const void* src, size_t bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MUXI_SPMV_H_ */
