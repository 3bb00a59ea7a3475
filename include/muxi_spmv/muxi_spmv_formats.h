/*
 * muxi_spmv_formats.h — COO, BSR and CSR2 containers and device-side
 * conversions between them and CSR, plus a deterministic CSR transpose.
 *
 * Replaces the reference's COOMatrix / CSR2Matrix / BSRMatrix structs
 * (src/formats/sparse_formats.h:113-269, same field order) and the format
 * conversions it declares without defining (coo_to_csr, csr_to_coo,
 * bsr_to_csr, convertCSRToCSR5: sparse_formats.h:389-405), and the
 * declared-only spmv_csr_transpose (src/spmv/csr/spmv_csr.cuh:1107-1114).
 *
 * Every conversion reads the input's DEVICE arrays and fills the output's
 * device arrays, which the library allocates (ownsDeviceMemory = true; host
 * pointers NULL); free them with muxi_csr_free_device / muxi_coo_free_device
 * / muxi_bsr_free_device (or hipFree each array). Values are FLOAT, DOUBLE
 * or BF16 (`vtype`) and are moved, never rounded, so every conversion is
 * exact and bit-reproducible: the order of the output is fixed by the rules
 * below, not by thread timing. Each call synchronizes `stream` before it
 * returns (the output sizes are read back). Indices are 0-based.
 *
 *  - CSR entries keep their row-major order; columns within a row are kept in
 *    the order given (a conversion that sorts says so).
 *  - BSR blocks are blockSize x blockSize, stored row-major inside a block
 *    (values[b*bs*bs + i*bs + j] is row i, column j of block b); block columns
 *    are ascending within a block row; the last block row / column is padded
 *    with zeros when numRows / numCols is not a multiple of blockSize.
 */
#ifndef MUXI_SPMV_FORMATS_H_
#define MUXI_SPMV_FORMATS_H_

#include <stdint.h>

#include "muxi_spmv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* COOMatrix<T> (src/formats/sparse_formats.h:113-165). */
typedef struct {
    int numRows;
    int numCols;
    int nnz;
    int* rowIdx;
    int* colIdx;
    void* values;
    int* d_rowIdx;
    int* d_colIdx;
    void* d_values;
    bool ownsHostMemory;
    bool ownsDeviceMemory;
} muxi_coo_t;

/* CSR2Matrix<T> (src/formats/sparse_formats.h:179-208): declared for ABI
 * compatibility only; the reference defines neither a conversion to it nor a
 * kernel for it, and neither does this library. */
typedef struct {
    int numRows;
    int numCols;
    int nnz;
    int* rowPtr;
    void* values;
    uint8_t* metadata;
    int* d_rowPtr;
    void* d_values;
    uint8_t* d_metadata;
    bool ownsHostMemory;
    bool ownsDeviceMemory;
} muxi_csr2_t;

/* BSRMatrix<T, BlockSize> (src/formats/sparse_formats.h:221-267); the block
 * size is a run-time field here (the reference's template parameter). */
typedef struct {
    int numRows;
    int numCols;
    int numBlockRows;
    int numBlockCols;
    int numBlocks;
    int blockSize;
    int* rowPtr;
    int* colIdx;
    void* values;
    int* d_rowPtr;
    int* d_colIdx;
    void* d_values;
    bool ownsHostMemory;
    bool ownsDeviceMemory;
} muxi_bsr_t;

/* csr_to_coo (sparse_formats.h:393-394): rowIdx expanded from rowPtr,
 * colIdx and values copied; same entry order. */
spmv_status_t muxi_csr_to_coo(const muxi_csr_t* A, spmv_type_t vtype, muxi_coo_t* out, void* stream);

/* coo_to_csr (sparse_formats.h:390-391): entries stably sorted by (row,
 * column) (already sorted input is detected and not re-sorted); duplicate
 * (row, column) entries are kept, in input order. An index outside
 * numRows x numCols is SPMV_ERROR_INVALID_MATRIX. */
spmv_status_t muxi_coo_to_csr(const muxi_coo_t* A, spmv_type_t vtype, muxi_csr_t* out, void* stream);

/* CSR -> BSR with blockSize in [1, 64] (the reference has no such conversion;
 * it is the inverse of bsr_to_csr). Duplicate (row, column) entries are
 * SPMV_ERROR_INVALID_MATRIX (they would land on one block slot). */
spmv_status_t muxi_csr_to_bsr(const muxi_csr_t* A, spmv_type_t vtype, int blockSize, muxi_bsr_t* out,
                              void* stream);

/* bsr_to_csr (sparse_formats.h:396-397): every stored block entry inside
 * numRows x numCols becomes a CSR entry, explicit zeros included; within a row
 * the entries follow the block order, then the column inside the block. */
spmv_status_t muxi_bsr_to_csr(const muxi_bsr_t* A, spmv_type_t vtype, muxi_csr_t* out, void* stream);

/* A^T as a CSR matrix (numRows = A.numCols): entries of each output row in
 * ascending original row order, i.e. a stable sort of A's entries by column.
 * A plan built on the result computes y = alpha*A^T*x + beta*y with any
 * kernel, deterministically (spmv_csr_transpose, spmv_csr.cuh:1107-1114). */
spmv_status_t muxi_csr_transpose(const muxi_csr_t* A, spmv_type_t vtype, muxi_csr_t* out, void* stream);

/* Free device arrays that the calls above allocated (ownsDeviceMemory) and
 * clear the pointers; matrices the library does not own are left alone. */
spmv_status_t muxi_csr_free_device(muxi_csr_t* A);
spmv_status_t muxi_coo_free_device(muxi_coo_t* A);
spmv_status_t muxi_bsr_free_device(muxi_bsr_t* A);

#ifdef __cplusplus
}
#endif

#endif /* MUXI_SPMV_FORMATS_H_ */
