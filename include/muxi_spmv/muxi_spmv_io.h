/*
 * muxi_spmv_io.h — Matrix Market (coordinate) files to and from CSR.
 *
 * Replaces the reference's src/generators/mtx_io.h (readMTXHeader,
 * readMTXFile<T>, writeMTXFile<T>); host arrays, like the reference's.
 *
 * Reading follows the reference: the banner must read
 * "%%MatrixMarket matrix coordinate <type> <symmetry>", <type> real, integer
 * or pattern (pattern entries get the value 1), comment lines start with '%',
 * indices are 1-based in the file, "symmetric" files have their off-diagonal
 * entries mirrored when expandSymmetric is set ("skew-symmetric" and
 * "hermitian" are read as stored), and entries are sorted by (row, column).
 * Differences, all toward strictness: complex files are rejected
 * (SPMV_ERROR_UNSUPPORTED_TYPE; the reference misreads them), an index
 * outside the stated dimensions or a short file is an error, and the sort is
 * stable, so duplicate (row, column) entries keep their file order (the
 * reference's std::sort leaves it unspecified; duplicates are kept, not
 * summed, in both).
 * Writing follows writeMTXFile (banner, two comment lines, dimensions,
 * "row col value" lines; writeSymmetric writes the upper triangle only) but
 * prints values with %.9g (FP32) / %.17g (FP64) instead of %g, so that a
 * write followed by a read returns the same values.
 */
#ifndef MUXI_SPMV_IO_H_
#define MUXI_SPMV_IO_H_

#include "muxi_spmv.h"
#include "muxi_spmv_formats.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MTXHeader (src/generators/mtx_io.h:23-31), same field order. */
typedef struct {
    bool isSymmetric;
    bool isComplex;
    bool isPattern;
    bool isInteger;
    int numRows;
    int numCols;
    int nnz; /* entries stored in the file */
} muxi_mtx_header_t;

/* readMTXHeader (src/generators/mtx_io.cu:70-112). */
spmv_status_t muxi_mtx_read_header(const char* filename, muxi_mtx_header_t* header);

/* readMTXFile<float/double> (src/generators/mtx_io.cu:114-207): fills out with
 * host arrays the library allocates (ownsHostMemory = true, device pointers
 * NULL); release them with muxi_csr_free_host. */
spmv_status_t muxi_mtx_read_csr_f32(const char* filename, int expandSymmetric, muxi_csr_t* out);
spmv_status_t muxi_mtx_read_csr_f64(const char* filename, int expandSymmetric, muxi_csr_t* out);

/* writeMTXFile<float/double> (src/generators/mtx_io.cu:209-254) from the
 * matrix's host arrays. */
spmv_status_t muxi_mtx_write_csr_f32(const char* filename, const muxi_csr_t* A, int writeSymmetric);
spmv_status_t muxi_mtx_write_csr_f64(const char* filename, const muxi_csr_t* A, int writeSymmetric);

/* readMTXToCOO<float/double> (src/generators/mtx_io.cu:273-299): the entries
 * as stored in the file (symmetric files are not mirrored), in (row, column)
 * order, as host arrays (ownsHostMemory = true); free with muxi_coo_free_host. */
spmv_status_t muxi_mtx_read_coo_f32(const char* filename, muxi_coo_t* out);
spmv_status_t muxi_mtx_read_coo_f64(const char* filename, muxi_coo_t* out);

/* writeCOOToMTX<float/double> (src/generators/mtx_io.cu:301-324): a "general"
 * file with the host entries in their order; an entry outside the matrix is
 * SPMV_ERROR_INVALID_MATRIX. */
spmv_status_t muxi_mtx_write_coo_f32(const char* filename, const muxi_coo_t* A);
spmv_status_t muxi_mtx_write_coo_f64(const char* filename, const muxi_coo_t* A);

/* Frees host arrays that muxi_mtx_read_* allocated and clears the pointers. */
spmv_status_t muxi_csr_free_host(muxi_csr_t* A);
spmv_status_t muxi_coo_free_host(muxi_coo_t* A);

#ifdef __cplusplus
}
#endif

#endif /* MUXI_SPMV_IO_H_ */
