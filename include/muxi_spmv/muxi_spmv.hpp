// muxi_spmv.hpp — header-only C++ shim that restores the reference's
// `namespace muxi_spmv` names on top of the C-ABI, so reference callers
// compile with `cudaStream_t` replaced by `hipStream_t` and nothing else.
//
//   CSRMatrix<T>      src/formats/sparse_formats.h:24-103 (move-only here)
//   CSR5Matrix<T>     src/formats/sparse_formats.h:284-386
//   ELLPACKMatrix<T>  src/spmv/ellpack/spmv_ellpack.cuh:26-39
//   spmv_csr<T>, spmv_csr_exec<T>, spmv_csr5<T>, csr5_preprocess<T>,
//   spmv_merge_based<T>, spmv_merge_based_v2<T>, spmv_sparse_optimized<T>,
//   csr_to_ellpack<T>, spmv_ellpack<T>, spmv_exec<T>, spmv_exec_transpose<T>,
//   handle functions (src/api/spmv_api.h:33-92)
//   COOMatrix<T>, BSRMatrix<T, B>, coo_to_csr, csr_to_coo, bsr_to_csr,
//   convertCSRToCSR5 (src/formats/sparse_formats.h:113-267, 389-405),
//   spmv_csr_transpose (src/spmv/csr/spmv_csr.cuh:1107-1114)
//   io::MTXHeader, io::readMTXHeader, io::readMTXFile, io::writeMTXFile,
//   io::readMTXToCOO, io::writeCOOToMTX (src/generators/mtx_io.h)
//   benchmark::PerformanceResult, benchmark::BenchmarkConfig, calculateGFlops, calculateBandwidth,
//   measureCSRPerformance<T>, measureCSR5Performance<T>, compareWithCusparse<T> (no counterpart: unsupported)
//   and compareWithRocsparse<T> (the vendor library's SpMV timed the same way),
//   writeResultsJSON, printPerformanceResult (src/benchmark/performance_benchmark.h)
// Instantiated for float and double, as the reference is.
#ifndef MUXI_SPMV_HPP_
#define MUXI_SPMV_HPP_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdlib>
#include <type_traits>

#include <vector>

#include "muxi_spmv.h"
#include "muxi_spmv_bench.h"
#include "muxi_spmv_formats.h"
#include "muxi_spmv_io.h"
#include "muxi_spmv_prep.h"

namespace muxi_spmv {

using ::spmv_format_t;
using ::spmv_operation_t;
using ::spmv_opts_t;
using ::spmv_status_t;
using ::spmv_type_t;
using spmv_handle_t = ::muxi_spmv_handle_t;

inline spmv_opts_t spmv_default_opts() { return ::muxi_spmv_default_opts(); }

template <typename FloatType>
struct CSRMatrix {
    int numRows = 0, numCols = 0, nnz = 0;
    int* rowPtr = nullptr;
    int* colIdx = nullptr;
    FloatType* values = nullptr;
    int* d_rowPtr = nullptr;
    int* d_colIdx = nullptr;
    FloatType* d_values = nullptr;
    bool ownsHostMemory = false;
    bool ownsDeviceMemory = false;

    CSRMatrix() = default;
    CSRMatrix(const CSRMatrix&) = delete;  // the reference's implicit copy double-frees
    CSRMatrix& operator=(const CSRMatrix&) = delete;
    CSRMatrix(CSRMatrix&& o) noexcept { *this = static_cast<CSRMatrix&&>(o); }
    CSRMatrix& operator=(CSRMatrix&& o) noexcept {
        if (this != &o) {
            release();
            numRows = o.numRows; numCols = o.numCols; nnz = o.nnz;
            rowPtr = o.rowPtr; colIdx = o.colIdx; values = o.values;
            d_rowPtr = o.d_rowPtr; d_colIdx = o.d_colIdx; d_values = o.d_values;
            ownsHostMemory = o.ownsHostMemory; ownsDeviceMemory = o.ownsDeviceMemory;
            o.ownsHostMemory = o.ownsDeviceMemory = false;
        }
        return *this;
    }
    ~CSRMatrix() { release(); }

    void release() {
        if (ownsHostMemory) { std::free(rowPtr); std::free(colIdx); std::free(values); }
        if (ownsDeviceMemory) { (void)hipFree(d_rowPtr); (void)hipFree(d_colIdx); (void)hipFree(d_values); }
        ownsHostMemory = ownsDeviceMemory = false;
    }
    void allocateHost(int rows, int cols, int nonzeros) {
        numRows = rows; numCols = cols; nnz = nonzeros;
        rowPtr = (int*)std::malloc((size_t)(numRows + 1) * sizeof(int));
        colIdx = (int*)std::malloc((size_t)nnz * sizeof(int));
        values = (FloatType*)std::malloc((size_t)nnz * sizeof(FloatType));
        ownsHostMemory = true;
    }
    void allocateDevice() {
        (void)hipMalloc(&d_rowPtr, (size_t)(numRows + 1) * sizeof(int));
        (void)hipMalloc(&d_colIdx, (size_t)nnz * sizeof(int));
        (void)hipMalloc(&d_values, (size_t)nnz * sizeof(FloatType));
        ownsDeviceMemory = true;
    }
    void copyToDevice(hipStream_t stream = 0) {
        (void)hipMemcpyAsync(d_rowPtr, rowPtr, (size_t)(numRows + 1) * sizeof(int), hipMemcpyHostToDevice, stream);
        (void)hipMemcpyAsync(d_colIdx, colIdx, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, stream);
        (void)hipMemcpyAsync(d_values, values, (size_t)nnz * sizeof(FloatType), hipMemcpyHostToDevice, stream);
    }
    void copyFromDevice(hipStream_t stream = 0) {
        (void)hipMemcpyAsync(rowPtr, d_rowPtr, (size_t)(numRows + 1) * sizeof(int), hipMemcpyDeviceToHost, stream);
        (void)hipMemcpyAsync(colIdx, d_colIdx, (size_t)nnz * sizeof(int), hipMemcpyDeviceToHost, stream);
        (void)hipMemcpyAsync(values, d_values, (size_t)nnz * sizeof(FloatType), hipMemcpyDeviceToHost, stream);
    }
    // The C-ABI view (same field order).
    ::muxi_csr_t abi() const {
        ::muxi_csr_t c;
        c.numRows = numRows; c.numCols = numCols; c.nnz = nnz;
        c.rowPtr = rowPtr; c.colIdx = colIdx; c.values = values;
        c.d_rowPtr = d_rowPtr; c.d_colIdx = d_colIdx; c.d_values = d_values;
        c.ownsHostMemory = false; c.ownsDeviceMemory = false;
        return c;
    }
};

// The reference's public fields in its order (sparse_formats.h:284-315), laid out as muxi_csr5_t (checked
// below), so reference callers read csr5.numTiles, csr5.tile_row_ptr[t], ... directly.
template <typename FloatType>
struct CSR5Matrix {
    int numRows = 0;
    int numCols = 0;
    int nnz = 0;
    int* rowPtr = nullptr;
    int* colIdx = nullptr;
    FloatType* values = nullptr;
    int sigma = 256;
    int numTiles = 0;
    int* tile_row_ptr = nullptr;     // host, numTiles (filled by csr5_preprocess)
    int* tile_nnz_offset = nullptr;  // host, numTiles
    int* d_rowPtr = nullptr;
    int* d_colIdx = nullptr;
    FloatType* d_values = nullptr;
    int* d_tile_row_ptr = nullptr;
    int* d_tile_nnz_offset = nullptr;
    bool ownsCsrData = false;
    bool ownsTileData = false;
    bool ownsDeviceMemory = false;
    double conversionTimeMs = 0.0;

    CSR5Matrix() = default;
    CSR5Matrix(const CSR5Matrix&) = delete;  // the reference's implicit copy double-frees
    CSR5Matrix& operator=(const CSR5Matrix&) = delete;
    ~CSR5Matrix() { ::muxi_csr5_destroy(abi()); }

    // Reference CSR data from an existing CSRMatrix (no copy), as the reference's referenceFromCSR.
    void referenceFromCSR(const CSRMatrix<FloatType>& csr) {
        numRows = csr.numRows; numCols = csr.numCols; nnz = csr.nnz;
        rowPtr = csr.rowPtr; colIdx = csr.colIdx; values = csr.values;
        d_rowPtr = csr.d_rowPtr; d_colIdx = csr.d_colIdx; d_values = csr.d_values;
        ownsCsrData = false;
    }
    void allocateTileMetadata(int tileCount) {
        numTiles = tileCount;
        tile_row_ptr = (int*)std::malloc((size_t)numTiles * sizeof(int));
        tile_nnz_offset = (int*)std::malloc((size_t)numTiles * sizeof(int));
        ownsTileData = true;
    }
    void allocateDeviceTileMetadata() {
        (void)hipMalloc(&d_tile_row_ptr, (size_t)numTiles * sizeof(int));
        (void)hipMalloc(&d_tile_nnz_offset, (size_t)numTiles * sizeof(int));
        ownsDeviceMemory = true;
    }
    void copyTileMetadataToDevice(hipStream_t stream = 0) {
        (void)hipMemcpyAsync(d_tile_row_ptr, tile_row_ptr, (size_t)numTiles * sizeof(int), hipMemcpyHostToDevice,
                             stream);
        (void)hipMemcpyAsync(d_tile_nnz_offset, tile_nnz_offset, (size_t)numTiles * sizeof(int),
                             hipMemcpyHostToDevice, stream);
    }

    // The C-ABI view: this object itself.
    ::muxi_csr5_t* abi() { return reinterpret_cast<::muxi_csr5_t*>(this); }
    const ::muxi_csr5_t* abi() const { return reinterpret_cast<const ::muxi_csr5_t*>(this); }
};

namespace detail {
template <typename T>
constexpr bool csr5_layout_matches() {
    using M = CSR5Matrix<T>;
    return std::is_standard_layout<M>::value && sizeof(M) == sizeof(::muxi_csr5_t) &&
           offsetof(M, numRows) == offsetof(::muxi_csr5_t, numRows) &&
           offsetof(M, numCols) == offsetof(::muxi_csr5_t, numCols) &&
           offsetof(M, nnz) == offsetof(::muxi_csr5_t, nnz) &&
           offsetof(M, rowPtr) == offsetof(::muxi_csr5_t, rowPtr) &&
           offsetof(M, colIdx) == offsetof(::muxi_csr5_t, colIdx) &&
           offsetof(M, values) == offsetof(::muxi_csr5_t, values) &&
           offsetof(M, sigma) == offsetof(::muxi_csr5_t, sigma) &&
           offsetof(M, numTiles) == offsetof(::muxi_csr5_t, numTiles) &&
           offsetof(M, tile_row_ptr) == offsetof(::muxi_csr5_t, tile_row_ptr) &&
           offsetof(M, tile_nnz_offset) == offsetof(::muxi_csr5_t, tile_nnz_offset) &&
           offsetof(M, d_rowPtr) == offsetof(::muxi_csr5_t, d_rowPtr) &&
           offsetof(M, d_colIdx) == offsetof(::muxi_csr5_t, d_colIdx) &&
           offsetof(M, d_values) == offsetof(::muxi_csr5_t, d_values) &&
           offsetof(M, d_tile_row_ptr) == offsetof(::muxi_csr5_t, d_tile_row_ptr) &&
           offsetof(M, d_tile_nnz_offset) == offsetof(::muxi_csr5_t, d_tile_nnz_offset) &&
           offsetof(M, ownsCsrData) == offsetof(::muxi_csr5_t, ownsCsrData) &&
           offsetof(M, ownsTileData) == offsetof(::muxi_csr5_t, ownsTileData) &&
           offsetof(M, ownsDeviceMemory) == offsetof(::muxi_csr5_t, ownsDeviceMemory) &&
           offsetof(M, conversionTimeMs) == offsetof(::muxi_csr5_t, conversionTimeMs);
}
}  // namespace detail
static_assert(detail::csr5_layout_matches<float>(), "CSR5Matrix<float> must be laid out as muxi_csr5_t");
static_assert(detail::csr5_layout_matches<double>(), "CSR5Matrix<double> must be laid out as muxi_csr5_t");

template <typename FloatType>
struct ELLPACKMatrix {
    int numRows = 0, numCols = 0, numElements = 0, nnz = 0;
    int* d_colIdx = nullptr;
    FloatType* d_values = nullptr;  // the caller frees, as in the reference
};

namespace detail {
template <typename T> struct is_fp : std::integral_constant<bool, std::is_same<T, float>::value || std::is_same<T, double>::value> {};
}

template <typename T>
inline spmv_status_t spmv_csr(const CSRMatrix<T>& A, const T* x, T* y, T alpha, T beta, const spmv_opts_t& opts) {
    static_assert(detail::is_fp<T>::value, "float or double");
    ::muxi_csr_t c = A.abi();
    if constexpr (std::is_same<T, float>::value) return ::muxi_spmv_csr_f32(&c, x, y, alpha, beta, &opts);
    else return ::muxi_spmv_csr_f64(&c, x, y, alpha, beta, &opts);
}

template <typename T>
inline spmv_status_t spmv_csr_exec(const CSRMatrix<T>& A, const T* x, T* y, T alpha = 1.0, T beta = 0.0,
                                   const spmv_opts_t& opts = spmv_default_opts()) {
    return spmv_csr<T>(A, x, y, alpha, beta, opts);
}

template <typename T>
inline spmv_status_t spmv_merge_based(const CSRMatrix<T>& A, const T* x, T* y, hipStream_t stream) {
    ::muxi_csr_t c = A.abi();
    if constexpr (std::is_same<T, float>::value) return ::muxi_spmv_merge_based_f32(&c, x, y, stream);
    else return ::muxi_spmv_merge_based_f64(&c, x, y, stream);
}

template <typename T>
inline spmv_status_t spmv_merge_based_v2(const CSRMatrix<T>& A, const T* x, T* y, hipStream_t stream, bool useOptimized) {
    ::muxi_csr_t c = A.abi();
    if constexpr (std::is_same<T, float>::value) return ::muxi_spmv_merge_based_v2_f32(&c, x, y, stream, useOptimized);
    else return ::muxi_spmv_merge_based_v2_f64(&c, x, y, stream, useOptimized);
}

template <typename T>
inline spmv_status_t spmv_sparse_optimized(const CSRMatrix<T>& A, const T* x, T* y, hipStream_t stream) {
    ::muxi_csr_t c = A.abi();
    if constexpr (std::is_same<T, float>::value) return ::muxi_spmv_sparse_optimized_f32(&c, x, y, stream);
    else return ::muxi_spmv_sparse_optimized_f64(&c, x, y, stream);
}

template <typename T>
inline spmv_status_t csr5_preprocess(const CSRMatrix<T>& csr, CSR5Matrix<T>& csr5, int sigma, hipStream_t stream) {
    ::muxi_csr_t c = csr.abi();
    return ::muxi_csr5_preprocess(&c, csr5.abi(), sigma, stream);
}

template <typename T>
inline spmv_status_t spmv_csr5(const CSR5Matrix<T>& A, const T* x, T* y, T alpha, T beta, const spmv_opts_t& opts) {
    if constexpr (std::is_same<T, float>::value) return ::muxi_spmv_csr5_f32(A.abi(), x, y, alpha, beta, &opts);
    else return ::muxi_spmv_csr5_f64(A.abi(), x, y, alpha, beta, &opts);
}

template <typename T>
inline spmv_status_t csr_to_ellpack(const CSRMatrix<T>& csr, ELLPACKMatrix<T>& ell, int maxElements = 0) {
    ::muxi_csr_t c = csr.abi();
    ::muxi_ell_t e{};
    spmv_status_t st;
    if constexpr (std::is_same<T, float>::value) st = ::muxi_csr_to_ellpack_f32(&c, &e, maxElements);
    else st = ::muxi_csr_to_ellpack_f64(&c, &e, maxElements);
    ell.numRows = e.numRows; ell.numCols = e.numCols; ell.numElements = e.numElements; ell.nnz = e.nnz;
    ell.d_colIdx = e.d_colIdx; ell.d_values = (T*)e.d_values;
    return st;
}

template <typename T>
inline spmv_status_t spmv_ellpack(const ELLPACKMatrix<T>& ell, const T* d_x, T* d_y, hipStream_t stream) {
    ::muxi_ell_t e{ell.numRows, ell.numCols, ell.numElements, ell.nnz, ell.d_colIdx, (void*)ell.d_values};
    if constexpr (std::is_same<T, float>::value) return ::muxi_spmv_ellpack_f32(&e, d_x, d_y, stream);
    else return ::muxi_spmv_ellpack_f64(&e, d_x, d_y, stream);
}

inline spmv_status_t spmv_create_handle(spmv_handle_t** h) { return ::muxi_spmv_create_handle(h); }
inline spmv_status_t spmv_destroy_handle(spmv_handle_t* h) { return ::muxi_spmv_destroy_handle(h); }
inline spmv_status_t spmv_set_stream(spmv_handle_t* h, hipStream_t s) { return ::muxi_spmv_set_stream(h, s); }

template <typename T>
inline spmv_status_t spmv_exec(spmv_handle_t* h, spmv_format_t format, const void* matrix, const T* x, T* y,
                               T alpha = 1.0, T beta = 0.0) {
    // matrix: CSRMatrix<T>* for SPMV_FORMAT_CSR, CSR5Matrix<T>* for SPMV_FORMAT_CSR5
    ::muxi_csr_t c;
    const void* m = matrix;
    if (format == SPMV_FORMAT_CSR) { c = static_cast<const CSRMatrix<T>*>(matrix)->abi(); m = &c; }
    else if (format == SPMV_FORMAT_CSR5) m = static_cast<const CSR5Matrix<T>*>(matrix)->abi();
    if constexpr (std::is_same<T, float>::value) return ::muxi_spmv_exec_f32(h, format, m, x, y, alpha, beta);
    else return ::muxi_spmv_exec_f64(h, format, m, x, y, alpha, beta);
}

template <typename T>
inline spmv_status_t spmv_exec_transpose(spmv_handle_t* h, spmv_format_t format, const void* matrix, const T* x,
                                         T* y, T alpha = 1.0, T beta = 0.0) {
    ::muxi_csr_t c;
    const void* m = matrix;
    if (format == SPMV_FORMAT_CSR) { c = static_cast<const CSRMatrix<T>*>(matrix)->abi(); m = &c; }
    else if (format == SPMV_FORMAT_CSR5) m = static_cast<const CSR5Matrix<T>*>(matrix)->abi();
    if constexpr (std::is_same<T, float>::value) return ::muxi_spmv_exec_transpose_f32(h, format, m, x, y, alpha, beta);
    else return ::muxi_spmv_exec_transpose_f64(h, format, m, x, y, alpha, beta);
}

// ---- COO / BSR and the conversions ------------------------------------------
// The reference declares coo_to_csr / csr_to_coo / bsr_to_csr without a status
// result; here they convert the DEVICE arrays on the GPU (stream 0) and return
// a matrix owning new device arrays. On failure the result is empty
// (numRows = -1) and muxi_spmv_last_error() says why.
template <typename FloatType>
struct COOMatrix {
    int numRows = 0, numCols = 0, nnz = 0;
    int* rowIdx = nullptr;
    int* colIdx = nullptr;
    FloatType* values = nullptr;
    int* d_rowIdx = nullptr;
    int* d_colIdx = nullptr;
    FloatType* d_values = nullptr;
    bool ownsHostMemory = false;
    bool ownsDeviceMemory = false;

    COOMatrix() = default;
    COOMatrix(const COOMatrix&) = delete;
    COOMatrix& operator=(const COOMatrix&) = delete;
    COOMatrix(COOMatrix&& o) noexcept { *this = static_cast<COOMatrix&&>(o); }
    COOMatrix& operator=(COOMatrix&& o) noexcept {
        if (this != &o) {
            release();
            numRows = o.numRows; numCols = o.numCols; nnz = o.nnz;
            rowIdx = o.rowIdx; colIdx = o.colIdx; values = o.values;
            d_rowIdx = o.d_rowIdx; d_colIdx = o.d_colIdx; d_values = o.d_values;
            ownsHostMemory = o.ownsHostMemory; ownsDeviceMemory = o.ownsDeviceMemory;
            o.ownsHostMemory = o.ownsDeviceMemory = false;
        }
        return *this;
    }
    ~COOMatrix() { release(); }
    void release() {
        if (ownsHostMemory) { std::free(rowIdx); std::free(colIdx); std::free(values); }
        if (ownsDeviceMemory) { (void)hipFree(d_rowIdx); (void)hipFree(d_colIdx); (void)hipFree(d_values); }
        ownsHostMemory = ownsDeviceMemory = false;
    }
    void allocateHost(int rows, int cols, int nonzeros) {
        numRows = rows; numCols = cols; nnz = nonzeros;
        rowIdx = (int*)std::malloc((size_t)nnz * sizeof(int));
        colIdx = (int*)std::malloc((size_t)nnz * sizeof(int));
        values = (FloatType*)std::malloc((size_t)nnz * sizeof(FloatType));
        ownsHostMemory = true;
    }
    void allocateDevice() {
        (void)hipMalloc(&d_rowIdx, (size_t)nnz * sizeof(int));
        (void)hipMalloc(&d_colIdx, (size_t)nnz * sizeof(int));
        (void)hipMalloc(&d_values, (size_t)nnz * sizeof(FloatType));
        ownsDeviceMemory = true;
    }
    ::muxi_coo_t abi() const {
        ::muxi_coo_t c;
        c.numRows = numRows; c.numCols = numCols; c.nnz = nnz;
        c.rowIdx = rowIdx; c.colIdx = colIdx; c.values = values;
        c.d_rowIdx = d_rowIdx; c.d_colIdx = d_colIdx; c.d_values = d_values;
        c.ownsHostMemory = false; c.ownsDeviceMemory = false;
        return c;
    }
};

template <typename FloatType, int BlockSize>
struct BSRMatrix {
    int numRows = 0, numCols = 0, numBlockRows = 0, numBlockCols = 0, numBlocks = 0;
    int blockSize = BlockSize;
    int* rowPtr = nullptr;
    int* colIdx = nullptr;
    FloatType* values = nullptr;
    int* d_rowPtr = nullptr;
    int* d_colIdx = nullptr;
    FloatType* d_values = nullptr;
    bool ownsHostMemory = false;
    bool ownsDeviceMemory = false;

    BSRMatrix() = default;
    BSRMatrix(const BSRMatrix&) = delete;
    BSRMatrix& operator=(const BSRMatrix&) = delete;
    BSRMatrix(BSRMatrix&& o) noexcept { *this = static_cast<BSRMatrix&&>(o); }
    BSRMatrix& operator=(BSRMatrix&& o) noexcept {
        if (this != &o) {
            release();
            numRows = o.numRows; numCols = o.numCols; numBlockRows = o.numBlockRows;
            numBlockCols = o.numBlockCols; numBlocks = o.numBlocks; blockSize = o.blockSize;
            rowPtr = o.rowPtr; colIdx = o.colIdx; values = o.values;
            d_rowPtr = o.d_rowPtr; d_colIdx = o.d_colIdx; d_values = o.d_values;
            ownsHostMemory = o.ownsHostMemory; ownsDeviceMemory = o.ownsDeviceMemory;
            o.ownsHostMemory = o.ownsDeviceMemory = false;
        }
        return *this;
    }
    ~BSRMatrix() { release(); }
    void release() {
        if (ownsHostMemory) { std::free(rowPtr); std::free(colIdx); std::free(values); }
        if (ownsDeviceMemory) { (void)hipFree(d_rowPtr); (void)hipFree(d_colIdx); (void)hipFree(d_values); }
        ownsHostMemory = ownsDeviceMemory = false;
    }
    ::muxi_bsr_t abi() const {
        ::muxi_bsr_t c;
        c.numRows = numRows; c.numCols = numCols; c.numBlockRows = numBlockRows; c.numBlockCols = numBlockCols;
        c.numBlocks = numBlocks; c.blockSize = blockSize;
        c.rowPtr = rowPtr; c.colIdx = colIdx; c.values = values;
        c.d_rowPtr = d_rowPtr; c.d_colIdx = d_colIdx; c.d_values = d_values;
        c.ownsHostMemory = false; c.ownsDeviceMemory = false;
        return c;
    }
};

namespace detail {
template <typename T> constexpr spmv_type_t vtype() { return std::is_same<T, float>::value ? SPMV_TYPE_FLOAT : SPMV_TYPE_DOUBLE; }
template <typename T>
inline CSRMatrix<T> csr_from(const ::muxi_csr_t& c, spmv_status_t st) {
    CSRMatrix<T> m;
    if (st != SPMV_SUCCESS) { m.numRows = -1; return m; }
    m.numRows = c.numRows; m.numCols = c.numCols; m.nnz = c.nnz;
    m.d_rowPtr = c.d_rowPtr; m.d_colIdx = c.d_colIdx; m.d_values = (T*)c.d_values;
    m.ownsDeviceMemory = true;
    return m;
}
}  // namespace detail

template <typename T>
inline COOMatrix<T> csr_to_coo(const CSRMatrix<T>& csr) {
    ::muxi_csr_t c = csr.abi();
    ::muxi_coo_t o{};
    COOMatrix<T> m;
    if (::muxi_csr_to_coo(&c, detail::vtype<T>(), &o, nullptr) != SPMV_SUCCESS) { m.numRows = -1; return m; }
    m.numRows = o.numRows; m.numCols = o.numCols; m.nnz = o.nnz;
    m.d_rowIdx = o.d_rowIdx; m.d_colIdx = o.d_colIdx; m.d_values = (T*)o.d_values;
    m.ownsDeviceMemory = true;
    return m;
}

template <typename T>
inline CSRMatrix<T> coo_to_csr(const COOMatrix<T>& coo) {
    ::muxi_coo_t c = coo.abi();
    ::muxi_csr_t o{};
    spmv_status_t st = ::muxi_coo_to_csr(&c, detail::vtype<T>(), &o, nullptr);
    return detail::csr_from<T>(o, st);
}

template <typename T, int B>
inline CSRMatrix<T> bsr_to_csr(const BSRMatrix<T, B>& bsr) {
    ::muxi_bsr_t c = bsr.abi();
    ::muxi_csr_t o{};
    spmv_status_t st = ::muxi_bsr_to_csr(&c, detail::vtype<T>(), &o, nullptr);
    return detail::csr_from<T>(o, st);
}

// The inverse of bsr_to_csr (no counterpart in the reference).
template <typename T, int B>
inline spmv_status_t csr_to_bsr(const CSRMatrix<T>& csr, BSRMatrix<T, B>& bsr, hipStream_t stream = 0) {
    ::muxi_csr_t c = csr.abi();
    ::muxi_bsr_t o{};
    spmv_status_t st = ::muxi_csr_to_bsr(&c, detail::vtype<T>(), B, &o, stream);
    if (st != SPMV_SUCCESS) return st;
    bsr.release();
    bsr.numRows = o.numRows; bsr.numCols = o.numCols; bsr.numBlockRows = o.numBlockRows;
    bsr.numBlockCols = o.numBlockCols; bsr.numBlocks = o.numBlocks; bsr.blockSize = o.blockSize;
    bsr.d_rowPtr = o.d_rowPtr; bsr.d_colIdx = o.d_colIdx; bsr.d_values = (T*)o.d_values;
    bsr.ownsDeviceMemory = true;
    return SPMV_SUCCESS;
}

// A^T as a CSR matrix on the device (deterministic: a stable sort by column).
template <typename T>
inline spmv_status_t csr_transpose(const CSRMatrix<T>& A, CSRMatrix<T>& AT, hipStream_t stream = 0) {
    ::muxi_csr_t c = A.abi();
    ::muxi_csr_t o{};
    spmv_status_t st = ::muxi_csr_transpose(&c, detail::vtype<T>(), &o, stream);
    if (st == SPMV_SUCCESS) AT = detail::csr_from<T>(o, st);
    return st;
}

template <typename T>
inline spmv_status_t convertCSRToCSR5(const CSRMatrix<T>& csr, CSR5Matrix<T>& csr5, int sigma = 256,
                                      hipStream_t stream = 0) {
    return csr5_preprocess<T>(csr, csr5, sigma, stream);
}

// y = alpha*A^T*x + beta*y without atomics: A^T is built, a plan chosen by the
// library's selector runs it, both are released. A caller that repeats the
// product keeps csr_transpose's result and a muxi_plan instead.
template <typename T>
inline spmv_status_t spmv_csr_transpose(const CSRMatrix<T>& A, const T* x, T* y, T alpha, T beta,
                                        const spmv_opts_t& opts) {
    hipStream_t s = (hipStream_t)opts.stream;
    CSRMatrix<T> AT;
    spmv_status_t st = csr_transpose<T>(A, AT, s);
    if (st != SPMV_SUCCESS) return st;
    ::muxi_csr_t c = AT.abi();
    ::muxi_plan* p = nullptr;
    st = ::muxi_plan_create(&c, detail::vtype<T>(), detail::vtype<T>(), MUXI_KERNEL_AUTO, 0, 0, s, &p);
    if (st != SPMV_SUCCESS) return st;
    st = ::muxi_plan_exec(p, x, y, (double)alpha, (double)beta, 1, s);
    if (st == SPMV_SUCCESS) st = hipStreamSynchronize(s) == hipSuccess ? SPMV_SUCCESS : SPMV_ERROR_DEVICE;
    ::muxi_plan_destroy(p);
    return st;
}

// ---- Matrix Market files (host arrays) -----------------------------------------
namespace io {
using MTXHeader = ::muxi_mtx_header_t;

inline spmv_status_t readMTXHeader(const char* filename, MTXHeader& header) {
    return ::muxi_mtx_read_header(filename, &header);
}

template <typename T>
inline spmv_status_t readMTXFile(const char* filename, CSRMatrix<T>& matrix, bool expandSymmetric = true) {
    ::muxi_csr_t c{};
    spmv_status_t st = std::is_same<T, float>::value ? ::muxi_mtx_read_csr_f32(filename, expandSymmetric, &c)
                                                     : ::muxi_mtx_read_csr_f64(filename, expandSymmetric, &c);
    if (st != SPMV_SUCCESS) return st;
    matrix.release();
    matrix.numRows = c.numRows; matrix.numCols = c.numCols; matrix.nnz = c.nnz;
    matrix.rowPtr = c.rowPtr; matrix.colIdx = c.colIdx; matrix.values = (T*)c.values;
    matrix.ownsHostMemory = true;  // malloc'd by the library, released with std::free
    return SPMV_SUCCESS;
}

template <typename T>
inline spmv_status_t writeMTXFile(const char* filename, const CSRMatrix<T>& matrix, bool writeSymmetric = false) {
    ::muxi_csr_t c = matrix.abi();
    return std::is_same<T, float>::value ? ::muxi_mtx_write_csr_f32(filename, &c, writeSymmetric)
                                         : ::muxi_mtx_write_csr_f64(filename, &c, writeSymmetric);
}

template <typename T>
inline spmv_status_t readMTXToCOO(const char* filename, COOMatrix<T>& matrix) {
    ::muxi_coo_t c{};
    spmv_status_t st = std::is_same<T, float>::value ? ::muxi_mtx_read_coo_f32(filename, &c)
                                                     : ::muxi_mtx_read_coo_f64(filename, &c);
    if (st != SPMV_SUCCESS) return st;
    matrix.release();
    matrix.numRows = c.numRows; matrix.numCols = c.numCols; matrix.nnz = c.nnz;
    matrix.rowIdx = c.rowIdx; matrix.colIdx = c.colIdx; matrix.values = (T*)c.values;
    matrix.ownsHostMemory = true;
    return SPMV_SUCCESS;
}

template <typename T>
inline spmv_status_t writeCOOToMTX(const char* filename, const COOMatrix<T>& matrix) {
    ::muxi_coo_t c = matrix.abi();
    return std::is_same<T, float>::value ? ::muxi_mtx_write_coo_f32(filename, &c)
                                         : ::muxi_mtx_write_coo_f64(filename, &c);
}
}  // namespace io

// ---- the benchmark layer (src/benchmark/performance_benchmark.h), over muxi_spmv_bench.h
namespace benchmark {

using PerformanceResult = ::muxi_perf_result_t;  // the reference's fields in its order

// BenchmarkConfig with the reference's defaults (performance_benchmark.h:58-71), laid out as muxi_bench_config_t
struct BenchmarkConfig {
    int warmupIterations = 3;
    int measureIterations = 10;
    bool checkCorrectness = true;
    bool syncAfterEach = true;
    hipStream_t stream = 0;
    spmv_type_t floatType = SPMV_TYPE_FLOAT;
    bool compareCusparse = false;

    const ::muxi_bench_config_t* abi() const { return reinterpret_cast<const ::muxi_bench_config_t*>(this); }
};
static_assert(std::is_standard_layout<BenchmarkConfig>::value && sizeof(BenchmarkConfig) == sizeof(::muxi_bench_config_t) &&
                  offsetof(BenchmarkConfig, warmupIterations) == offsetof(::muxi_bench_config_t, warmupIterations) &&
                  offsetof(BenchmarkConfig, measureIterations) == offsetof(::muxi_bench_config_t, measureIterations) &&
                  offsetof(BenchmarkConfig, checkCorrectness) == offsetof(::muxi_bench_config_t, checkCorrectness) &&
                  offsetof(BenchmarkConfig, syncAfterEach) == offsetof(::muxi_bench_config_t, syncAfterEach) &&
                  offsetof(BenchmarkConfig, stream) == offsetof(::muxi_bench_config_t, stream) &&
                  offsetof(BenchmarkConfig, floatType) == offsetof(::muxi_bench_config_t, floatType) &&
                  offsetof(BenchmarkConfig, compareCusparse) == offsetof(::muxi_bench_config_t, compareCusparse),
              "BenchmarkConfig must be laid out as muxi_bench_config_t");

inline double calculateGFlops(int nnz, double timeMs) { return ::muxi_calculate_gflops(nnz, timeMs); }
inline double calculateBandwidth(int nnz, int numRows, int numCols, double timeMs, int floatBytes = 4) {
    return ::muxi_calculate_bandwidth(nnz, numRows, numCols, timeMs, floatBytes);
}

namespace detail {
template <typename T>
inline BenchmarkConfig typed(const BenchmarkConfig& c) {
    BenchmarkConfig t = c;
    t.floatType = std::is_same<T, double>::value ? SPMV_TYPE_DOUBLE : SPMV_TYPE_FLOAT;
    return t;
}
}  // namespace detail

template <typename FloatType>
inline spmv_status_t measureCSRPerformance(const CSRMatrix<FloatType>& matrix, const FloatType* d_x, FloatType* d_y,
                                           const BenchmarkConfig& config, spmv_handle_t* handle,
                                           PerformanceResult& result) {
    static_assert(muxi_spmv::detail::is_fp<FloatType>::value, "float or double");
    const ::muxi_csr_t c = matrix.abi();
    const BenchmarkConfig t = detail::typed<FloatType>(config);
    return ::muxi_measure_csr_performance(&c, d_x, d_y, t.abi(), handle, &result);
}

template <typename FloatType>
inline spmv_status_t measureCSR5Performance(CSRMatrix<FloatType>& csr, CSR5Matrix<FloatType>& csr5,
                                            const FloatType* d_x, FloatType* d_y, const BenchmarkConfig& config,
                                            spmv_handle_t* handle, PerformanceResult& result, int sigma) {
    static_assert(muxi_spmv::detail::is_fp<FloatType>::value, "float or double");
    const ::muxi_csr_t c = csr.abi();
    const BenchmarkConfig t = detail::typed<FloatType>(config);
    return ::muxi_measure_csr5_performance(&c, csr5.abi(), d_x, d_y, t.abi(), handle, &result, sigma);
}

// The reference compiles its cuSPARSE comparison out (ENABLE_CUSPARSE is never defined) and returns this status.
template <typename FloatType>
inline spmv_status_t compareWithCusparse(const CSRMatrix<FloatType>&, const FloatType*, FloatType*,
                                         const BenchmarkConfig&, double&, double&) {
    return SPMV_ERROR_UNSUPPORTED_FORMAT;
}

// The vendor library's CSR SpMV timed the same way (muxi_compare_with_rocsparse; alg 0 adaptive, 1 rowsplit,
// 2 LRB, 3 nnzsplit), in the place of compareWithCusparse's outputs; *analysisMs, when given, receives the
// library's preprocess time.
template <typename FloatType>
inline spmv_status_t compareWithRocsparse(const CSRMatrix<FloatType>& matrix, const FloatType* d_x, FloatType* d_y,
                                          const BenchmarkConfig& config, double& rocsparseTime,
                                          double& rocsparseGflops, int alg = 0, double* analysisMs = nullptr) {
    static_assert(muxi_spmv::detail::is_fp<FloatType>::value, "float or double");
    const ::muxi_csr_t c = matrix.abi();
    const BenchmarkConfig t = detail::typed<FloatType>(config);
    return ::muxi_compare_with_rocsparse(&c, d_x, d_y, t.abi(), alg, &rocsparseTime, &rocsparseGflops, analysisMs);
}

inline void printPerformanceResult(const PerformanceResult& result) { ::muxi_print_performance_result(&result); }

inline spmv_status_t writeResultsJSON(const std::vector<PerformanceResult>& results, const char* filename) {
    return ::muxi_write_results_json(results.data(), (int)results.size(), filename);
}

}  // namespace benchmark

}  // namespace muxi_spmv

#endif  // MUXI_SPMV_HPP_
