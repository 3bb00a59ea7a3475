// C-ABI layer: handle, reference-compatible entry points and the plan API.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <string>

#include "kernels.hpp"
#include "plan.hpp"

namespace muxi {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

// Stream-ordered scratch for the per-call reference entry points (merge
// partitions, carries): freed on the same stream once the kernels are done.
struct AsyncBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    ~AsyncBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
};

static spmv_status_t alloc_async(AsyncBuf& b, size_t bytes, hipStream_t s) {
    b.s = s;
    hipError_t e = hipMallocAsync(&b.p, bytes ? bytes : 16, s);
    if (e != hipSuccess) {
        b.p = nullptr;
        set_error(std::string("hipMallocAsync: ") + hipGetErrorString(e));
        return SPMV_ERROR_MEMORY_ALLOC;
    }
    return SPMV_SUCCESS;
}

static DevCsr dev_of(const muxi_csr_t* A) { return DevCsr{A->numRows, A->numCols, A->nnz, A->d_rowPtr, A->d_colIdx, A->d_values}; }

static spmv_status_t check_csr(const muxi_csr_t* A) {
    if (!A) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (A->numRows < 0 || A->numCols < 0 || A->nnz < 0) {
        set_error("negative matrix dimension");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if ((long)A->numRows + (long)A->nnz > 2147483647L) {
        set_error("numRows + nnz exceeds INT32_MAX (merge-path coordinates are int32)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (A->nnz > 0 && (!A->d_rowPtr || !A->d_colIdx || !A->d_values)) {
        set_error("matrix device arrays are NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    return SPMV_SUCCESS;
}

// Merge-path SpMV for one call: partitions sized for this device, scratch
// from the stream-ordered allocator.
template <typename VT, typename XT, typename AT, bool COMP>
static spmv_status_t merge_once(const muxi_csr_t* A, int ipt, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s) {
    const int L = A->numRows + A->nnz;
    const int P = merge_default_partitions(L, ipt);
    AsyncBuf b1, b2;
    spmv_status_t st;
    if ((st = alloc_async(b1, 2 * sizeof(int) * (size_t)(P + 1), s)) != SPMV_SUCCESS) return st;
    if ((st = alloc_async(b2, merge_carry_bytes(P, TypeOf<XT>::v, COMP), s)) != SPMV_SUCCESS) return st;
    int* pos = (int*)b1.p;
    int* row = pos + (P + 1);
    if ((st = launch_merge_partitions(A->numRows, A->nnz, P, pos, s)) != SPMV_SUCCESS) return st;
    if ((st = launch_merge_coords(A->d_rowPtr, A->numRows, A->nnz, pos, P + 1, row, nullptr, s)) != SPMV_SUCCESS)
        return st;
    return launch_merge<VT, XT, AT, COMP>(dev_of(A), ipt, pos, row, P, b2.p, x, y, alpha, beta, s);
}

template <typename T>
static spmv_status_t spmv_csr_impl(const muxi_csr_t* A, const T* x, T* y, T alpha, T beta, const spmv_opts_t* o) {
    spmv_opts_t opts = o ? *o : muxi_spmv_default_opts();
    if (!A) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (A->nnz == 0) return SPMV_SUCCESS;  // spmv_csr.cu:61-63: y untouched
    spmv_status_t st = check_csr(A);
    if (st != SPMV_SUCCESS) return st;
    if (!x || !y) {
        set_error("x or y is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    hipStream_t s = (hipStream_t)opts.stream;
    if (opts.operation == SPMV_OP_TRANSPOSE || opts.operation == SPMV_OP_CONJUGATE) {
        if ((st = launch_scale_y<T>(y, A->numCols, beta, s)) != SPMV_SUCCESS) return st;
        st = launch_transpose_scalar<T, T>(dev_of(A), x, y, alpha, s);
    } else {
        st = merge_once<T, T, T, false>(A, 8, x, y, alpha, beta, s);
    }
    if (st != SPMV_SUCCESS) return st;
    if (opts.sync) MUXI_HIP(hipStreamSynchronize(s));
    return SPMV_SUCCESS;
}

template <typename T>
static spmv_status_t merge_based_impl(const muxi_csr_t* A, const T* x, T* y, void* stream, int ipt) {
    if (!A) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (A->nnz == 0) return SPMV_SUCCESS;  // spmv_csr5.cu:937-939
    spmv_status_t st = check_csr(A);
    if (st != SPMV_SUCCESS) return st;
    if (!x || !y) {
        set_error("x or y is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    return merge_once<T, T, T, false>(A, ipt, x, y, T(1), T(0), (hipStream_t)stream);
}

template <typename T>
static spmv_status_t sparse_optimized_impl(const muxi_csr_t* A, const T* x, T* y, void* stream) {
    if (A && A->nnz == 0) {  // spmv_sparse_optimized.cu:87-90 zeroes y
        if (A->numRows > 0 && y) MUXI_HIP(hipMemsetAsync(y, 0, sizeof(T) * (size_t)A->numRows, (hipStream_t)stream));
        return SPMV_SUCCESS;
    }
    return merge_based_impl<T>(A, x, y, stream, 8);
}

template <typename T>
static spmv_status_t spmv_csr5_impl(const muxi_csr5_t* M, const T* x, T* y, T alpha, T beta, const spmv_opts_t* o) {
    spmv_opts_t opts = o ? *o : muxi_spmv_default_opts();
    if (!M) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (M->nnz == 0) return SPMV_SUCCESS;  // spmv_csr5.cu:285-287
    if (!M->d_rowPtr || !M->d_colIdx || !M->d_values || !M->d_tile_row_ptr || !M->d_tile_nnz_offset) {
        set_error("CSR5 device arrays are NULL (run muxi_csr5_preprocess first)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (!x || !y) {
        set_error("x or y is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    hipStream_t s = (hipStream_t)opts.stream;
    DevCsr A{M->numRows, M->numCols, M->nnz, M->d_rowPtr, M->d_colIdx, M->d_values};
    spmv_status_t st;
    if (csr5_sigma_supported(M->sigma)) {
        AsyncBuf b;
        if ((st = alloc_async(b, csr5_carry_bytes(M->numTiles, TypeOf<T>::v, false), s)) != SPMV_SUCCESS) return st;
        st = launch_csr5<T, T, T, false>(A, M->sigma, M->d_tile_row_ptr, M->numTiles, b.p, x, y, alpha, beta, s);
    } else {
        // tile sizes the wave kernel has no variant for run on the merge path
        muxi_csr_t c{};
        c.numRows = M->numRows;
        c.numCols = M->numCols;
        c.nnz = M->nnz;
        c.d_rowPtr = M->d_rowPtr;
        c.d_colIdx = M->d_colIdx;
        c.d_values = M->d_values;
        st = merge_once<T, T, T, false>(&c, 8, x, y, alpha, beta, s);
    }
    if (st != SPMV_SUCCESS) return st;
    if (opts.sync) MUXI_HIP(hipStreamSynchronize(s));
    return SPMV_SUCCESS;
}

template <typename T>
static spmv_status_t csr_to_ell_impl(const muxi_csr_t* A, muxi_ell_t* E, int maxElements) {
    if (!A || !E) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    spmv_status_t st = check_csr(A);
    if (st != SPMV_SUCCESS) return st;
    if (A->numRows > 0 && !A->d_rowPtr) {
        set_error("CSR device arrays are NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    hipStream_t s = nullptr;
    int width = maxElements;
    if (width <= 0) {
        int* d_w = nullptr;
        MUXI_HIP(hipMalloc(&d_w, sizeof(int)));
        st = launch_max_row_len(A->d_rowPtr, A->numRows, d_w, s);
        if (st == SPMV_SUCCESS) {
            if (hipMemcpy(&width, d_w, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) st = SPMV_ERROR_MEMORY_COPY;
        }
        (void)hipFree(d_w);
        if (st != SPMV_SUCCESS) return st;
    }
    long slots = (long)A->numRows * (long)width;
    if (slots > 2147483647L) {
        set_error("ELLPACK numRows*width exceeds INT32_MAX");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    E->numRows = A->numRows;
    E->numCols = A->numCols;
    E->numElements = width;
    E->nnz = A->nnz;
    E->d_colIdx = nullptr;
    E->d_values = nullptr;
    if (slots == 0) return SPMV_SUCCESS;
    if (hipMalloc(&E->d_colIdx, sizeof(int) * (size_t)slots) != hipSuccess ||
        hipMalloc(&E->d_values, sizeof(T) * (size_t)slots) != hipSuccess) {
        (void)hipFree(E->d_colIdx);
        E->d_colIdx = nullptr;
        set_error("ELLPACK allocation failed");
        return SPMV_ERROR_MEMORY_ALLOC;
    }
    st = launch_csr_to_ell<T>(dev_of(A), width, E->d_colIdx, (T*)E->d_values, s);
    if (st != SPMV_SUCCESS) return st;
    MUXI_HIP(hipStreamSynchronize(s));
    return SPMV_SUCCESS;
}

template <typename T>
static spmv_status_t spmv_ell_impl(const muxi_ell_t* E, const T* x, T* y, void* stream) {
    if (!E) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (E->numRows == 0) return SPMV_SUCCESS;
    if (E->numElements > 0 && (!E->d_colIdx || !E->d_values)) {
        set_error("ELLPACK device arrays are NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (!x || !y) {
        set_error("x or y is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    return launch_ell<T, T, T, false>(E->numRows, E->numCols, E->numElements, E->d_colIdx, (const T*)E->d_values, x,
                                      y, T(1), T(0), (hipStream_t)stream);
}

template <typename T>
static spmv_status_t exec_impl(muxi_spmv_handle_t* h, spmv_format_t format, const void* matrix, const T* x, T* y,
                               T alpha, T beta, spmv_operation_t op) {
    if (!h) {
        set_error("handle is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    spmv_opts_t o = h->opts;
    o.stream = h->stream;
    o.operation = op;
    if (format == SPMV_FORMAT_CSR) return spmv_csr_impl<T>((const muxi_csr_t*)matrix, x, y, alpha, beta, &o);
    if (format == SPMV_FORMAT_CSR5) {
        if (op != SPMV_OP_NORMAL) {
            set_error("CSR5 transpose is not supported");
            return SPMV_ERROR_UNSUPPORTED_FORMAT;
        }
        return spmv_csr5_impl<T>((const muxi_csr5_t*)matrix, x, y, alpha, beta, &o);
    }
    set_error("format not supported");
    return SPMV_ERROR_UNSUPPORTED_FORMAT;
}

}  // namespace muxi

using namespace muxi;

extern "C" {

spmv_opts_t muxi_spmv_default_opts(void) {
    spmv_opts_t o;
    o.operation = SPMV_OP_NORMAL;
    o.sync = 0;
    o.check_inputs = 0;
    o.use_tensor_core = 0;
    o.use_streaming = 0;
    o.stream_id = 0;
    o.stream = nullptr;
    return o;
}

const char* muxi_spmv_last_error(void) { return get_error(); }
const char* muxi_spmv_version(void) { return "muxi_spmv_amd 0.1.0 gfx950"; }

spmv_status_t muxi_spmv_create_handle(muxi_spmv_handle_t** hp) {
    if (!hp) {
        set_error("handle pointer is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    muxi_spmv_handle_t* h = (muxi_spmv_handle_t*)std::calloc(1, sizeof(muxi_spmv_handle_t));
    if (!h) return SPMV_ERROR_MEMORY_ALLOC;
    if (hipGetDevice(&h->deviceId) != hipSuccess) {
        std::free(h);
        set_error("hipGetDevice failed");
        return SPMV_ERROR_DEVICE;
    }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, h->deviceId) != hipSuccess) {
        std::free(h);
        set_error("hipGetDeviceProperties failed");
        return SPMV_ERROR_DEVICE;
    }
    h->warpSize = p.warpSize;
    h->smCount = p.multiProcessorCount;
    h->maxThreadsPerBlock = p.maxThreadsPerBlock;
    h->sharedMemPerSM = p.maxSharedMemoryPerMultiProcessor;
    h->stream = nullptr;
    h->opts = muxi_spmv_default_opts();
    *hp = h;
    return SPMV_SUCCESS;
}

spmv_status_t muxi_spmv_destroy_handle(muxi_spmv_handle_t* h) {
    if (!h) {
        set_error("handle is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    std::free(h);
    return SPMV_SUCCESS;
}

spmv_status_t muxi_spmv_set_stream(muxi_spmv_handle_t* h, void* stream) {
    if (!h) {
        set_error("handle is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    h->stream = stream;
    h->opts.stream = stream;
    return SPMV_SUCCESS;
}

spmv_status_t muxi_spmv_csr_f32(const muxi_csr_t* A, const float* x, float* y, float alpha, float beta,
                                const spmv_opts_t* opts) {
    return spmv_csr_impl<float>(A, x, y, alpha, beta, opts);
}
spmv_status_t muxi_spmv_csr_f64(const muxi_csr_t* A, const double* x, double* y, double alpha, double beta,
                                const spmv_opts_t* opts) {
    return spmv_csr_impl<double>(A, x, y, alpha, beta, opts);
}

spmv_status_t muxi_spmv_exec_f32(muxi_spmv_handle_t* h, spmv_format_t f, const void* m, const float* x, float* y,
                                 float a, float b) {
    return exec_impl<float>(h, f, m, x, y, a, b, SPMV_OP_NORMAL);
}
spmv_status_t muxi_spmv_exec_f64(muxi_spmv_handle_t* h, spmv_format_t f, const void* m, const double* x, double* y,
                                 double a, double b) {
    return exec_impl<double>(h, f, m, x, y, a, b, SPMV_OP_NORMAL);
}
spmv_status_t muxi_spmv_exec_transpose_f32(muxi_spmv_handle_t* h, spmv_format_t f, const void* m, const float* x,
                                           float* y, float a, float b) {
    return exec_impl<float>(h, f, m, x, y, a, b, SPMV_OP_TRANSPOSE);
}
spmv_status_t muxi_spmv_exec_transpose_f64(muxi_spmv_handle_t* h, spmv_format_t f, const void* m, const double* x,
                                           double* y, double a, double b) {
    return exec_impl<double>(h, f, m, x, y, a, b, SPMV_OP_TRANSPOSE);
}

spmv_status_t muxi_spmv_merge_based_f32(const muxi_csr_t* A, const float* x, float* y, void* s) {
    return merge_based_impl<float>(A, x, y, s, 8);
}
spmv_status_t muxi_spmv_merge_based_f64(const muxi_csr_t* A, const double* x, double* y, void* s) {
    return merge_based_impl<double>(A, x, y, s, 8);
}
spmv_status_t muxi_spmv_merge_based_v2_f32(const muxi_csr_t* A, const float* x, float* y, void* s, bool opt) {
    return merge_based_impl<float>(A, x, y, s, opt ? 4 : 8);
}
spmv_status_t muxi_spmv_merge_based_v2_f64(const muxi_csr_t* A, const double* x, double* y, void* s, bool opt) {
    return merge_based_impl<double>(A, x, y, s, opt ? 4 : 8);
}
spmv_status_t muxi_spmv_sparse_optimized_f32(const muxi_csr_t* A, const float* x, float* y, void* s) {
    return sparse_optimized_impl<float>(A, x, y, s);
}
spmv_status_t muxi_spmv_sparse_optimized_f64(const muxi_csr_t* A, const double* x, double* y, void* s) {
    return sparse_optimized_impl<double>(A, x, y, s);
}

int muxi_merge_partition_count(int numRows, int nnz, int ePP, int mult, int numSMs) {
    // spmv_csr5.cu:942-972 with warpsPerBlock = 256/64
    long L = (long)numRows + (long)nnz;
    long target = ePP > 0 ? L / ePP : 0;
    long maxp = (long)numSMs * (256 / kWave) * mult;
    long p = target < maxp ? target : maxp;
    return (int)(p > 1 ? p : 1);
}

spmv_status_t muxi_merge_partitions(int numRows, int nnz, int P, int* d_pos, void* s) {
    if (P <= 0 || !d_pos) {
        set_error("invalid partition count or NULL output");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    return launch_merge_partitions(numRows, nnz, P, d_pos, (hipStream_t)s);
}

spmv_status_t muxi_merge_path_coords(const int* d_rowPtr, int numRows, int nnz, const int* d_pos, int count,
                                     int* d_rowIdx, int* d_nnzIdx, void* s) {
    if (!d_rowPtr || !d_pos || !d_rowIdx) {
        set_error("NULL argument");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    return launch_merge_coords(d_rowPtr, numRows, nnz, d_pos, count, d_rowIdx, d_nnzIdx, (hipStream_t)s);
}

spmv_status_t muxi_csr5_preprocess(const muxi_csr_t* csr, muxi_csr5_t* c5, int sigma, void* stream) {
    if (!csr || !c5) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (csr->nnz == 0) return SPMV_SUCCESS;  // spmv_csr5.cu:229-231
    spmv_status_t st = check_csr(csr);
    if (st != SPMV_SUCCESS) return st;
    hipStream_t s = (hipStream_t)stream;
    // referenceFromCSR (sparse_formats.h:351-362)
    c5->numRows = csr->numRows;
    c5->numCols = csr->numCols;
    c5->nnz = csr->nnz;
    c5->rowPtr = csr->rowPtr;
    c5->colIdx = csr->colIdx;
    c5->values = csr->values;
    c5->d_rowPtr = csr->d_rowPtr;
    c5->d_colIdx = csr->d_colIdx;
    c5->d_values = csr->d_values;
    c5->ownsCsrData = false;
    if (sigma <= 0) sigma = 256;  // getCSR5Config for a 64-wide wavefront (spmv_csr5.cu:19-21)
    c5->sigma = sigma;
    const int T = (int)(((long)csr->nnz + sigma - 1) / sigma);
    c5->numTiles = T;
    c5->tile_row_ptr = (int*)std::malloc(sizeof(int) * (size_t)T);
    c5->tile_nnz_offset = (int*)std::malloc(sizeof(int) * (size_t)T);
    c5->ownsTileData = true;
    if (!c5->tile_row_ptr || !c5->tile_nnz_offset) return SPMV_ERROR_MEMORY_ALLOC;
    MUXI_HIP(hipMalloc(&c5->d_tile_row_ptr, sizeof(int) * (size_t)T));
    MUXI_HIP(hipMalloc(&c5->d_tile_nnz_offset, sizeof(int) * (size_t)T));
    c5->ownsDeviceMemory = true;
    hipEvent_t e0, e1;
    MUXI_HIP(hipEventCreate(&e0));
    MUXI_HIP(hipEventCreate(&e1));
    MUXI_HIP(hipEventRecord(e0, s));
    st = launch_csr5_tiles(csr->d_rowPtr, csr->numRows, csr->nnz, sigma, T, c5->d_tile_row_ptr,
                           c5->d_tile_nnz_offset, s);
    MUXI_HIP(hipEventRecord(e1, s));
    MUXI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    c5->conversionTimeMs = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != SPMV_SUCCESS) return st;
    MUXI_HIP(hipMemcpy(c5->tile_row_ptr, c5->d_tile_row_ptr, sizeof(int) * (size_t)T, hipMemcpyDeviceToHost));
    MUXI_HIP(hipMemcpy(c5->tile_nnz_offset, c5->d_tile_nnz_offset, sizeof(int) * (size_t)T, hipMemcpyDeviceToHost));
    return SPMV_SUCCESS;
}

spmv_status_t muxi_csr5_destroy(muxi_csr5_t* c5) {
    if (!c5) return SPMV_ERROR_INVALID_MATRIX;
    if (c5->ownsTileData) {
        std::free(c5->tile_row_ptr);
        std::free(c5->tile_nnz_offset);
    }
    if (c5->ownsDeviceMemory) {
        (void)hipFree(c5->d_tile_row_ptr);
        (void)hipFree(c5->d_tile_nnz_offset);
        if (c5->ownsCsrData) {
            (void)hipFree(c5->d_rowPtr);
            (void)hipFree(c5->d_colIdx);
            (void)hipFree(c5->d_values);
        }
    }
    if (c5->ownsCsrData) {
        std::free(c5->rowPtr);
        std::free(c5->colIdx);
        std::free(c5->values);
    }
    c5->tile_row_ptr = c5->tile_nnz_offset = nullptr;
    c5->d_tile_row_ptr = c5->d_tile_nnz_offset = nullptr;
    c5->ownsTileData = c5->ownsDeviceMemory = c5->ownsCsrData = false;
    return SPMV_SUCCESS;
}

spmv_status_t muxi_spmv_csr5_f32(const muxi_csr5_t* A, const float* x, float* y, float a, float b,
                                 const spmv_opts_t* o) {
    return spmv_csr5_impl<float>(A, x, y, a, b, o);
}
spmv_status_t muxi_spmv_csr5_f64(const muxi_csr5_t* A, const double* x, double* y, double a, double b,
                                 const spmv_opts_t* o) {
    return spmv_csr5_impl<double>(A, x, y, a, b, o);
}

spmv_status_t muxi_csr_to_ellpack_f32(const muxi_csr_t* c, muxi_ell_t* e, int m) { return csr_to_ell_impl<float>(c, e, m); }
spmv_status_t muxi_csr_to_ellpack_f64(const muxi_csr_t* c, muxi_ell_t* e, int m) { return csr_to_ell_impl<double>(c, e, m); }
spmv_status_t muxi_spmv_ellpack_f32(const muxi_ell_t* e, const float* x, float* y, void* s) {
    return spmv_ell_impl<float>(e, x, y, s);
}
spmv_status_t muxi_spmv_ellpack_f64(const muxi_ell_t* e, const double* x, double* y, void* s) {
    return spmv_ell_impl<double>(e, x, y, s);
}
spmv_status_t muxi_ellpack_destroy(muxi_ell_t* e) {
    if (!e) return SPMV_ERROR_INVALID_MATRIX;
    (void)hipFree(e->d_colIdx);
    (void)hipFree(e->d_values);
    e->d_colIdx = nullptr;
    e->d_values = nullptr;
    return SPMV_SUCCESS;
}

spmv_status_t muxi_spmv_oracle_f64(const muxi_csr_t* A, spmv_type_t vt, spmv_type_t xt, const void* x, double* y,
                                   void* stream) {
    spmv_status_t st = check_csr(A);
    if (st != SPMV_SUCCESS) return st;
    if (!x || !y) return SPMV_ERROR_INVALID_VECTOR;
    hipStream_t s = (hipStream_t)stream;
    DevCsr d = dev_of(A);
    if (xt == SPMV_TYPE_FLOAT) {
        if (vt == SPMV_TYPE_FLOAT) return launch_oracle_f64<float, float>(d, (const float*)x, y, s);
        if (vt == SPMV_TYPE_DOUBLE) return launch_oracle_f64<double, float>(d, (const float*)x, y, s);
        if (vt == SPMV_TYPE_BF16) return launch_oracle_f64<bf16_t, float>(d, (const float*)x, y, s);
    } else if (xt == SPMV_TYPE_DOUBLE) {
        if (vt == SPMV_TYPE_FLOAT) return launch_oracle_f64<float, double>(d, (const double*)x, y, s);
        if (vt == SPMV_TYPE_DOUBLE) return launch_oracle_f64<double, double>(d, (const double*)x, y, s);
        if (vt == SPMV_TYPE_BF16) return launch_oracle_f64<bf16_t, double>(d, (const double*)x, y, s);
    }
    set_error("oracle: unsupported type pair");
    return SPMV_ERROR_UNSUPPORTED_TYPE;
}

spmv_status_t muxi_convert_values(const void* src, spmv_type_t st, void* dst, spmv_type_t dt, size_t n,
                                  void* stream) {
    if (n == 0) return SPMV_SUCCESS;
    if (!src || !dst) return SPMV_ERROR_INVALID_VECTOR;
    return launch_convert(src, st, dst, dt, n, (hipStream_t)stream);
}

spmv_status_t muxi_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return SPMV_SUCCESS;
    if (!dst || !src) return SPMV_ERROR_INVALID_VECTOR;
    MUXI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    MUXI_HIP(hipStreamSynchronize((hipStream_t)stream));
    return SPMV_SUCCESS;
}

}  // extern "C"
