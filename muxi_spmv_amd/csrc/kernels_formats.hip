// Format conversions on the device: CSR <-> COO, CSR <-> BSR and the CSR
// transpose. All of them are index work plus value moves, so they are exact;
// their output order is fixed by stable sorts (rocPRIM's LSD radix sort) and
// closed-form offsets, never by atomics, so repeated calls give identical
// bytes. Reference: the declarations in src/formats/sparse_formats.h:389-405
// and src/spmv/csr/spmv_csr.cuh:1107-1114 (no definitions there).
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "common.hpp"
#include "muxi_spmv/muxi_spmv_formats.h"

namespace muxi {
namespace {

// Temporary device memory released when it leaves scope.
struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T* as() const {
        return (T*)p;
    }
};

spmv_status_t dalloc(DevBuf& b, size_t bytes) {
    MUXI_HIP(hipMalloc(&b.p, bytes > 0 ? bytes : 1));
    return SPMV_SUCCESS;
}

#define FMT_TRY(expr)                        \
    do {                                     \
        spmv_status_t st_ = (expr);          \
        if (st_ != SPMV_SUCCESS) return st_; \
    } while (0)

int elem_bytes(spmv_type_t t) {
    switch (t) {
        case SPMV_TYPE_FLOAT: return 4;
        case SPMV_TYPE_DOUBLE: return 8;
        case SPMV_TYPE_BF16: return 2;
        default: return 0;
    }
}

// Bits that hold every key in [0, max_key].
unsigned key_bits(unsigned long long max_key) {
    unsigned b = 1;
    while (b < 64 && (max_key >> b) != 0) ++b;
    return b;
}

unsigned grid_cap(long work, int per_block) {
    long g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    return (unsigned)(g > 65536 ? 65536 : g);
}

// ---- validation --------------------------------------------------------------
// flag bit 1: rowPtr not monotone / wrong ends; bit 2: index out of range;
// bit 4: (row, column) order descends (COO only); bit 8: duplicate BSR slot.
__global__ void check_rowptr_kernel(const int* __restrict__ rp, int R, int N, int* flag) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r <= R; r += (long)gridDim.x * blockDim.x) {
        const int a = rp[r];
        const bool bad = (r == 0 && a != 0) || (r == R && a != N) || (r < R && rp[r + 1] < a);
        if (bad) atomicOr(flag, 1);
    }
}

__global__ void check_range_kernel(const int* __restrict__ a, long n, int hi, int* flag) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const int v = a[k];
        if (v < 0 || v >= hi) atomicOr(flag, 2);
    }
}

__global__ void check_coo_kernel(const int* __restrict__ ri, const int* __restrict__ ci, long n, int R, int C,
                                 int* flag) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const int r = ri[k], c = ci[k];
        if (r < 0 || r >= R || c < 0 || c >= C) atomicOr(flag, 2);
        if (k > 0) {
            const int pr = ri[k - 1], pc = ci[k - 1];
            if (r < pr || (r == pr && c < pc)) atomicOr(flag, 4);
        }
    }
}

spmv_status_t read_flag(const int* d_flag, hipStream_t s, int* h) {
    MUXI_HIP(hipMemcpyAsync(h, d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    MUXI_HIP(hipStreamSynchronize(s));
    return SPMV_SUCCESS;
}

// rowPtr[0..R] and columns of a CSR matrix are well formed.
spmv_status_t validate_csr(const muxi_csr_t* A, hipStream_t s) {
    if (A->numRows < 0 || A->numCols < 0 || A->nnz < 0 || (A->numRows > 0 && !A->d_rowPtr) ||
        (A->nnz > 0 && (!A->d_colIdx || !A->d_values))) {
        set_error("invalid CSR matrix (dimensions or device arrays)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (A->numRows == 0) {
        if (A->nnz != 0) {
            set_error("CSR matrix with no rows has nonzeros");
            return SPMV_ERROR_INVALID_MATRIX;
        }
        return SPMV_SUCCESS;
    }
    DevBuf f;
    FMT_TRY(dalloc(f, sizeof(int)));
    MUXI_HIP(hipMemsetAsync(f.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(check_rowptr_kernel, dim3(grid_cap(A->numRows + 1L, 256)), dim3(256), 0, s, A->d_rowPtr,
                       A->numRows, A->nnz, f.as<int>());
    MUXI_LAUNCH_CHECK();
    if (A->nnz > 0) {
        hipLaunchKernelGGL(check_range_kernel, dim3(grid_cap(A->nnz, 256)), dim3(256), 0, s, A->d_colIdx,
                           (long)A->nnz, A->numCols, f.as<int>());
        MUXI_LAUNCH_CHECK();
    }
    int h = 0;
    FMT_TRY(read_flag(f.as<int>(), s, &h));
    if (h & 1) {
        set_error("CSR rowPtr is not a monotone offset array from 0 to nnz");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (h & 2) {
        set_error("CSR column index outside [0, numCols)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    return SPMV_SUCCESS;
}

// ---- building blocks ---------------------------------------------------------
__global__ void iota_kernel(int* a, long n) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        a[k] = (int)k;
}

// Row of each nonzero: the largest r with rp[r] <= k (skips empty rows).
__global__ void rows_of_kernel(const int* __restrict__ rp, int R, long N, int* __restrict__ row) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        int lo = 0, hi = R;  // rp[lo] <= k < rp[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rp[mid] <= k) lo = mid;
            else hi = mid;
        }
        row[k] = lo;
    }
}

// out[j] = in[perm[j]] for elements of E bytes.
template <typename E>
__global__ void gather_kernel(const E* __restrict__ in, const int* __restrict__ perm, long n, E* __restrict__ out) {
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x)
        out[j] = in[perm[j]];
}

spmv_status_t gather_bytes(const void* in, const int* perm, long n, int eb, void* out, hipStream_t s) {
    if (n <= 0) return SPMV_SUCCESS;
    const dim3 g(grid_cap(n, 256)), b(256);
    if (eb == 2) hipLaunchKernelGGL(gather_kernel<uint16_t>, g, b, 0, s, (const uint16_t*)in, perm, n, (uint16_t*)out);
    else if (eb == 4) hipLaunchKernelGGL(gather_kernel<uint32_t>, g, b, 0, s, (const uint32_t*)in, perm, n, (uint32_t*)out);
    else hipLaunchKernelGGL(gather_kernel<uint64_t>, g, b, 0, s, (const uint64_t*)in, perm, n, (uint64_t*)out);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

// rp[r] = first j with key[j] >= r * stride, r in [0, R] (keys ascending).
template <typename K>
__global__ void offsets_from_sorted_kernel(const K* __restrict__ key, long n, int R, K stride, int* __restrict__ rp) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r <= R; r += (long)gridDim.x * blockDim.x) {
        const K t = (K)r * stride;
        long lo = 0, hi = n;  // first index in [lo, hi] with key >= t
        while (lo < hi) {
            const long mid = (lo + hi) >> 1;
            if (key[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        rp[r] = (int)lo;
    }
}

template <typename K>
spmv_status_t offsets_from_sorted(const K* key, long n, int R, K stride, int* rp, hipStream_t s) {
    hipLaunchKernelGGL(offsets_from_sorted_kernel<K>, dim3(grid_cap(R + 1L, 256)), dim3(256), 0, s, key, n, R, stride,
                       rp);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template <typename K>
spmv_status_t sort_pairs(const K* kin, K* kout, const int* vin, int* vout, long n, unsigned bits, hipStream_t s) {
    size_t tb = 0;
    MUXI_HIP(rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, (size_t)n, 0u, bits, s));
    DevBuf t;
    FMT_TRY(dalloc(t, tb));
    MUXI_HIP(rocprim::radix_sort_pairs(t.p, tb, kin, kout, vin, vout, (size_t)n, 0u, bits, s));
    return SPMV_SUCCESS;
}

template <typename T>
spmv_status_t exclusive_scan(const T* in, T* out, long n, hipStream_t s) {
    size_t tb = 0;
    MUXI_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, (T)0, (size_t)n, rocprim::plus<T>(), s));
    DevBuf t;
    FMT_TRY(dalloc(t, tb));
    MUXI_HIP(rocprim::exclusive_scan(t.p, tb, in, out, (T)0, (size_t)n, rocprim::plus<T>(), s));
    return SPMV_SUCCESS;
}

// Output arrays owned by a conversion until it succeeds.
struct OutArrays {
    void* p[3] = {nullptr, nullptr, nullptr};
    bool keep = false;
    ~OutArrays() {
        if (!keep)
            for (void* q : p)
                if (q) (void)hipFree(q);
    }
};

spmv_status_t alloc_out(OutArrays& o, int i, size_t bytes) {
    MUXI_HIP(hipMalloc(&o.p[i], bytes > 0 ? bytes : 1));
    return SPMV_SUCCESS;
}

// ---- CSR -> BSR pieces -------------------------------------------------------
__global__ void block_keys_kernel(const int* __restrict__ row, const int* __restrict__ ci, long N, int b,
                                  unsigned long long nbc, unsigned long long* __restrict__ key) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x)
        key[k] = (unsigned long long)(row[k] / b) * nbc + (unsigned long long)(ci[k] / b);
}

__global__ void block_cols_kernel(const unsigned long long* __restrict__ ukey, long nb, unsigned long long nbc,
                                  int* __restrict__ bci) {
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < nb; j += (long)gridDim.x * blockDim.x)
        bci[j] = (int)(ukey[j] % nbc);
}

// Each nonzero claims its slot in its block (the first claim wins; a second
// claim is a duplicate entry), then the owner writes the value.
template <typename E>
__global__ void bsr_scatter_kernel(const int* __restrict__ row, const int* __restrict__ ci, const E* __restrict__ v,
                                   const unsigned long long* __restrict__ key, long N, const unsigned long long* __restrict__ ukey,
                                   long nb, int b, int* __restrict__ owner, E* __restrict__ bv, int* flag) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const unsigned long long t = key[k];
        long lo = 0, hi = nb - 1;  // t is present in ukey
        while (lo < hi) {
            const long mid = (lo + hi) >> 1;
            if (ukey[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        const long slot = lo * (long)b * b + (long)(row[k] % b) * b + (ci[k] % b);
        if (atomicCAS(&owner[slot], -1, (int)k) != -1) atomicOr(flag, 8);
        else bv[slot] = v[k];
    }
}

// ---- BSR -> CSR pieces -------------------------------------------------------
// Width of each block's column range inside numCols.
__global__ void block_width_kernel(const int* __restrict__ bci, long nb, int b, int C, long long* __restrict__ w) {
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j <= nb; j += (long)gridDim.x * blockDim.x) {
        if (j == nb) {
            w[j] = 0;
        } else {
            const long c0 = (long)bci[j] * b;
            const long rem = C - c0;
            w[j] = rem < b ? rem : b;
        }
    }
}

// Entries of each block row: (its row length) x (its rows inside numRows).
__global__ void blockrow_total_kernel(const int* __restrict__ brp, const long long* __restrict__ G, int nbr, int b, int R,
                                      long long* __restrict__ T) {
    for (long br = blockIdx.x * (long)blockDim.x + threadIdx.x; br <= nbr; br += (long)gridDim.x * blockDim.x) {
        if (br == nbr) {
            T[br] = 0;
        } else {
            const long long L = G[brp[br + 1]] - G[brp[br]];
            const long rem = R - br * (long)b;
            T[br] = L * (rem < b ? rem : b);
        }
    }
}

__global__ void bsr_rowptr_kernel(const int* __restrict__ brp, const long long* __restrict__ G,
                                  const long long* __restrict__ S, int nbr, int b, int R, int* __restrict__ rp) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r <= R; r += (long)gridDim.x * blockDim.x) {
        if (r == R) {
            rp[r] = (int)S[nbr];
        } else {
            const int br = (int)(r / b), i = (int)(r % b);
            const long long L = G[brp[br + 1]] - G[brp[br]];
            rp[r] = (int)(S[br] + (long long)i * L);
        }
    }
}

template <typename E>
__global__ void bsr_fill_kernel(const int* __restrict__ brp, const int* __restrict__ bci, const E* __restrict__ bv,
                                const long long* __restrict__ G, const long long* __restrict__ S, int nbr, long nb, int b,
                                int R, int C, int* __restrict__ ci, E* __restrict__ v) {
    const long bb = (long)b * b;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nb * bb; t += (long)gridDim.x * blockDim.x) {
        const long j = t / bb;
        const int e = (int)(t - j * bb), i = e / b, jj = e - i * b;
        int lo = 0, hi = nbr;  // block row: brp[lo] <= j < brp[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (brp[mid] <= j) lo = mid;
            else hi = mid;
        }
        const long r = (long)lo * b + i, c = (long)bci[j] * b + jj;
        if (r >= R || c >= C) continue;
        const long long L = G[brp[lo + 1]] - G[brp[lo]];
        const long long pos = S[lo] + i * L + (G[j] - G[brp[lo]]) + jj;
        ci[pos] = (int)c;
        v[pos] = bv[t];
    }
}

}  // namespace
}  // namespace muxi

using namespace muxi;

extern "C" spmv_status_t muxi_csr_to_coo(const muxi_csr_t* A, spmv_type_t vt, muxi_coo_t* out, void* stream) {
    if (!A || !out) return SPMV_ERROR_INVALID_MATRIX;
    const int eb = elem_bytes(vt);
    if (!eb) {
        set_error("csr_to_coo: values must be FLOAT, DOUBLE or BF16");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    hipStream_t s = (hipStream_t)stream;
    FMT_TRY(validate_csr(A, s));
    const long N = A->nnz;
    OutArrays o;
    FMT_TRY(alloc_out(o, 0, sizeof(int) * N));
    FMT_TRY(alloc_out(o, 1, sizeof(int) * N));
    FMT_TRY(alloc_out(o, 2, (size_t)eb * N));
    if (N > 0) {
        hipLaunchKernelGGL(rows_of_kernel, dim3(grid_cap(N, 256)), dim3(256), 0, s, A->d_rowPtr, A->numRows, N,
                           (int*)o.p[0]);
        MUXI_LAUNCH_CHECK();
        MUXI_HIP(hipMemcpyAsync(o.p[1], A->d_colIdx, sizeof(int) * N, hipMemcpyDeviceToDevice, s));
        MUXI_HIP(hipMemcpyAsync(o.p[2], A->d_values, (size_t)eb * N, hipMemcpyDeviceToDevice, s));
    }
    MUXI_HIP(hipStreamSynchronize(s));
    std::memset(out, 0, sizeof(*out));
    out->numRows = A->numRows;
    out->numCols = A->numCols;
    out->nnz = (int)N;
    out->d_rowIdx = (int*)o.p[0];
    out->d_colIdx = (int*)o.p[1];
    out->d_values = o.p[2];
    out->ownsDeviceMemory = true;
    o.keep = true;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_coo_to_csr(const muxi_coo_t* A, spmv_type_t vt, muxi_csr_t* out, void* stream) {
    if (!A || !out) return SPMV_ERROR_INVALID_MATRIX;
    const int eb = elem_bytes(vt);
    if (!eb) {
        set_error("coo_to_csr: values must be FLOAT, DOUBLE or BF16");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    if (A->numRows < 0 || A->numCols < 0 || A->nnz < 0 ||
        (A->nnz > 0 && (!A->d_rowIdx || !A->d_colIdx || !A->d_values))) {
        set_error("invalid COO matrix (dimensions or device arrays)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    hipStream_t s = (hipStream_t)stream;
    const long N = A->nnz;
    const int R = A->numRows;
    int sorted = 1;
    if (N > 0) {
        DevBuf f;
        FMT_TRY(dalloc(f, sizeof(int)));
        MUXI_HIP(hipMemsetAsync(f.p, 0, sizeof(int), s));
        hipLaunchKernelGGL(check_coo_kernel, dim3(grid_cap(N, 256)), dim3(256), 0, s, A->d_rowIdx, A->d_colIdx, N, R,
                           A->numCols, f.as<int>());
        MUXI_LAUNCH_CHECK();
        int h = 0;
        FMT_TRY(read_flag(f.as<int>(), s, &h));
        if (h & 2) {
            set_error("COO index outside numRows x numCols");
            return SPMV_ERROR_INVALID_MATRIX;
        }
        sorted = !(h & 4);
    }
    OutArrays o;
    FMT_TRY(alloc_out(o, 0, sizeof(int) * ((size_t)R + 1)));
    FMT_TRY(alloc_out(o, 1, sizeof(int) * N));
    FMT_TRY(alloc_out(o, 2, (size_t)eb * N));
    DevBuf key, skey, idx, sidx, srow;
    const int* rows = A->d_rowIdx;
    if (N > 0 && sorted) {
        MUXI_HIP(hipMemcpyAsync(o.p[1], A->d_colIdx, sizeof(int) * N, hipMemcpyDeviceToDevice, s));
        MUXI_HIP(hipMemcpyAsync(o.p[2], A->d_values, (size_t)eb * N, hipMemcpyDeviceToDevice, s));
    } else if (N > 0) {
        using K = unsigned long long;
        FMT_TRY(dalloc(key, sizeof(K) * N));
        FMT_TRY(dalloc(skey, sizeof(K) * N));
        FMT_TRY(dalloc(idx, sizeof(int) * N));
        FMT_TRY(dalloc(sidx, sizeof(int) * N));
        FMT_TRY(dalloc(srow, sizeof(int) * N));
        hipLaunchKernelGGL(block_keys_kernel, dim3(grid_cap(N, 256)), dim3(256), 0, s, A->d_rowIdx, A->d_colIdx, N, 1,
                           (K)A->numCols, key.as<K>());  // key = row * numCols + col
        MUXI_LAUNCH_CHECK();
        hipLaunchKernelGGL(iota_kernel, dim3(grid_cap(N, 256)), dim3(256), 0, s, idx.as<int>(), N);
        MUXI_LAUNCH_CHECK();
        const unsigned bits = key_bits((K)R * (K)A->numCols - 1);
        FMT_TRY(sort_pairs<K>(key.as<K>(), skey.as<K>(), idx.as<int>(), sidx.as<int>(), N, bits, s));
        FMT_TRY(gather_bytes(A->d_colIdx, sidx.as<int>(), N, 4, o.p[1], s));
        FMT_TRY(gather_bytes(A->d_values, sidx.as<int>(), N, eb, o.p[2], s));
        FMT_TRY(gather_bytes(A->d_rowIdx, sidx.as<int>(), N, 4, srow.p, s));
        rows = srow.as<int>();
    }
    FMT_TRY(offsets_from_sorted<int>(rows, N, R, 1, (int*)o.p[0], s));
    MUXI_HIP(hipStreamSynchronize(s));
    std::memset(out, 0, sizeof(*out));
    out->numRows = R;
    out->numCols = A->numCols;
    out->nnz = (int)N;
    out->d_rowPtr = (int*)o.p[0];
    out->d_colIdx = (int*)o.p[1];
    out->d_values = o.p[2];
    out->ownsDeviceMemory = true;
    o.keep = true;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_csr_transpose(const muxi_csr_t* A, spmv_type_t vt, muxi_csr_t* out, void* stream) {
    if (!A || !out) return SPMV_ERROR_INVALID_MATRIX;
    const int eb = elem_bytes(vt);
    if (!eb) {
        set_error("csr_transpose: values must be FLOAT, DOUBLE or BF16");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    hipStream_t s = (hipStream_t)stream;
    FMT_TRY(validate_csr(A, s));
    const long N = A->nnz;
    const int C = A->numCols;
    OutArrays o;
    FMT_TRY(alloc_out(o, 0, sizeof(int) * ((size_t)C + 1)));
    FMT_TRY(alloc_out(o, 1, sizeof(int) * N));
    FMT_TRY(alloc_out(o, 2, (size_t)eb * N));
    DevBuf skey, idx, sidx, row;
    FMT_TRY(dalloc(skey, sizeof(unsigned) * N));
    if (N > 0) {
        FMT_TRY(dalloc(idx, sizeof(int) * N));
        FMT_TRY(dalloc(sidx, sizeof(int) * N));
        FMT_TRY(dalloc(row, sizeof(int) * N));
        hipLaunchKernelGGL(iota_kernel, dim3(grid_cap(N, 256)), dim3(256), 0, s, idx.as<int>(), N);
        MUXI_LAUNCH_CHECK();
        // columns are in [0, C): as unsigned keys they sort in column order
        FMT_TRY(sort_pairs<unsigned>((const unsigned*)A->d_colIdx, skey.as<unsigned>(), idx.as<int>(), sidx.as<int>(),
                                     N, key_bits(C > 0 ? C - 1 : 0), s));
        hipLaunchKernelGGL(rows_of_kernel, dim3(grid_cap(N, 256)), dim3(256), 0, s, A->d_rowPtr, A->numRows, N,
                           row.as<int>());
        MUXI_LAUNCH_CHECK();
        FMT_TRY(gather_bytes(row.p, sidx.as<int>(), N, 4, o.p[1], s));
        FMT_TRY(gather_bytes(A->d_values, sidx.as<int>(), N, eb, o.p[2], s));
    }
    FMT_TRY(offsets_from_sorted<unsigned>(skey.as<unsigned>(), N, C, 1u, (int*)o.p[0], s));
    MUXI_HIP(hipStreamSynchronize(s));
    std::memset(out, 0, sizeof(*out));
    out->numRows = C;
    out->numCols = A->numRows;
    out->nnz = (int)N;
    out->d_rowPtr = (int*)o.p[0];
    out->d_colIdx = (int*)o.p[1];
    out->d_values = o.p[2];
    out->ownsDeviceMemory = true;
    o.keep = true;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_csr_to_bsr(const muxi_csr_t* A, spmv_type_t vt, int b, muxi_bsr_t* out,
                                         void* stream) {
    if (!A || !out) return SPMV_ERROR_INVALID_MATRIX;
    const int eb = elem_bytes(vt);
    if (!eb) {
        set_error("csr_to_bsr: values must be FLOAT, DOUBLE or BF16");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    if (b < 1 || b > 64) {
        set_error("csr_to_bsr: blockSize must be in [1, 64]");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    hipStream_t s = (hipStream_t)stream;
    FMT_TRY(validate_csr(A, s));
    using K = unsigned long long;
    const long N = A->nnz;
    const int R = A->numRows, C = A->numCols;
    const int nbr = (int)(((long)R + b - 1) / b), nbc = (int)(((long)C + b - 1) / b);
    DevBuf row, key, skey, ukey, cnt;
    long nb = 0;
    if (N > 0) {
        FMT_TRY(dalloc(row, sizeof(int) * N));
        FMT_TRY(dalloc(key, sizeof(K) * N));
        FMT_TRY(dalloc(skey, sizeof(K) * N));
        FMT_TRY(dalloc(ukey, sizeof(K) * N));
        FMT_TRY(dalloc(cnt, sizeof(size_t)));
        hipLaunchKernelGGL(rows_of_kernel, dim3(grid_cap(N, 256)), dim3(256), 0, s, A->d_rowPtr, R, N, row.as<int>());
        MUXI_LAUNCH_CHECK();
        hipLaunchKernelGGL(block_keys_kernel, dim3(grid_cap(N, 256)), dim3(256), 0, s, row.as<int>(), A->d_colIdx, N, b,
                           (K)nbc, key.as<K>());
        MUXI_LAUNCH_CHECK();
        const unsigned bits = key_bits((K)nbr * (K)nbc - 1);
        size_t tb = 0;
        MUXI_HIP(rocprim::radix_sort_keys(nullptr, tb, key.as<K>(), skey.as<K>(), (size_t)N, 0u, bits, s));
        {
            DevBuf t;
            FMT_TRY(dalloc(t, tb));
            MUXI_HIP(rocprim::radix_sort_keys(t.p, tb, key.as<K>(), skey.as<K>(), (size_t)N, 0u, bits, s));
        }
        tb = 0;
        MUXI_HIP(rocprim::unique(nullptr, tb, skey.as<K>(), ukey.as<K>(), cnt.as<size_t>(), (size_t)N,
                                 rocprim::equal_to<K>(), s));
        {
            DevBuf t;
            FMT_TRY(dalloc(t, tb));
            MUXI_HIP(rocprim::unique(t.p, tb, skey.as<K>(), ukey.as<K>(), cnt.as<size_t>(), (size_t)N,
                                     rocprim::equal_to<K>(), s));
        }
        size_t h = 0;
        MUXI_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(size_t), hipMemcpyDeviceToHost, s));
        MUXI_HIP(hipStreamSynchronize(s));
        nb = (long)h;
    }
    const size_t slots = (size_t)nb * b * b;
    OutArrays o;
    FMT_TRY(alloc_out(o, 0, sizeof(int) * ((size_t)nbr + 1)));
    FMT_TRY(alloc_out(o, 1, sizeof(int) * nb));
    FMT_TRY(alloc_out(o, 2, (size_t)eb * slots));
    if (nb > 0) {
        DevBuf owner, f;
        FMT_TRY(dalloc(owner, sizeof(int) * slots));
        FMT_TRY(dalloc(f, sizeof(int)));
        MUXI_HIP(hipMemsetAsync(owner.p, 0xff, sizeof(int) * slots, s));
        MUXI_HIP(hipMemsetAsync(f.p, 0, sizeof(int), s));
        MUXI_HIP(hipMemsetAsync(o.p[2], 0, (size_t)eb * slots, s));
        hipLaunchKernelGGL(block_cols_kernel, dim3(grid_cap(nb, 256)), dim3(256), 0, s, ukey.as<K>(), nb, (K)nbc,
                           (int*)o.p[1]);
        MUXI_LAUNCH_CHECK();
        const dim3 g(grid_cap(N, 256)), bl(256);
        if (eb == 2)
            hipLaunchKernelGGL(bsr_scatter_kernel<uint16_t>, g, bl, 0, s, row.as<int>(), A->d_colIdx,
                               (const uint16_t*)A->d_values, key.as<K>(), N, ukey.as<K>(), nb, b, owner.as<int>(),
                               (uint16_t*)o.p[2], f.as<int>());
        else if (eb == 4)
            hipLaunchKernelGGL(bsr_scatter_kernel<uint32_t>, g, bl, 0, s, row.as<int>(), A->d_colIdx,
                               (const uint32_t*)A->d_values, key.as<K>(), N, ukey.as<K>(), nb, b, owner.as<int>(),
                               (uint32_t*)o.p[2], f.as<int>());
        else
            hipLaunchKernelGGL(bsr_scatter_kernel<uint64_t>, g, bl, 0, s, row.as<int>(), A->d_colIdx,
                               (const uint64_t*)A->d_values, key.as<K>(), N, ukey.as<K>(), nb, b, owner.as<int>(),
                               (uint64_t*)o.p[2], f.as<int>());
        MUXI_LAUNCH_CHECK();
        int h = 0;
        FMT_TRY(read_flag(f.as<int>(), s, &h));
        if (h & 8) {
            set_error("csr_to_bsr: duplicate (row, column) entries");
            return SPMV_ERROR_INVALID_MATRIX;
        }
    }
    FMT_TRY(offsets_from_sorted<K>(nb > 0 ? ukey.as<K>() : nullptr, nb, nbr, (K)nbc, (int*)o.p[0], s));
    MUXI_HIP(hipStreamSynchronize(s));
    std::memset(out, 0, sizeof(*out));
    out->numRows = R;
    out->numCols = C;
    out->numBlockRows = nbr;
    out->numBlockCols = nbc;
    out->numBlocks = (int)nb;
    out->blockSize = b;
    out->d_rowPtr = (int*)o.p[0];
    out->d_colIdx = (int*)o.p[1];
    out->d_values = o.p[2];
    out->ownsDeviceMemory = true;
    o.keep = true;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_bsr_to_csr(const muxi_bsr_t* A, spmv_type_t vt, muxi_csr_t* out, void* stream) {
    if (!A || !out) return SPMV_ERROR_INVALID_MATRIX;
    const int eb = elem_bytes(vt);
    if (!eb) {
        set_error("bsr_to_csr: values must be FLOAT, DOUBLE or BF16");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    const int b = A->blockSize, R = A->numRows, C = A->numCols, nbr = A->numBlockRows;
    const long nb = A->numBlocks;
    if (b < 1 || b > 64 || R < 0 || C < 0 || nb < 0 || nbr != (int)(((long)R + b - 1) / b) ||
        A->numBlockCols != (int)(((long)C + b - 1) / b) || (nbr > 0 && !A->d_rowPtr) ||
        (nb > 0 && (!A->d_colIdx || !A->d_values))) {
        set_error("invalid BSR matrix (blockSize in [1, 64], numBlockRows = ceil(numRows / blockSize), "
                  "numBlockCols = ceil(numCols / blockSize), device arrays)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    hipStream_t s = (hipStream_t)stream;
    // the block structure is a CSR matrix of numBlockRows x numBlockCols
    muxi_csr_t blk{};
    blk.numRows = nbr;
    blk.numCols = A->numBlockCols;
    blk.nnz = (int)nb;
    blk.d_rowPtr = A->d_rowPtr;
    blk.d_colIdx = A->d_colIdx;
    blk.d_values = A->d_values;
    if (nbr > 0) FMT_TRY(validate_csr(&blk, s));
    DevBuf w, G, T, S;
    FMT_TRY(dalloc(w, sizeof(long long) * (nb + 1)));
    FMT_TRY(dalloc(G, sizeof(long long) * (nb + 1)));
    FMT_TRY(dalloc(T, sizeof(long long) * ((size_t)nbr + 1)));
    FMT_TRY(dalloc(S, sizeof(long long) * ((size_t)nbr + 1)));
    long long total = 0;
    if (nbr > 0) {
        hipLaunchKernelGGL(block_width_kernel, dim3(grid_cap(nb + 1, 256)), dim3(256), 0, s, A->d_colIdx, nb, b, C,
                           w.as<long long>());
        MUXI_LAUNCH_CHECK();
        FMT_TRY(exclusive_scan<long long>(w.as<long long>(), G.as<long long>(), nb + 1, s));
        hipLaunchKernelGGL(blockrow_total_kernel, dim3(grid_cap(nbr + 1L, 256)), dim3(256), 0, s, A->d_rowPtr,
                           G.as<long long>(), nbr, b, R, T.as<long long>());
        MUXI_LAUNCH_CHECK();
        FMT_TRY(exclusive_scan<long long>(T.as<long long>(), S.as<long long>(), nbr + 1L, s));
        MUXI_HIP(hipMemcpyAsync(&total, S.as<long long>() + nbr, sizeof(long long), hipMemcpyDeviceToHost, s));
        MUXI_HIP(hipStreamSynchronize(s));
    }
    if (total > 2147483647LL) {
        set_error("bsr_to_csr: more than INT32_MAX entries");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const long N = (long)total;
    OutArrays o;
    FMT_TRY(alloc_out(o, 0, sizeof(int) * ((size_t)R + 1)));
    FMT_TRY(alloc_out(o, 1, sizeof(int) * N));
    FMT_TRY(alloc_out(o, 2, (size_t)eb * N));
    if (nbr > 0) {
        hipLaunchKernelGGL(bsr_rowptr_kernel, dim3(grid_cap(R + 1L, 256)), dim3(256), 0, s, A->d_rowPtr, G.as<long long>(),
                           S.as<long long>(), nbr, b, R, (int*)o.p[0]);
        MUXI_LAUNCH_CHECK();
    } else {
        MUXI_HIP(hipMemsetAsync(o.p[0], 0, sizeof(int) * ((size_t)R + 1), s));
    }
    if (N > 0) {
        const dim3 g(grid_cap(nb * (long)b * b, 256)), bl(256);
        if (eb == 2)
            hipLaunchKernelGGL(bsr_fill_kernel<uint16_t>, g, bl, 0, s, A->d_rowPtr, A->d_colIdx,
                               (const uint16_t*)A->d_values, G.as<long long>(), S.as<long long>(), nbr, nb, b, R, C,
                               (int*)o.p[1], (uint16_t*)o.p[2]);
        else if (eb == 4)
            hipLaunchKernelGGL(bsr_fill_kernel<uint32_t>, g, bl, 0, s, A->d_rowPtr, A->d_colIdx,
                               (const uint32_t*)A->d_values, G.as<long long>(), S.as<long long>(), nbr, nb, b, R, C,
                               (int*)o.p[1], (uint32_t*)o.p[2]);
        else
            hipLaunchKernelGGL(bsr_fill_kernel<uint64_t>, g, bl, 0, s, A->d_rowPtr, A->d_colIdx,
                               (const uint64_t*)A->d_values, G.as<long long>(), S.as<long long>(), nbr, nb, b, R, C,
                               (int*)o.p[1], (uint64_t*)o.p[2]);
        MUXI_LAUNCH_CHECK();
    }
    MUXI_HIP(hipStreamSynchronize(s));
    std::memset(out, 0, sizeof(*out));
    out->numRows = R;
    out->numCols = C;
    out->nnz = (int)N;
    out->d_rowPtr = (int*)o.p[0];
    out->d_colIdx = (int*)o.p[1];
    out->d_values = o.p[2];
    out->ownsDeviceMemory = true;
    o.keep = true;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_csr_free_device(muxi_csr_t* A) {
    if (!A) return SPMV_ERROR_INVALID_MATRIX;
    if (A->ownsDeviceMemory) {
        (void)hipFree(A->d_rowPtr);
        (void)hipFree(A->d_colIdx);
        (void)hipFree(A->d_values);
        A->d_rowPtr = A->d_colIdx = nullptr;
        A->d_values = nullptr;
        A->ownsDeviceMemory = false;
    }
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_coo_free_device(muxi_coo_t* A) {
    if (!A) return SPMV_ERROR_INVALID_MATRIX;
    if (A->ownsDeviceMemory) {
        (void)hipFree(A->d_rowIdx);
        (void)hipFree(A->d_colIdx);
        (void)hipFree(A->d_values);
        A->d_rowIdx = A->d_colIdx = nullptr;
        A->d_values = nullptr;
        A->ownsDeviceMemory = false;
    }
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_bsr_free_device(muxi_bsr_t* A) {
    if (!A) return SPMV_ERROR_INVALID_MATRIX;
    if (A->ownsDeviceMemory) {
        (void)hipFree(A->d_rowPtr);
        (void)hipFree(A->d_colIdx);
        (void)hipFree(A->d_values);
        A->d_rowPtr = A->d_colIdx = nullptr;
        A->d_values = nullptr;
        A->ownsDeviceMemory = false;
    }
    return SPMV_SUCCESS;
}
