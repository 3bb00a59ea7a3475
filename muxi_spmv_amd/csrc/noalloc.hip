// Merge-path SpMV that allocates nothing: muxi_spmv_merge_based_noalloc_f32.
//
// The per-call merge-path entry point (muxi_spmv_merge_based_f32) takes its
// scratch -- partition diagonals, their row coordinates and one carry per
// partition -- from the stream-ordered allocator. Callers that may not
// allocate (code under stream capture, tight serving loops) use this variant:
// the same kernels, with the scratch in a static device array of
// kNoAllocScratchBytes per GPU. The array's device address comes from
// hipGetSymbolAddress, once per device with that device current, and is kept
// per device; host code never passes the array's own name to a kernel (on the
// host that name is only a stand-in for the symbol).
//
// Nothing in the scratch has to start at zero: the partition kernel writes
// every diagonal, the coordinate kernel every row coordinate and the SpMV
// kernel every carry before any kernel reads them. Every call on one GPU uses
// that GPU's one array, so calls that overlap on a GPU (two streams, or two
// host threads) race on it; calls on one stream are ordered and safe.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "kernels.hpp"
#include "plan.hpp"

namespace muxi {

constexpr size_t kNoAllocScratchBytes = 2u << 20;  // 2 MiB per GPU
constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocIpt = 8;  // items per thread, as muxi_spmv_merge_based_f32

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The scratch's device address on the current device (looked up on the first call there).
static spmv_status_t scratch_on_current_device(void** out) {
    int dev = -1;
    MUXI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("merge_based_noalloc: device id " + std::to_string(dev) + " beyond the " +
                  std::to_string(kNoAllocMaxDevices) + " devices the static scratch table covers");
        return SPMV_ERROR_DEVICE;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("merge_based_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return SPMV_ERROR_DEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SPMV_SUCCESS;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Scratch layout for P partitions: diagonals [0, P], row coordinates [0, P], then (256-aligned) P carries.
static size_t noalloc_carry_offset(int P) { return align256(2 * sizeof(int) * ((size_t)P + 1)); }
static size_t noalloc_bytes(int P) {
    return noalloc_carry_offset(P) + merge_carry_bytes(P, SPMV_TYPE_FLOAT, false);
}

}  // namespace muxi

using namespace muxi;

extern "C" spmv_status_t muxi_spmv_merge_based_noalloc_f32(const muxi_csr_t* A, const float* x, float* y,
                                                           void* stream) {
    // every argument check comes before the first HIP call
    if (!A) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (A->numRows < 0 || A->numCols < 0 || A->nnz < 0) {
        set_error("negative matrix dimension");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (A->nnz == 0) return SPMV_SUCCESS;  // as muxi_spmv_merge_based_f32: y untouched
    if ((long)A->numRows + (long)A->nnz > 2147483647L) {
        set_error("numRows + nnz exceeds INT32_MAX (merge-path coordinates are int32)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (!A->d_rowPtr || !A->d_colIdx || !A->d_values) {
        set_error("matrix device arrays are NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (!x || !y) {
        set_error("x or y is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    const int L = A->numRows + A->nnz;
    const int P = merge_default_partitions(L, kNoAllocIpt);
    if (noalloc_bytes(P) > kNoAllocScratchBytes) {
        long pmax = 1;
        while (noalloc_bytes((int)(pmax * 2)) <= kNoAllocScratchBytes) pmax *= 2;
        long lo = pmax, hi = pmax * 2;  // largest P that fits: noalloc_bytes(lo) fits, (hi) does not
        while (hi - lo > 1) {
            const long mid = (lo + hi) / 2;
            (noalloc_bytes((int)mid) <= kNoAllocScratchBytes ? lo : hi) = mid;
        }
        set_error("merge_based_noalloc: numRows + nnz = " + std::to_string(L) + " needs " +
                  std::to_string(noalloc_bytes(P)) + " bytes of scratch; the static scratch holds " +
                  std::to_string(kNoAllocScratchBytes) + " (numRows + nnz up to " +
                  std::to_string(lo * 256L * kNoAllocIpt) + "); use muxi_spmv_merge_based_f32 or a plan");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    void* base = nullptr;
    spmv_status_t st = scratch_on_current_device(&base);
    if (st != SPMV_SUCCESS) return st;
    hipStream_t s = (hipStream_t)stream;
    int* pos = (int*)base;
    int* row = pos + (P + 1);
    void* carry = (unsigned char*)base + noalloc_carry_offset(P);
    if ((st = launch_merge_partitions(A->numRows, A->nnz, P, pos, s)) != SPMV_SUCCESS) return st;
    if ((st = launch_merge_coords(A->d_rowPtr, A->numRows, A->nnz, pos, P + 1, row, nullptr, s)) != SPMV_SUCCESS)
        return st;
    const DevCsr d{A->numRows, A->numCols, A->nnz, A->d_rowPtr, A->d_colIdx, A->d_values};
    return launch_merge<float, float, float, false>(d, kNoAllocIpt, pos, row, P, carry, x, y, 1.0f, 0.0f, s);
}
