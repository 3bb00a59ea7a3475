// Device application of row and column permutations.
#include "common.hpp"
#include "muxi_spmv/muxi_spmv_prep.h"

namespace muxi {

__global__ void remap_kernel(const int* __restrict__ in, const int* __restrict__ map, int* __restrict__ out, int n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = map[in[i]];
}

template <typename T>
__global__ void gather_kernel(const T* __restrict__ src, const int* __restrict__ idx, T* __restrict__ dst, int n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

template <typename T>
__global__ void scatter_kernel(const T* __restrict__ src, const int* __restrict__ idx, T* __restrict__ dst, int n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[idx[i]] = src[i];
}

// One wavefront per output row copies the source row with coalesced lanes.
template <typename T>
__global__ void permute_rows_kernel(int R, const int* __restrict__ rp, const int* __restrict__ ci,
                                    const T* __restrict__ v, const int* __restrict__ rowMap,
                                    const int* __restrict__ nrp, int* __restrict__ nci, T* __restrict__ nv) {
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= R) return;
    const int o = rowMap[wave];
    const int b = rp[o], len = rp[o + 1] - b, d = nrp[wave];
    for (int i = lane; i < len; i += 64) {
        nci[d + i] = ci[b + i];
        nv[d + i] = v[b + i];
    }
}

static unsigned grid_cap(long n) {
    long g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace muxi

using namespace muxi;

extern "C" spmv_status_t muxi_remap_indices_dev(const int* in, const int* map, int* out, int n, void* stream) {
    if (n <= 0) return SPMV_SUCCESS;
    if (!in || !map || !out) return SPMV_ERROR_INVALID_VECTOR;
    hipLaunchKernelGGL(remap_kernel, dim3(grid_cap(n)), dim3(256), 0, (hipStream_t)stream, in, map, out, n);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template <typename T>
static void launch_move(bool gather, const void* src, const int* idx, void* dst, int n, hipStream_t s) {
    if (gather)
        hipLaunchKernelGGL(gather_kernel<T>, dim3(grid_cap(n)), dim3(256), 0, s, (const T*)src, idx, (T*)dst, n);
    else
        hipLaunchKernelGGL(scatter_kernel<T>, dim3(grid_cap(n)), dim3(256), 0, s, (const T*)src, idx, (T*)dst, n);
}

static spmv_status_t by_size(bool gather, const void* src, const int* idx, void* dst, int n, int eb, hipStream_t s) {
    switch (eb) {
        case 2: launch_move<uint16_t>(gather, src, idx, dst, n, s); break;
        case 4: launch_move<uint32_t>(gather, src, idx, dst, n, s); break;
        case 8: launch_move<uint64_t>(gather, src, idx, dst, n, s); break;
        default: set_error("element size must be 2, 4 or 8"); return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_gather_dev(const void* src, const int* idx, void* dst, int n, int eb, void* stream) {
    if (n <= 0) return SPMV_SUCCESS;
    if (!src || !idx || !dst) return SPMV_ERROR_INVALID_VECTOR;
    return by_size(true, src, idx, dst, n, eb, (hipStream_t)stream);
}

extern "C" spmv_status_t muxi_scatter_dev(const void* src, const int* idx, void* dst, int n, int eb, void* stream) {
    if (n <= 0) return SPMV_SUCCESS;
    if (!src || !idx || !dst) return SPMV_ERROR_INVALID_VECTOR;
    return by_size(false, src, idx, dst, n, eb, (hipStream_t)stream);
}

extern "C" spmv_status_t muxi_permute_rows_dev(int R, const int* rp, const int* ci, const void* v, int eb,
                                               const int* rowMap, const int* nrp, int* nci, void* nv, void* stream) {
    if (R <= 0) return SPMV_SUCCESS;
    if (!rp || !rowMap || !nrp) return SPMV_ERROR_INVALID_MATRIX;
    hipStream_t s = (hipStream_t)stream;
    dim3 g(grid_for((long)R * 64, 256));
    switch (eb) {
        case 2: hipLaunchKernelGGL(permute_rows_kernel<uint16_t>, g, dim3(256), 0, s, R, rp, ci, (const uint16_t*)v, rowMap, nrp, nci, (uint16_t*)nv); break;
        case 4: hipLaunchKernelGGL(permute_rows_kernel<uint32_t>, g, dim3(256), 0, s, R, rp, ci, (const uint32_t*)v, rowMap, nrp, nci, (uint32_t*)nv); break;
        case 8: hipLaunchKernelGGL(permute_rows_kernel<uint64_t>, g, dim3(256), 0, s, R, rp, ci, (const uint64_t*)v, rowMap, nrp, nci, (uint64_t*)nv); break;
        default: set_error("element size must be 2, 4 or 8"); return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}
