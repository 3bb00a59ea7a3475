// ELLPACK build and SpMV.
//
// Layout follows csr_to_ellpack (src/spmv/ellpack/spmv_ellpack.cu:11-70):
// width = max row length unless given; slot (row, i) lives at row + i*R
// (column-major, so lane-adjacent rows are address-adjacent for every i);
// padding is col = -1, value = 0; rows longer than the width are truncated.
// Unlike the reference, the packing runs on the device from the CSR device
// arrays, one lane per row, instead of a host loop plus blocking copies.
#include "kernels.hpp"

namespace muxi {

__global__ void max_row_len_kernel(const int* __restrict__ rp, int R, int* __restrict__ out) {
    int m = 0;
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (long)gridDim.x * blockDim.x) {
        int len = rp[r + 1] - rp[r];
        m = len > m ? len : m;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        int o = __shfl_xor(m, d);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

spmv_status_t launch_max_row_len(const int* rp, int R, int* d_out, hipStream_t s) {
    MUXI_HIP(hipMemsetAsync(d_out, 0, sizeof(int), s));
    if (R <= 0) return SPMV_SUCCESS;
    unsigned g = grid_for(R, 256);
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(max_row_len_kernel, dim3(g), dim3(256), 0, s, rp, R, d_out);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template <typename VT>
__global__ void csr_to_ell_kernel(int R, int width, const int* __restrict__ rp, const int* __restrict__ ci,
                                  const VT* __restrict__ v, int* __restrict__ ecol, VT* __restrict__ eval) {
    long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= R) return;
    int b = rp[row], len = rp[row + 1] - b;
    for (int i = 0; i < width; ++i) {
        long idx = row + (long)i * R;
        if (i < len) {
            ecol[idx] = ci[b + i];
            eval[idx] = v[b + i];
        } else {
            ecol[idx] = -1;
            VT z;
            __builtin_memset(&z, 0, sizeof(VT));
            eval[idx] = z;
        }
    }
}

template <typename VT>
spmv_status_t launch_csr_to_ell(const DevCsr& A, int width, int* d_col, VT* d_val, hipStream_t s) {
    if (A.R <= 0 || width <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL(csr_to_ell_kernel<VT>, dim3(grid_for(A.R, 256)), dim3(256), 0, s, A.R, width, A.rp, A.ci,
                       (const VT*)A.v, d_col, d_val);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

// One lane per row; each slot index i is one coalesced 64-lane load of
// columns and values. Padding (col < 0) contributes nothing.
template <typename VT, typename XT, typename AT, bool COMP>
__global__ __launch_bounds__(256) void ell_spmv_kernel(int R, int C, int width, const int* __restrict__ ecol,
                                                       const VT* __restrict__ eval, const XT* __restrict__ x,
                                                       XT* __restrict__ y, AT alpha, AT beta) {
    long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= R) return;
    Acc<AT, COMP> a;
    a.zero();
    // Slots in groups of U: all U column/value loads, then all U gathers, are
    // issued before any is used. A padding slot gathers x[0] (its column is
    // -1) and its product is skipped, so no load sits behind a branch. With no
    // columns every slot is padding and x is never read.
    constexpr int U = 8;
    if (C <= 0) width = 0;
    int i = 0;
    for (; i + U <= width; i += U) {
        int c[U];
        VT v[U];
        XT xv[U];
#pragma unroll
        for (int k = 0; k < U; ++k) c[k] = ntl(ecol + row + (long)(i + k) * R);
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = ntl(eval + row + (long)(i + k) * R);
#pragma unroll
        for (int k = 0; k < U; ++k) xv[k] = x[c[k] > 0 ? c[k] : 0];
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (c[k] >= 0) a.fma(v[k], xv[k]);
    }
    if (i < width) {
        int c[U];
        VT v[U];
        XT xv[U];
        const int n = width - i;
#pragma unroll
        for (int k = 0; k < U; ++k) c[k] = ntl(ecol + row + (long)(i + (k < n ? k : n - 1)) * R);
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = ntl(eval + row + (long)(i + (k < n ? k : n - 1)) * R);
#pragma unroll
        for (int k = 0; k < U; ++k) xv[k] = x[c[k] > 0 ? c[k] : 0];
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (k < n && c[k] >= 0) a.fma(v[k], xv[k]);
    }
    store_y(y, row, a.value(), alpha, beta);
}

template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_ell(int R, int C, int width, const int* d_col, const VT* d_val, const XT* x, XT* y, AT alpha,
                         AT beta, hipStream_t s) {
    if (R <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL((ell_spmv_kernel<VT, XT, AT, COMP>), dim3(grid_for(R, 256)), dim3(256), 0, s, R, C, width,
                       d_col, d_val, x, y, alpha, beta);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template spmv_status_t launch_csr_to_ell<float>(const DevCsr&, int, int*, float*, hipStream_t);
template spmv_status_t launch_csr_to_ell<double>(const DevCsr&, int, int*, double*, hipStream_t);
template spmv_status_t launch_csr_to_ell<bf16_t>(const DevCsr&, int, int*, bf16_t*, hipStream_t);

#define INST(VT, XT, AT, COMP)                                                                                       \
    template spmv_status_t launch_ell<VT, XT, AT, COMP>(int, int, int, const int*, const VT*, const XT*, XT*, AT, AT, \
                                                        hipStream_t);
MUXI_FOR_EACH_COMBO(INST)
#undef INST

}  // namespace muxi
