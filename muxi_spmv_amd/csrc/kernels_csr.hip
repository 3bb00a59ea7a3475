// Row-parallel CSR kernels for wave64: one lane per row (scalar), one
// wavefront per row (vector) and a virtual wavefront of L lanes per row with
// the block's row pointers staged in LDS.
//
// Reference counterparts: spmv_csr_scalar_kernel (src/spmv/csr/spmv_csr.cuh:83-116),
// spmv_csr_vector_kernel (:121-152) and the virtual-warp family
// (tests/test_virtual_warp_size.cu:30-67, tests/test_adaptive_warp_optimized.cu:31-230).
// Grids here cover every row (the reference caps grids at 65535 blocks,
// spmv_csr.cu:44-49, and then skips rows).
#include "kernels.hpp"

namespace muxi {

template <typename VT, typename XT, typename AT, bool COMP>
__global__ __launch_bounds__(256) void scalar_kernel(int R, const int* __restrict__ rp, const int* __restrict__ ci,
                                                     const VT* __restrict__ v, const XT* __restrict__ x,
                                                     XT* __restrict__ y, AT alpha, AT beta) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= R) return;
    Acc<AT, COMP> a;
    a.zero();
    int b = rp[row], e = rp[row + 1];
    for (int i = b; i < e; ++i) a.fma(v[i], x[ci[i]]);
    store_y(y, row, a.value(), alpha, beta);
}

// L lanes per row, 256/L rows per block. The block's R/L+1 row pointers are
// read once, coalesced, into LDS.
template <typename VT, typename XT, typename AT, bool COMP, int L, bool STAGE>
__global__ __launch_bounds__(256) void vwave_kernel(int R, const int* __restrict__ rp, const int* __restrict__ ci,
                                                    const VT* __restrict__ v, const XT* __restrict__ x,
                                                    XT* __restrict__ y, AT alpha, AT beta) {
    constexpr int RPB = 256 / L;
    __shared__ int s_rp[RPB + 1];
    const long row0 = (long)blockIdx.x * RPB;
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    const long row = row0 + g;
    int b = 0, e = 0;
    if (STAGE) {
        for (int i = threadIdx.x; i <= RPB; i += 256) {
            long r = row0 + i;
            s_rp[i] = rp[r < R ? r : R];
        }
        __syncthreads();
        if (row < R) {
            b = s_rp[g];
            e = s_rp[g + 1];
        }
    } else if (row < R) {
        b = rp[row];
        e = rp[row + 1];
    }
    Acc<AT, COMP> a;
    a.zero();
    int i = b + lane;
    // two independent streams per lane keep more gathers in flight
    Acc<AT, COMP> a2;
    a2.zero();
    for (; i + L < e; i += 2 * L) {
        int c0 = ci[i], c1 = ci[i + L];
        VT v0 = v[i], v1 = v[i + L];
        a.fma(v0, x[c0]);
        a2.fma(v1, x[c1]);
    }
    if (i < e) a.fma(v[i], x[ci[i]]);
    a.add(a2);
    a = group_reduce<L>(a);
    if (lane == 0 && row < R) store_y(y, row, a.value(), alpha, beta);
}

template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_scalar(const DevCsr& A, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s) {
    if (A.R <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL((scalar_kernel<VT, XT, AT, COMP>), dim3(grid_for(A.R, 256)), dim3(256), 0, s, A.R, A.rp, A.ci,
                       (const VT*)A.v, x, y, alpha, beta);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template <typename VT, typename XT, typename AT, bool COMP, int L, bool STAGE>
static spmv_status_t vwave_go(const DevCsr& A, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s) {
    constexpr int RPB = 256 / L;
    hipLaunchKernelGGL((vwave_kernel<VT, XT, AT, COMP, L, STAGE>), dim3(grid_for(A.R, RPB)), dim3(256), 0, s, A.R,
                       A.rp, A.ci, (const VT*)A.v, x, y, alpha, beta);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_vector(const DevCsr& A, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s) {
    if (A.R <= 0) return SPMV_SUCCESS;
    return vwave_go<VT, XT, AT, COMP, 64, false>(A, x, y, alpha, beta, s);
}

template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_vwave(const DevCsr& A, int lanes, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s) {
    if (A.R <= 0) return SPMV_SUCCESS;
    switch (lanes) {
        case 1: return vwave_go<VT, XT, AT, COMP, 1, true>(A, x, y, alpha, beta, s);
        case 2: return vwave_go<VT, XT, AT, COMP, 2, true>(A, x, y, alpha, beta, s);
        case 4: return vwave_go<VT, XT, AT, COMP, 4, true>(A, x, y, alpha, beta, s);
        case 8: return vwave_go<VT, XT, AT, COMP, 8, true>(A, x, y, alpha, beta, s);
        case 16: return vwave_go<VT, XT, AT, COMP, 16, true>(A, x, y, alpha, beta, s);
        case 32: return vwave_go<VT, XT, AT, COMP, 32, true>(A, x, y, alpha, beta, s);
        case 64: return vwave_go<VT, XT, AT, COMP, 64, true>(A, x, y, alpha, beta, s);
        default: set_error("vwave: lanes per row must be a power of two in [1, 64]"); return SPMV_ERROR_INTERNAL;
    }
}

#define INST(VT, XT, AT, COMP)                                                                                   \
    template spmv_status_t launch_scalar<VT, XT, AT, COMP>(const DevCsr&, const XT*, XT*, AT, AT, hipStream_t);   \
    template spmv_status_t launch_vector<VT, XT, AT, COMP>(const DevCsr&, const XT*, XT*, AT, AT, hipStream_t);   \
    template spmv_status_t launch_vwave<VT, XT, AT, COMP>(const DevCsr&, int, const XT*, XT*, AT, AT, hipStream_t);
MUXI_FOR_EACH_COMBO(INST)
#undef INST

}  // namespace muxi
