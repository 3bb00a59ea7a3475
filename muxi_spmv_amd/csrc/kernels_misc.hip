// Device FP64 oracle SpMV, element conversion, y scaling and the transpose
// product.
#include "kernels.hpp"

namespace muxi {

// One lane per row; products and sums in FP64 with the Dot2 compensation
// carried in double (TwoProduct by FMA, TwoSum by Knuth), i.e. the result is
// as if summed in about 106-bit precision and rounded to FP64.
template <typename VT, typename XT>
__global__ __launch_bounds__(256) void oracle_f64_kernel(int R, const int* __restrict__ rp, const int* __restrict__ ci,
                                                         const VT* __restrict__ v, const XT* __restrict__ x,
                                                         double* __restrict__ y) {
    long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= R) return;
    double s = 0.0, c = 0.0;
    for (int i = rp[row]; i < rp[row + 1]; ++i) {
        double a = to_f64(v[i]), b = to_f64(x[ci[i]]);
        double p = __dmul_rn(a, b);
        double ep = __fma_rn(a, b, -p);
        double t = __dadd_rn(s, p);
        double bp = __dsub_rn(t, s);
        double e = __dadd_rn(__dsub_rn(s, __dsub_rn(t, bp)), __dsub_rn(p, bp));
        s = t;
        c = __dadd_rn(c, __dadd_rn(e, ep));
    }
    y[row] = __dadd_rn(s, c);
}

template <typename VT, typename XT>
spmv_status_t launch_oracle_f64(const DevCsr& A, const XT* x, double* y, hipStream_t s) {
    if (A.R <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL((oracle_f64_kernel<VT, XT>), dim3(grid_for(A.R, 256)), dim3(256), 0, s, A.R, A.rp, A.ci,
                       (const VT*)A.v, x, y);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}
template spmv_status_t launch_oracle_f64<float, float>(const DevCsr&, const float*, double*, hipStream_t);
template spmv_status_t launch_oracle_f64<double, float>(const DevCsr&, const float*, double*, hipStream_t);
template spmv_status_t launch_oracle_f64<bf16_t, float>(const DevCsr&, const float*, double*, hipStream_t);
template spmv_status_t launch_oracle_f64<float, double>(const DevCsr&, const double*, double*, hipStream_t);
template spmv_status_t launch_oracle_f64<double, double>(const DevCsr&, const double*, double*, hipStream_t);
template spmv_status_t launch_oracle_f64<bf16_t, double>(const DevCsr&, const double*, double*, hipStream_t);

template <typename S, typename D>
__global__ void convert_kernel(const S* __restrict__ src, D* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if constexpr (std::is_same<D, bf16_t>::value) {
            bf16_t b;
            b.bits = f32_to_bf16_bits((float)src[i]);
            dst[i] = b;
        } else if constexpr (std::is_same<S, bf16_t>::value) {
            dst[i] = (D)to_f32(src[i]);
        } else {
            dst[i] = (D)src[i];
        }
    }
}

template <typename S, typename D>
static spmv_status_t conv_go(const void* src, void* dst, size_t n, hipStream_t s) {
    unsigned g = (unsigned)((n + 255) / 256);
    if (g > 65536) g = 65536;
    if (g == 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL((convert_kernel<S, D>), dim3(g), dim3(256), 0, s, (const S*)src, (D*)dst, n);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

spmv_status_t launch_convert(const void* src, spmv_type_t st, void* dst, spmv_type_t dt, size_t n, hipStream_t s) {
    if (st == SPMV_TYPE_FLOAT && dt == SPMV_TYPE_BF16) return conv_go<float, bf16_t>(src, dst, n, s);
    if (st == SPMV_TYPE_DOUBLE && dt == SPMV_TYPE_BF16) return conv_go<double, bf16_t>(src, dst, n, s);
    if (st == SPMV_TYPE_FLOAT && dt == SPMV_TYPE_DOUBLE) return conv_go<float, double>(src, dst, n, s);
    if (st == SPMV_TYPE_DOUBLE && dt == SPMV_TYPE_FLOAT) return conv_go<double, float>(src, dst, n, s);
    if (st == SPMV_TYPE_BF16 && dt == SPMV_TYPE_FLOAT) return conv_go<bf16_t, float>(src, dst, n, s);
    if (st == SPMV_TYPE_BF16 && dt == SPMV_TYPE_DOUBLE) return conv_go<bf16_t, double>(src, dst, n, s);
    if (st == dt && (st == SPMV_TYPE_FLOAT || st == SPMV_TYPE_DOUBLE || st == SPMV_TYPE_BF16)) {
        size_t es = st == SPMV_TYPE_DOUBLE ? 8 : (st == SPMV_TYPE_FLOAT ? 4 : 2);
        MUXI_HIP(hipMemcpyAsync(dst, src, n * es, hipMemcpyDeviceToDevice, s));
        return SPMV_SUCCESS;
    }
    set_error("convert: unsupported type pair");
    return SPMV_ERROR_UNSUPPORTED_TYPE;
}

template <typename XT>
__global__ void scale_kernel(XT* __restrict__ y, long n, XT beta) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = beta == XT(0) ? XT(0) : beta * y[i];
}

template <typename XT>
spmv_status_t launch_scale_y(XT* y, long n, XT beta, hipStream_t s) {
    if (n <= 0) return SPMV_SUCCESS;
    unsigned g = grid_for(n, 256);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(scale_kernel<XT>, dim3(g), dim3(256), 0, s, y, n, beta);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}
template spmv_status_t launch_scale_y<float>(float*, long, float, hipStream_t);
template spmv_status_t launch_scale_y<double>(double*, long, double, hipStream_t);

// y += alpha * A^T x, one lane per row scattering with atomics: the
// counterpart of spmv_csr_scalar_kernel<T, true> (src/spmv/csr/spmv_csr.cuh:105-106).
// The caller scales y by beta first. Atomic order makes the last bits
// run-dependent; muxi_csr_transpose followed by a plan is deterministic instead.
template <typename VT, typename XT>
__global__ __launch_bounds__(256) void transpose_scalar_kernel(int R, const int* __restrict__ rp,
                                                               const int* __restrict__ ci, const VT* __restrict__ v,
                                                               const XT* __restrict__ x, XT* __restrict__ y, XT alpha) {
    long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= R) return;
    XT xr = alpha * x[row];
    for (int i = rp[row]; i < rp[row + 1]; ++i) {
        XT val = (XT)(sizeof(XT) == 8 ? to_f64(v[i]) : to_f32(v[i]));
        atomicAdd(&y[ci[i]], val * xr);
    }
}

template <typename VT, typename XT>
spmv_status_t launch_transpose_scalar(const DevCsr& A, const XT* x, XT* y, XT alpha, hipStream_t s) {
    if (A.R <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL((transpose_scalar_kernel<VT, XT>), dim3(grid_for(A.R, 256)), dim3(256), 0, s, A.R, A.rp, A.ci,
                       (const VT*)A.v, x, y, alpha);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}
template spmv_status_t launch_transpose_scalar<float, float>(const DevCsr&, const float*, float*, float, hipStream_t);
template spmv_status_t launch_transpose_scalar<double, double>(const DevCsr&, const double*, double*, double,
                                                               hipStream_t);

}  // namespace muxi
