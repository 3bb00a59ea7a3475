// Adaptive virtual-wavefront SpMV (row-length binned, LDS staged).
//
// The reference's adaptive family (tests/test_adaptive_thread.cu:178-210,
// tests/test_adaptive_warp_optimized.cu:31-230, spmv_csr_adaptive_kernel in
// src/spmv/csr/spmv_csr.cuh:1000-1056) picks lanes per row from row lengths
// and stages row pointers in shared memory. This version bins at workgroup
// granularity: the plan cuts the rows into blocks of consecutive rows holding
// at most NB nonzeros (a row longer than NB gets its own blocks of NB
// nonzeros each). A multi-row block streams its nonzeros' products into LDS
// with coalesced loads, then reduces its rows with a virtual wavefront whose
// width (1..64 lanes) follows the block's mean row length. A long row's
// blocks each reduce NB products to one partial, and a finalize kernel adds a
// row's partials in order. Fixed summation order: reproducible results.
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"

namespace muxi {

template <typename AT, bool COMP, int L, typename XT>
__device__ __forceinline__ void stream_rows(int nrows, int n0, const int* __restrict__ s_rp,
                                            const Prod<AT, COMP>* __restrict__ s_prod, long r0, XT* __restrict__ y,
                                            AT alpha, AT beta) {
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    for (int rr = 0; rr < nrows; rr += 256 / L) {
        const int row = rr + g;
        Acc<AT, COMP> a;
        a.zero();
        if (row < nrows) {
            const int b = s_rp[row] - n0, e = s_rp[row + 1] - n0;
            for (int i = b + lane; i < e; i += L) s_prod[i].into(a);
        }
        a = group_reduce<L>(a);
        if (lane == 0 && row < nrows) store_y(y, r0 + row, a.value(), alpha, beta);
    }
}

template <typename VT, typename XT, typename AT, bool COMP, int NB>
__global__ __launch_bounds__(256) void adaptive_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                       const VT* __restrict__ vals, const XT* __restrict__ x,
                                                       XT* __restrict__ y, const int* __restrict__ blk_row,
                                                       const int* __restrict__ blk_chunk, AT alpha, AT beta,
                                                       Acc<AT, COMP>* __restrict__ partial) {
    using A = Acc<AT, COMP>;
    using P = Prod<AT, COMP>;
    __shared__ P s_prod[NB];
    __shared__ int s_rp[NB + 1];
    __shared__ A s_w[4];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int r0 = blk_row[b], r1 = blk_row[b + 1], chunk = blk_chunk[b];
    if (chunk < 0) {
        const int nrows = r1 - r0;
        const int n0 = rp[r0], nn = rp[r1] - n0;
        constexpr int K = NB / 256;
        // clamped (unconditional) loads: all K columns and values, then all K
        // gathers of a lane are in flight together; nn > 0 is block-uniform
        if (nn > 0) {
            int c[K];
            VT v[K];
            XT xv[K];
#pragma unroll
            for (int k = 0; k < K; ++k) c[k] = ntl(ci + n0 + min(tid + k * 256, nn - 1));
#pragma unroll
            for (int k = 0; k < K; ++k) v[k] = ntl(vals + n0 + min(tid + k * 256, nn - 1));
#pragma unroll
            for (int k = 0; k < K; ++k) xv[k] = x[c[k]];
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (tid + k * 256 < nn) s_prod[tid + k * 256] = P::make(v[k], xv[k]);
        }
        for (int i = tid; i <= nrows; i += 256) s_rp[i] = rp[r0 + i];
        __syncthreads();
        const int avg = nrows > 0 ? nn / nrows : 0;
        if (avg <= 2)
            stream_rows<AT, COMP, 1>(nrows, n0, s_rp, s_prod, r0, y, alpha, beta);
        else if (avg <= 4)
            stream_rows<AT, COMP, 2>(nrows, n0, s_rp, s_prod, r0, y, alpha, beta);
        else if (avg <= 8)
            stream_rows<AT, COMP, 4>(nrows, n0, s_rp, s_prod, r0, y, alpha, beta);
        else if (avg <= 16)
            stream_rows<AT, COMP, 8>(nrows, n0, s_rp, s_prod, r0, y, alpha, beta);
        else if (avg <= 32)
            stream_rows<AT, COMP, 16>(nrows, n0, s_rp, s_prod, r0, y, alpha, beta);
        else if (avg <= 64)
            stream_rows<AT, COMP, 32>(nrows, n0, s_rp, s_prod, r0, y, alpha, beta);
        else
            stream_rows<AT, COMP, 64>(nrows, n0, s_rp, s_prod, r0, y, alpha, beta);
    } else {
        // one NB-nonzero piece of a long row r0
        const long e0 = (long)rp[r0] + (long)chunk * NB;
        const long e1 = e0 + NB < rp[r0 + 1] ? e0 + NB : rp[r0 + 1];
        A a;
        a.zero();
        constexpr int K = NB / 256;
        const long last = e1 - 1;  // a piece holds at least one nonzero
        auto at = [&](int k) { const long i = e0 + tid + k * 256; return i < last ? i : last; };
        int c[K];
        VT v[K];
        XT xv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) c[k] = ntl(ci + at(k));
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = ntl(vals + at(k));
#pragma unroll
        for (int k = 0; k < K; ++k) xv[k] = x[c[k]];
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (e0 + tid + k * 256 < e1) a.fma(v[k], xv[k]);
        a = group_reduce<64>(a);
        if ((tid & 63) == 0) s_w[tid >> 6] = a;
        __syncthreads();
        if (tid == 0) {
            A t = s_w[0];
            t.add(s_w[1]);
            t.add(s_w[2]);
            t.add(s_w[3]);
            partial[b] = t;
        }
    }
}

// One wavefront per long row: y[row] = alpha * sum of its pieces (block order) + beta * y[row].
template <typename XT, typename AT, bool COMP>
__global__ __launch_bounds__(256) void adaptive_finalize_kernel(const int* __restrict__ long_row,
                                                                const int* __restrict__ long_first, int nlong,
                                                                const Acc<AT, COMP>* __restrict__ partial,
                                                                XT* __restrict__ y, AT alpha, AT beta) {
    const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= nlong) return;
    const int f = long_first[2 * w], l = long_first[2 * w + 1];
    Acc<AT, COMP> a;
    a.zero();
    for (int i = f + lane; i < l; i += 64) a.add(partial[i]);
    a = group_reduce<64>(a);
    if (lane == 0) store_y(y, long_row[w], a.value(), alpha, beta);
}

template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_adaptive(const DevCsr& A, int nb, int nblocks, const int* blk_row, const int* blk_chunk,
                              int nlong, const int* long_row, const int* long_first, void* partial, const XT* x, XT* y,
                              AT alpha, AT beta, hipStream_t s, hipEvent_t ev) {
    if (A.R <= 0 || nblocks <= 0) return SPMV_SUCCESS;
    using Ap = Acc<AT, COMP>;
    switch (nb) {
        case 256:
            hipLaunchKernelGGL((adaptive_kernel<VT, XT, AT, COMP, 256>), dim3(nblocks), dim3(256), 0, s, A.rp, A.ci,
                               (const VT*)A.v, x, y, blk_row, blk_chunk, alpha, beta, (Ap*)partial);
            break;
        case 512:
            hipLaunchKernelGGL((adaptive_kernel<VT, XT, AT, COMP, 512>), dim3(nblocks), dim3(256), 0, s, A.rp, A.ci,
                               (const VT*)A.v, x, y, blk_row, blk_chunk, alpha, beta, (Ap*)partial);
            break;
        case 1024:
            hipLaunchKernelGGL((adaptive_kernel<VT, XT, AT, COMP, 1024>), dim3(nblocks), dim3(256), 0, s, A.rp, A.ci,
                               (const VT*)A.v, x, y, blk_row, blk_chunk, alpha, beta, (Ap*)partial);
            break;
        case 2048:
            hipLaunchKernelGGL((adaptive_kernel<VT, XT, AT, COMP, 2048>), dim3(nblocks), dim3(256), 0, s, A.rp, A.ci,
                               (const VT*)A.v, x, y, blk_row, blk_chunk, alpha, beta, (Ap*)partial);
            break;
        default: set_error("adaptive: block nonzeros must be 256, 512, 1024 or 2048"); return SPMV_ERROR_INTERNAL;
    }
    MUXI_LAUNCH_CHECK();
    if (ev) MUXI_HIP(hipEventRecord(ev, s));
    if (nlong > 0) {
        hipLaunchKernelGGL((adaptive_finalize_kernel<XT, AT, COMP>), dim3(grid_for((long)nlong * 64, 256)), dim3(256),
                           0, s, long_row, long_first, nlong, (const Ap*)partial, y, alpha, beta);
        MUXI_LAUNCH_CHECK();
    }
    return SPMV_SUCCESS;
}

#define INST(VT, XT, AT, COMP)                                                                                    \
    template spmv_status_t launch_adaptive<VT, XT, AT, COMP>(const DevCsr&, int, int, const int*, const int*, int, \
                                                             const int*, const int*, void*, const XT*, XT*, AT, AT,  \
                                                             hipStream_t, hipEvent_t);
MUXI_FOR_EACH_COMBO(INST)
#undef INST

// Host side of the plan: cut rows into blocks. Long rows (> nb nonzeros)
// become ceil(len/nb) chunk blocks whose partials land in partial[block].
spmv_status_t adaptive_build(const int* d_rp, int R, int nb, hipStream_t s, std::vector<int>& blk_row,
                             std::vector<int>& blk_chunk, std::vector<int>& long_row, std::vector<int>& long_first) {
    std::vector<int> rp((size_t)R + 1);
    MUXI_HIP(hipMemcpyAsync(rp.data(), d_rp, sizeof(int) * ((size_t)R + 1), hipMemcpyDeviceToHost, s));
    MUXI_HIP(hipStreamSynchronize(s));
    blk_row.clear();
    blk_chunk.clear();
    long_row.clear();
    long_first.clear();
    int r = 0;
    while (r < R) {
        const int len = rp[r + 1] - rp[r];
        if (len > nb) {
            long_row.push_back(r);
            long_first.push_back((int)blk_row.size());
            const int nch = (len + nb - 1) / nb;
            for (int c = 0; c < nch; ++c) {
                blk_row.push_back(r);
                blk_chunk.push_back(c);
            }
            long_first.push_back((int)blk_row.size());
            ++r;
            continue;
        }
        const int start = r;
        long nnz = 0;
        while (r < R && r - start < nb && nnz + (rp[r + 1] - rp[r]) <= nb) {
            nnz += rp[r + 1] - rp[r];
            ++r;
        }
        blk_row.push_back(start);
        blk_chunk.push_back(-1);
    }
    // blk_row[nblocks] = R ends the last stream block; chunk blocks read their
    // row's end from rowPtr, not from blk_row[b+1]. long_first holds
    // [first, last+1) block pairs per long row.
    blk_row.push_back(R);
    return SPMV_SUCCESS;
}

}  // namespace muxi
