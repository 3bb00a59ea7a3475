// Merge-path SpMV for wave64 with LDS staging and deterministic carries.
//
// The merge path interleaves the R row ends with the N nonzeros; diagonal k
// is at coordinate (r, k-r) for the unique r with rowPtr[r]+r <= k <=
// rowPtr[r+1]+r. The partition plan follows the reference exactly
// (compute_merge_partitions_kernel, src/spmv/csr5/spmv_csr5.cu:912-925;
// merge_path_search, :497-533): pos[t] = min(t*ceil(L/P), L), L = R+N.
//
// The SpMV itself is this library's own design (the reference's merge
// kernels, spmv_csr5.cu:543-902, lose row heads and double-count tails):
//  * one 256-lane workgroup per partition of at most 256*IPT path items;
//  * the partition's row ends and products val*x[col] are staged in LDS with
//    coalesced loads;
//  * each lane walks IPT consecutive path items (a binary search in LDS finds
//    its start), a workgroup segmented scan carries partial row sums across
//    lanes, and every row end writes y exactly once (no memset of y);
//  * the partial sum of the row left open at the partition end goes to a
//    carry slot; a fix-up kernel adds each row's carries in partition order.
// Every sum is formed in a fixed order, so results are bitwise reproducible.
#include <climits>

#include "merge_tile.hpp"

namespace muxi {

int merge_chunk_items(int ipt) { return 256 * ipt; }

__global__ void merge_partitions_kernel(int L, int P, int* __restrict__ pos) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > P) return;
    long per = ((long)L + P - 1) / P;
    long k = (long)t * per;
    pos[t] = (int)(k < L ? k : L);
}

spmv_status_t launch_merge_partitions(int R, int N, int P, int* d_pos, hipStream_t s) {
    int L = R + N;
    hipLaunchKernelGGL(merge_partitions_kernel, dim3(grid_for((long)P + 1, 256)), dim3(256), 0, s, L, P, d_pos);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

// r = number of row ends among the first k path items: smallest r in
// [max(0,k-N), min(k,R)] with rowPtr[r+1] >= k-r.
__device__ __forceinline__ int merge_search_global(const int* __restrict__ rp, int R, int N, int k) {
    int lo = k - N > 0 ? k - N : 0;
    int hi = k < R ? k : R;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (rp[mid + 1] <= k - mid - 1)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void merge_coords_kernel(const int* __restrict__ rp, int R, int N, const int* __restrict__ pos, int count,
                                    int* __restrict__ row, int* __restrict__ nnz) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int k = pos[t];
    int r = merge_search_global(rp, R, N, k);
    row[t] = r;
    if (nnz) nnz[t] = k - r;
}

spmv_status_t launch_merge_coords(const int* rp, int R, int N, const int* d_pos, int count, int* d_row, int* d_nnz,
                                  hipStream_t s) {
    if (count <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL(merge_coords_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, rp, R, N, d_pos, count,
                       d_row, d_nnz);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

size_t merge_carry_bytes(int P, spmv_type_t xtype, bool comp) {
    size_t e = xtype == SPMV_TYPE_DOUBLE ? sizeof(double) : (comp ? 2 * sizeof(float) : sizeof(float));
    return (size_t)(P > 0 ? P : 1) * e;
}

template <typename VT, typename XT, typename AT, bool COMP, int IPT>
__global__ __launch_bounds__(256) void merge_spmv_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                         const VT* __restrict__ vals, const XT* __restrict__ x,
                                                         XT* __restrict__ y, const int* __restrict__ pdiag,
                                                         const int* __restrict__ prow, int R, AT alpha, AT beta,
                                                         Acc<AT, COMP>* __restrict__ carry) {
    __shared__ MergeSmem<AT, COMP, IPT> sm;
    const int p = blockIdx.x;
    Acc<AT, COMP> total = merge_tile<VT, XT, AT, COMP, IPT>(sm, rp, ci, vals, x, y, pdiag[p], prow[p], pdiag[p + 1],
                                                            prow[p + 1], alpha, beta);
    if (threadIdx.x == 0) carry[p] = total;
}

// Adds carried partial sums to y. Carry t belongs to row key(t) =
// bounds[t+1] (R when t+1 == nbounds, i.e. no row). Keys never decrease, so
// each row's carries are contiguous; the wave that holds the first carry of
// a row sums the whole run in index order and adds alpha*sum to y once.
template <typename XT, typename AT, bool COMP>
__global__ __launch_bounds__(256) void carry_fixup_kernel(const int* __restrict__ bounds, int nbounds, int count,
                                                          const Acc<AT, COMP>* __restrict__ carry, XT* __restrict__ y,
                                                          int R, AT alpha) {
    using A = Acc<AT, COMP>;
    const int lane = threadIdx.x & 63;
    const long base = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64 * 64;
    if (base >= count) return;
    auto key_of = [&](long t) -> int {
        if (t >= count) return INT_MAX;
        return (t + 1 < nbounds) ? bounds[t + 1] : R;
    };
    const long t = base + lane;
    int key = key_of(t);
    A v;
    if (t < count)
        v = carry[t];
    else
        v.zero();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        A vu = v.shfl_up(d);
        int ku = __shfl_up(key, d);
        if (lane >= d && ku == key) {
            vu.add(v);
            v = vu;
        }
    }
    int knext = __shfl_down(key, 1);
    if (lane == 63) knext = key_of(base + 64);
    const int key0 = __shfl(key, 0);
    const bool first_run_foreign = (base > 0) && (key_of(base - 1) == key0);
    const bool owned = !(first_run_foreign && key == key0);
    if (knext != key && owned && key < R) {
        XT yo = y[key];
        y[key] = (XT)((AT)yo + alpha * v.value());
    }
    // a run that reaches the window end continues into the next windows
    const int k63 = __shfl(key, 63);
    const int kn63 = __shfl(knext, 63);
    const bool own63 = __shfl((int)owned, 63);
    if (kn63 == k63 && own63 && k63 < R) {
        A sum = v.shfl(63);
        for (long nb = base + 64; nb < count; nb += 64) {
            long tt = nb + lane;
            int kk = key_of(tt);
            A c;
            if (kk == k63)
                c = carry[tt];
            else
                c.zero();
            c = group_reduce<64>(c);
            sum.add(c);
            if (__shfl(kk, 63) != k63) break;
        }
        if (lane == 0) {
            XT yo = y[k63];
            y[k63] = (XT)((AT)yo + alpha * sum.value());
        }
    }
}

template <typename XT, typename AT, bool COMP>
spmv_status_t launch_fixup(const int* bounds, int nbounds, int count, const void* carry, XT* y, int R, AT alpha,
                           hipStream_t s) {
    if (count <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL((carry_fixup_kernel<XT, AT, COMP>), dim3(grid_for((long)count, 256)), dim3(256), 0, s, bounds,
                       nbounds, count, (const Acc<AT, COMP>*)carry, y, R, alpha);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template <typename VT, typename XT, typename AT, bool COMP, int IPT>
static spmv_status_t merge_go(const DevCsr& A, const int* d_pos, const int* d_row, int P, void* carry, const XT* x,
                              XT* y, AT alpha, AT beta, hipStream_t s, hipEvent_t ev) {
    hipLaunchKernelGGL((merge_spmv_kernel<VT, XT, AT, COMP, IPT>), dim3(P), dim3(256), 0, s, A.rp, A.ci,
                       (const VT*)A.v, x, y, d_pos, d_row, A.R, alpha, beta, (Acc<AT, COMP>*)carry);
    MUXI_LAUNCH_CHECK();
    if (ev) MUXI_HIP(hipEventRecord(ev, s));
    return launch_fixup<XT, AT, COMP>(d_row, P + 1, P, carry, y, A.R, alpha, s);
}

template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_merge(const DevCsr& A, int ipt, const int* d_pos, const int* d_row, int P, void* carry,
                           const XT* x, XT* y, AT alpha, AT beta, hipStream_t s, hipEvent_t ev) {
    if (A.R <= 0 || P <= 0) return SPMV_SUCCESS;
    switch (ipt) {
        case 2: return merge_go<VT, XT, AT, COMP, 2>(A, d_pos, d_row, P, carry, x, y, alpha, beta, s, ev);
        case 4: return merge_go<VT, XT, AT, COMP, 4>(A, d_pos, d_row, P, carry, x, y, alpha, beta, s, ev);
        case 8: return merge_go<VT, XT, AT, COMP, 8>(A, d_pos, d_row, P, carry, x, y, alpha, beta, s, ev);
        case 16: return merge_go<VT, XT, AT, COMP, 16>(A, d_pos, d_row, P, carry, x, y, alpha, beta, s, ev);
        default: set_error("merge: items per thread must be 2, 4, 8 or 16"); return SPMV_ERROR_INTERNAL;
    }
}

#define INST(VT, XT, AT, COMP)                                                                                 \
    template spmv_status_t launch_merge<VT, XT, AT, COMP>(const DevCsr&, int, const int*, const int*, int, void*, \
                                                          const XT*, XT*, AT, AT, hipStream_t, hipEvent_t);
MUXI_FOR_EACH_COMBO(INST)
#undef INST

template spmv_status_t launch_fixup<float, float, false>(const int*, int, int, const void*, float*, int, float,
                                                         hipStream_t);
template spmv_status_t launch_fixup<float, float, true>(const int*, int, int, const void*, float*, int, float,
                                                        hipStream_t);
template spmv_status_t launch_fixup<double, double, false>(const int*, int, int, const void*, double*, int, double,
                                                           hipStream_t);

}  // namespace muxi
