// SpMM  Y = alpha*A*X + beta*Y  with X (C x k) and Y (R x k) row-major.
//
// Merge-path partitions as in the SpMV kernel, but each merge item is worked
// by a group of G lanes, every lane owning 4 consecutive columns of X/Y, so
// one nonzero costs the group one contiguous k-wide read of X (float4 per
// lane when k % 4 == 0). Per partition (256 lanes = 256/G groups, IPT items
// per group): column indices, values and row ends are staged in LDS; each
// group walks its items once, writes every row it closes except its first,
// keeps that first row's partial in registers, and a workgroup segmented
// scan over groups supplies the sum carried into it. The open row at the
// partition end goes to a k-wide carry slot, added by the fix-up kernel in
// partition order: every sum has a fixed order, results are reproducible.
#include <climits>

#include "kernels.hpp"

namespace muxi {

template <typename XT>
struct V4 {
    XT v[4];
    __device__ __forceinline__ void zero() { v[0] = v[1] = v[2] = v[3] = XT(0); }
};

template <typename XT, bool VEC>
__device__ __forceinline__ void load_x4(const XT* __restrict__ X, long row, int k, int c0, XT out[4]) {
    const XT* p = X + row * k + c0;
    if (VEC) {
        if constexpr (sizeof(XT) == 4) {
            float4 t = *(const float4*)p;
            out[0] = t.x; out[1] = t.y; out[2] = t.z; out[3] = t.w;
        } else {
            double2 a = *(const double2*)p;
            double2 b = *(const double2*)(p + 2);
            out[0] = a.x; out[1] = a.y; out[2] = b.x; out[3] = b.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) out[j] = (c0 + j < k) ? p[j] : XT(0);
    }
}

template <typename XT>
__device__ __forceinline__ void store_y4(XT* __restrict__ Y, long row, int k, int c0, const XT a[4], XT alpha,
                                         XT beta) {
    XT* p = Y + row * k + c0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (c0 + j < k) {
            XT r = alpha * a[j];
            if (beta != XT(0)) r += beta * p[j];
            p[j] = r;
        }
}

// Items per partition: (256/G groups) x ipt, capped at 2048 so the LDS
// staging (12 or 16 bytes per item) stays at or under 32 KiB.
__host__ __device__ constexpr int spmm_chunk(int G, int ipt) { return (256 / G) * ipt > 2048 ? 2048 : (256 / G) * ipt; }

template <typename VT, typename XT, int G, int IPT0, bool VEC>
__global__ __launch_bounds__(256) void spmm_merge_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                         const VT* __restrict__ vals, const XT* __restrict__ X,
                                                         XT* __restrict__ Y, int k, const int* __restrict__ pdiag,
                                                         const int* __restrict__ prow, XT alpha, XT beta,
                                                         XT* __restrict__ carry) {
    constexpr int B = 256, NG = B / G, CH = spmm_chunk(G, IPT0), IPT = CH / NG, GPW = 64 / G;
    __shared__ int s_end[CH];
    __shared__ int s_col[CH];
    __shared__ XT s_val[CH];
    __shared__ XT s_agg[4][G][4];
    __shared__ int s_flag[4];

    const int p = blockIdx.x, tid = threadIdx.x;
    const int g = tid / G, gl = tid % G, lane = tid & 63, w = tid >> 6;
    const int c0 = gl * 4;
    const int d0 = pdiag[p], d1 = pdiag[p + 1];
    const int r0 = prow[p], r1 = prow[p + 1];
    const int n0 = d0 - r0, n1 = d1 - r1;
    const int nr = r1 - r0, nn = n1 - n0, tot = nr + nn;
    for (int j = tid; j < nn; j += B) {
        s_col[j] = ci[n0 + j];
        s_val[j] = (XT)(sizeof(XT) == 8 ? to_f64(vals[n0 + j]) : to_f32(vals[n0 + j]));
    }
    for (int i = tid; i < nr; i += B) s_end[i] = rp[r0 + 1 + i] - n0;
    __syncthreads();

    const int di = min(g * IPT, tot), de = min(di + IPT, tot);
    int lo = max(0, di - nn), hi = min(di, nr);
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (s_end[mid] <= di - mid - 1)
            lo = mid + 1;
        else
            hi = mid;
    }
    int r = lo, n = di - lo;
    XT acc[4] = {XT(0), XT(0), XT(0), XT(0)};
    XT first[4] = {XT(0), XT(0), XT(0), XT(0)};
    int first_row = -1;
    for (int it = di; it < de; ++it) {
        if (r < nr && s_end[r] <= n) {
            if (first_row < 0) {
                first_row = r;
#pragma unroll
                for (int j = 0; j < 4; ++j) first[j] = acc[j];
            } else {
                store_y4(Y, (long)r0 + r, k, c0, acc, alpha, beta);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] = XT(0);
            ++r;
        } else {
            XT xv[4];
            load_x4<XT, VEC>(X, s_col[n], k, c0, xv);
            const XT v = s_val[n];
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] += v * xv[j];
            ++n;
        }
    }
    // segmented scan over groups: (flag, acc) with reset at groups that closed a row
    int f = first_row >= 0;
#pragma unroll
    for (int d = 1; d < GPW; d <<= 1) {
        XT u[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) u[j] = __shfl_up(acc[j], d * G);
        int fu = __shfl_up(f, d * G);
        if (lane >= d * G) {
            if (!f) {
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] = u[j] + acc[j];
            }
            f |= fu;
        }
    }
    // acc/f now inclusive within the wave; publish each wave's last group
    if (lane >= 64 - G) {
#pragma unroll
        for (int j = 0; j < 4; ++j) s_agg[w][gl][j] = acc[j];
        if (gl == 0) s_flag[w] = f;
    }
    __syncthreads();
    XT pre[4] = {XT(0), XT(0), XT(0), XT(0)};
    for (int u = 0; u < w; ++u) {
#pragma unroll
        for (int j = 0; j < 4; ++j) pre[j] = s_flag[u] ? s_agg[u][gl][j] : pre[j] + s_agg[u][gl][j];
    }
    XT inc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) inc[j] = f ? acc[j] : pre[j] + acc[j];
    XT ex[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        XT t = __shfl_up(inc[j], G);
        ex[j] = (lane >= G) ? t : pre[j];
    }
    if (first_row >= 0) {
        XT a[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = ex[j] + first[j];
        store_y4(Y, (long)r0 + first_row, k, c0, a, alpha, beta);
    }
    // workgroup total = inclusive value of the last group
    if (g == NG - 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (c0 + j < k) carry[(long)p * k + c0 + j] = inc[j];
    }
}

// k-wide carry fix-up: blockIdx.y = column. Same run logic as the SpMV fix-up.
template <typename XT>
__global__ __launch_bounds__(256) void spmm_fixup_kernel(const int* __restrict__ bounds, int count,
                                                         const XT* __restrict__ carry, XT* __restrict__ Y, int R,
                                                         int k, XT alpha) {
    const int col = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const long base = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64 * 64;
    if (base >= count) return;
    auto key_of = [&](long t) -> int { return t < count ? bounds[t + 1] : INT_MAX; };
    const long t = base + lane;
    int key = key_of(t);
    XT v = t < count ? carry[t * k + col] : XT(0);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        XT vu = __shfl_up(v, d);
        int ku = __shfl_up(key, d);
        if (lane >= d && ku == key) v = vu + v;
    }
    int knext = __shfl_down(key, 1);
    if (lane == 63) knext = key_of(base + 64);
    const int key0 = __shfl(key, 0);
    const bool owned = !((base > 0) && (key_of(base - 1) == key0) && key == key0);
    if (knext != key && owned && key < R) Y[(long)key * k + col] += alpha * v;
    const int k63 = __shfl(key, 63), kn63 = __shfl(knext, 63);
    const bool own63 = __shfl((int)owned, 63);
    if (kn63 == k63 && own63 && k63 < R) {
        XT sum = __shfl(v, 63);
        for (long nb = base + 64; nb < count; nb += 64) {
            long tt = nb + lane;
            int kk = key_of(tt);
            XT c = (kk == k63) ? carry[tt * k + col] : XT(0);
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
            sum += c;
            if (__shfl(kk, 63) != k63) break;
        }
        if (lane == 0) Y[(long)k63 * k + col] += alpha * sum;
    }
}

int spmm_groups(int k) {
    int need = (k + 3) / 4;
    int g = 1;
    while (g < need) g <<= 1;
    return g;
}

int spmm_chunk_items(int k, int ipt) { return spmm_chunk(spmm_groups(k), ipt); }

template <typename VT, typename XT, int G, int IPT>
static spmv_status_t spmm_go(const DevCsr& A, int k, const int* d_pos, const int* d_row, int P, XT* carry,
                             const XT* X, XT* Y, XT alpha, XT beta, hipStream_t s, hipEvent_t ev) {
    // wide X loads only when the groups cover exactly k columns and X is aligned
    const bool vec = (k == 4 * G) && ((uintptr_t)X % (4 * sizeof(XT)) == 0);
    if (vec)
        hipLaunchKernelGGL((spmm_merge_kernel<VT, XT, G, IPT, true>), dim3(P), dim3(256), 0, s, A.rp, A.ci,
                           (const VT*)A.v, X, Y, k, d_pos, d_row, alpha, beta, carry);
    else
        hipLaunchKernelGGL((spmm_merge_kernel<VT, XT, G, IPT, false>), dim3(P), dim3(256), 0, s, A.rp, A.ci,
                           (const VT*)A.v, X, Y, k, d_pos, d_row, alpha, beta, carry);
    MUXI_LAUNCH_CHECK();
    if (ev) MUXI_HIP(hipEventRecord(ev, s));
    hipLaunchKernelGGL(spmm_fixup_kernel<XT>, dim3(grid_for(P, 256), k), dim3(256), 0, s, d_row, P, carry, Y, A.R,
                       k, alpha);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template <typename VT, typename XT, int G>
static spmv_status_t spmm_ipt(const DevCsr& A, int k, int ipt, const int* d_pos, const int* d_row, int P, XT* carry,
                              const XT* X, XT* Y, XT alpha, XT beta, hipStream_t s, hipEvent_t ev) {
    switch (ipt) {
        case 4: return spmm_go<VT, XT, G, 4>(A, k, d_pos, d_row, P, carry, X, Y, alpha, beta, s, ev);
        case 8: return spmm_go<VT, XT, G, 8>(A, k, d_pos, d_row, P, carry, X, Y, alpha, beta, s, ev);
        case 16: return spmm_go<VT, XT, G, 16>(A, k, d_pos, d_row, P, carry, X, Y, alpha, beta, s, ev);
        case 32: return spmm_go<VT, XT, G, 32>(A, k, d_pos, d_row, P, carry, X, Y, alpha, beta, s, ev);
        default: set_error("spmm: items per group must be 4, 8, 16 or 32"); return SPMV_ERROR_INTERNAL;
    }
}

template <typename VT, typename XT>
spmv_status_t launch_spmm(const DevCsr& A, int k, int ipt, const int* d_pos, const int* d_row, int P, void* carry,
                          const XT* X, XT* Y, XT alpha, XT beta, hipStream_t s, hipEvent_t ev) {
    if (A.R <= 0 || P <= 0 || k <= 0) return SPMV_SUCCESS;
    XT* c = (XT*)carry;
    switch (spmm_groups(k)) {
        case 1: return spmm_ipt<VT, XT, 1>(A, k, ipt, d_pos, d_row, P, c, X, Y, alpha, beta, s, ev);
        case 2: return spmm_ipt<VT, XT, 2>(A, k, ipt, d_pos, d_row, P, c, X, Y, alpha, beta, s, ev);
        case 4: return spmm_ipt<VT, XT, 4>(A, k, ipt, d_pos, d_row, P, c, X, Y, alpha, beta, s, ev);
        case 8: return spmm_ipt<VT, XT, 8>(A, k, ipt, d_pos, d_row, P, c, X, Y, alpha, beta, s, ev);
        case 16: return spmm_ipt<VT, XT, 16>(A, k, ipt, d_pos, d_row, P, c, X, Y, alpha, beta, s, ev);
        default: set_error("spmm: k must be in [1, 64]"); return SPMV_ERROR_INTERNAL;
    }
}

template spmv_status_t launch_spmm<float, float>(const DevCsr&, int, int, const int*, const int*, int, void*,
                                                 const float*, float*, float, float, hipStream_t, hipEvent_t);
template spmv_status_t launch_spmm<bf16_t, float>(const DevCsr&, int, int, const int*, const int*, int, void*,
                                                  const float*, float*, float, float, hipStream_t, hipEvent_t);
template spmv_status_t launch_spmm<double, float>(const DevCsr&, int, int, const int*, const int*, int, void*,
                                                  const float*, float*, float, float, hipStream_t, hipEvent_t);
template spmv_status_t launch_spmm<double, double>(const DevCsr&, int, int, const int*, const int*, int, void*,
                                                   const double*, double*, double, double, hipStream_t, hipEvent_t);

}  // namespace muxi
