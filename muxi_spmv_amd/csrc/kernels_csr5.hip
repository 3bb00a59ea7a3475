// CSR5-style nnz tiles for wave64.
//
// Tile descriptors follow the reference bit for bit
// (csr5_compute_tile_boundaries_kernel, src/spmv/csr5/spmv_csr5.cu:36-58, with
// csr5_binary_search_row, src/spmv/csr5/spmv_csr5.cuh:44-61):
//   tile_row_ptr[t]    = last r in [0, R) with rowPtr[r] <= t*sigma
//   tile_nnz_offset[t] = t*sigma - rowPtr[tile_row_ptr[t]]
//
// The SpMV kernel is this library's own: one wavefront per tile of sigma =
// 64*EPL nonzeros, lane l owning EPL consecutive nonzeros (one wide load
// each for values and columns). Tile t writes y for exactly the rows whose
// last nonzero lies in the tile (plus the empty rows that sit there), found
// from rowPtr staged in LDS; a wave segmented scan moves partial sums across
// lanes, and the row left open at the tile end is carried to the shared
// fix-up kernel. No atomics, no memset of y; results are reproducible.
#include <climits>

#include "kernels.hpp"

namespace muxi {

template <typename XT, typename AT, bool COMP>
spmv_status_t launch_fixup(const int* bounds, int nbounds, int count, const void* carry, XT* y, int R, AT alpha,
                           hipStream_t s);

__global__ void csr5_tiles_kernel(const int* __restrict__ rp, int R, int N, int sigma, int T,
                                  int* __restrict__ tile_row, int* __restrict__ tile_off) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    long start = (long)t * sigma;
    // last r in [0, R) with rowPtr[r] <= start
    int lo = 0, hi = R;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((long)rp[mid] <= start)
            lo = mid + 1;
        else
            hi = mid;
    }
    int row = lo - 1;
    tile_row[t] = row;
    tile_off[t] = (int)(start - rp[row]);
}

spmv_status_t launch_csr5_tiles(const int* rp, int R, int N, int sigma, int T, int* d_tile_row, int* d_tile_off,
                                hipStream_t s) {
    if (T <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL(csr5_tiles_kernel, dim3(grid_for(T, 256)), dim3(256), 0, s, rp, R, N, sigma, T, d_tile_row,
                       d_tile_off);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

bool csr5_sigma_supported(int sigma) {
    return sigma == 64 || sigma == 128 || sigma == 256 || sigma == 512 || sigma == 1024;
}

size_t csr5_carry_bytes(int T, spmv_type_t xtype, bool comp) {
    size_t e = xtype == SPMV_TYPE_DOUBLE ? sizeof(double) : (comp ? 2 * sizeof(float) : sizeof(float));
    return (size_t)(T > 0 ? T : 1) * e;
}

template <typename T, int N>
struct alignas(sizeof(T) * N) VecN {
    T v[N];
};

template <typename VT, typename XT, typename AT, bool COMP, int EPL>
__global__ __launch_bounds__(256) void csr5_spmv_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                        const VT* __restrict__ vals, const XT* __restrict__ x,
                                                        XT* __restrict__ y, const int* __restrict__ tile_row, int T,
                                                        int N, int R, AT alpha, AT beta,
                                                        Acc<AT, COMP>* __restrict__ carry, int vec) {
    constexpr int SIG = 64 * EPL;
    constexpr int CAP = SIG + 64;  // staged row pointers per wave
    using A = Acc<AT, COMP>;
    using P = Prod<AT, COMP>;
    __shared__ int s_rp[4][CAP];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + w;
    if (t >= T) return;  // whole wave leaves; no workgroup barrier below
    const long ts = (long)t * SIG;
    const long te = ts + SIG < N ? ts + SIG : N;
    const int ra = tile_row[t];                       // row holding nonzero ts
    const int rb = (t + 1 < T) ? tile_row[t + 1] : R;  // first row not closed in this tile
    // rows [ra, rb] need rowPtr[ra .. rb] (rb's start for the open row)
    const int cnt = (rb < R ? rb : R) - ra + 1;
    const bool staged = cnt <= CAP;
    int* srp = s_rp[w];
    auto RP = [&](int r) -> int { return staged ? srp[r - ra] : rp[r]; };

    // Issue the lane's column and value loads first, then stage the row
    // pointers, then the x gathers, so the three latencies overlap. Lanes past
    // the matrix end load a clamped (valid) element and drop its product;
    // no load sits behind a per-element branch.
    const long e0 = ts + (long)lane * EPL;
    const int nv = e0 < te ? (int)((te - e0) < EPL ? (te - e0) : EPL) : 0;
    const long last = te - 1;
    int cv[EPL];
    VT vv[EPL];
    if (vec && nv == EPL) {
        const VecN<int, EPL> c4 = *(const VecN<int, EPL>*)(ci + e0);
        const VecN<VT, EPL> v4 = *(const VecN<VT, EPL>*)(vals + e0);
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            cv[j] = c4.v[j];
            vv[j] = v4.v[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            const long i = e0 + j < last ? e0 + j : last;
            cv[j] = ci[i];
            vv[j] = vals[i];
        }
    }
    if (staged) {
        for (int i = lane; i < cnt; i += 64) srp[i] = rp[ra + i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    P pr[EPL];
    {
        XT xv[EPL];
#pragma unroll
        for (int j = 0; j < EPL; ++j) xv[j] = x[cv[j]];
#pragma unroll
        for (int j = 0; j < EPL; ++j) pr[j] = (j < nv) ? P::make(vv[j], xv[j]) : P::zero();
    }
    // lane's first row: last r in [ra, min(rb, R-1)] with rowPtr[r] <= e0
    int rf = ra;
    if (nv > 0) {
        int lo = ra, hi = (rb < R ? rb : R - 1);
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (RP(mid) <= e0)
                lo = mid;
            else
                hi = mid - 1;
        }
        rf = lo;
    }
    // pass 1
    A acc;
    acc.zero();
    int fl = 0;
    {
        int r = rf;
        int nxt = (nv > 0 && r < rb) ? RP(r + 1) : INT_MAX;
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            if (j < nv) {
                pr[j].into(acc);
                long e1 = e0 + j + 1;
                if (e1 == nxt) {
                    fl = 1;
                    acc.zero();
                    ++r;
                    while (r < rb && RP(r + 1) == e1) ++r;
                    nxt = (r < rb) ? RP(r + 1) : INT_MAX;
                }
            }
        }
    }
    // wave segmented scan
    A v = acc;
    int f = fl;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        A vu = v.shfl_up(d);
        int fu = __shfl_up(f, d);
        if (lane >= d) {
            if (!f) {
                vu.add(v);
                v = vu;
            }
            f |= fu;
        }
    }
    A ex = v.shfl_up(1);
    if (lane == 0) ex.zero();
    // leading empty rows of the matrix
    if (t == 0 && lane == 0) {
        for (int r = 0; r < ra; ++r) store_y(y, r, AT(0), alpha, beta);
    }
    // pass 2
    {
        A a = ex;
        int r = rf;
        int nxt = (nv > 0 && r < rb) ? RP(r + 1) : INT_MAX;
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            if (j < nv) {
                pr[j].into(a);
                long e1 = e0 + j + 1;
                if (e1 == nxt) {
                    store_y(y, r, a.value(), alpha, beta);
                    a.zero();
                    ++r;
                    while (r < rb && RP(r + 1) == e1) {
                        store_y(y, r, AT(0), alpha, beta);
                        ++r;
                    }
                    nxt = (r < rb) ? RP(r + 1) : INT_MAX;
                }
            }
        }
    }
    if (lane == 63) carry[t] = v;
}

template <typename VT, typename XT, typename AT, bool COMP, int EPL>
static spmv_status_t csr5_go(const DevCsr& A, const int* d_tile_row, int T, void* carry, const XT* x, XT* y, AT alpha,
                             AT beta, hipStream_t s, hipEvent_t ev) {
    // wide loads need the arrays aligned to one lane's EPL elements
    const int vec = ((uintptr_t)A.ci % (sizeof(int) * EPL) == 0) && ((uintptr_t)A.v % (sizeof(VT) * EPL) == 0);
    hipLaunchKernelGGL((csr5_spmv_kernel<VT, XT, AT, COMP, EPL>), dim3(grid_for(T, 4)), dim3(256), 0, s, A.rp, A.ci,
                       (const VT*)A.v, x, y, d_tile_row, T, A.N, A.R, alpha, beta, (Acc<AT, COMP>*)carry, vec);
    MUXI_LAUNCH_CHECK();
    if (ev) MUXI_HIP(hipEventRecord(ev, s));
    return launch_fixup<XT, AT, COMP>(d_tile_row, T, T, carry, y, A.R, alpha, s);
}

template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_csr5(const DevCsr& A, int sigma, const int* d_tile_row, int T, void* carry, const XT* x, XT* y,
                          AT alpha, AT beta, hipStream_t s, hipEvent_t ev) {
    if (A.R <= 0 || T <= 0) return SPMV_SUCCESS;
    switch (sigma) {
        case 64: return csr5_go<VT, XT, AT, COMP, 1>(A, d_tile_row, T, carry, x, y, alpha, beta, s, ev);
        case 128: return csr5_go<VT, XT, AT, COMP, 2>(A, d_tile_row, T, carry, x, y, alpha, beta, s, ev);
        case 256: return csr5_go<VT, XT, AT, COMP, 4>(A, d_tile_row, T, carry, x, y, alpha, beta, s, ev);
        case 512: return csr5_go<VT, XT, AT, COMP, 8>(A, d_tile_row, T, carry, x, y, alpha, beta, s, ev);
        case 1024: return csr5_go<VT, XT, AT, COMP, 16>(A, d_tile_row, T, carry, x, y, alpha, beta, s, ev);
        default: set_error("csr5: sigma must be 64, 128, 256, 512 or 1024 for the SpMV kernel"); return SPMV_ERROR_INTERNAL;
    }
}

#define INST(VT, XT, AT, COMP)                                                                                 \
    template spmv_status_t launch_csr5<VT, XT, AT, COMP>(const DevCsr&, int, const int*, int, void*, const XT*, \
                                                         XT*, AT, AT, hipStream_t, hipEvent_t);
MUXI_FOR_EACH_COMBO(INST)
#undef INST

}  // namespace muxi
