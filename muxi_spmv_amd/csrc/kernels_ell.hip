// ELLPACK build and SpMV.
//
// Layout follows csr_to_ellpack (src/spmv/ellpack/spmv_ellpack.cu:11-70):
// width = max row length unless given; slot (row, i) lives at row + i*R
// (column-major, so lane-adjacent rows are address-adjacent for every i);
// padding is col = -1, value = 0; rows longer than the width are truncated.
// Unlike the reference, the packing runs on the device from the CSR device
// arrays, one lane per row, instead of a host loop plus blocking copies.
#include <type_traits>

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

// ---- the plan's ELL layout: a row's slots grouped by four --------------------
// Slots 4g .. 4g+3 of row r sit together at 4*g*R + 4*r, so a lane loads its
// row's four columns with one 16-byte load (and its four values with one
// vector load); after the n4 = width/4 groups come a pair (2 slots at
// base2 + 2r, when width % 4 >= 2) and a single (base1 + r, when width is
// odd). Lane-adjacent rows stay address-adjacent within every group, so the
// x gathers of one slot still come from neighbouring rows.
struct EllGroups {
    int n4, two, one;
    long R, base2, base1;
};

inline EllGroups ell_groups(int R, int width) {
    EllGroups g;
    g.R = R;
    g.n4 = width / 4;
    g.two = (width % 4) >= 2;
    g.one = width % 2;
    g.base2 = 4L * R * g.n4;
    g.base1 = g.base2 + (g.two ? 2L * R : 0);
    return g;
}

__device__ __forceinline__ long ellg_index(const EllGroups& g, long r, int i) {
    if (i < 4 * g.n4) return 4L * g.R * (i / 4) + 4 * r + (i % 4);
    if (g.two && i < 4 * g.n4 + 2) return g.base2 + 2 * r + (i - 4 * g.n4);
    return g.base1 + r;
}

// Column indices of the grouped layout come in two widths: int (the column
// itself, padding -1) or, when every stored column lies within 32767 of its
// row (banded and stencil matrices), a 16-bit offset col - row with padding
// ELL_PAD16. The offsets halve the index stream, which for a 5-point stencil
// in FP64 is a quarter of the bytes a SpMV moves.
constexpr short ELL_PAD16 = -32768;

__device__ __forceinline__ int ell_col(int c, long) { return c; }
__device__ __forceinline__ int ell_col(short c, long row) { return c == ELL_PAD16 ? -1 : (int)row + c; }
__device__ __forceinline__ void ell_put(int* p, int col, long) { *p = col; }
__device__ __forceinline__ void ell_put(short* p, int col, long row) {
    *p = col < 0 ? ELL_PAD16 : (short)(col - row);
}

// *bad = 1 when a column among the first `width` of some row lies further
// than 32767 from its row (the 16-bit offsets cannot hold it). Row r of a
// row block is row r + rb of its full matrix; offsets are taken from that.
// out[1] = the largest |column - row| (capped at 32768), which sizes the x window of the windowed kernels.
__global__ void ell_offsets_fit_kernel(int R, int width, int rb, const int* __restrict__ rp,
                                       const int* __restrict__ ci, int* __restrict__ bad) {
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int m = 0;
    if (row < R) {
        const int s0 = rp[row], n = rp[row + 1] - s0, len = n < width ? n : width;
        for (int i = 0; i < len && m <= 32767; ++i) {
            long d = (long)ci[s0 + i] - (row + rb);
            d = d < 0 ? -d : d;
            m = d > m ? (d > 32768 ? 32768 : (int)d) : m;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const int t = __shfl_xor(m, o);
        m = t > m ? t : m;
    }
    if ((threadIdx.x & 63) == 0) {
        if (m > 32767) atomicOr(bad, 1);
        atomicMax(bad + 1, m);
    }
}

spmv_status_t launch_ell_offsets_fit(const DevCsr& A, int width, int rb, int* d_bad, hipStream_t s) {
    MUXI_HIP(hipMemsetAsync(d_bad, 0, 2 * sizeof(int), s));
    if (A.R <= 0 || width <= 0) return SPMV_SUCCESS;
    hipLaunchKernelGGL(ell_offsets_fit_kernel, dim3(grid_for(A.R, 256)), dim3(256), 0, s, A.R, width, rb, A.rp,
                       A.ci, d_bad);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template <typename VT, typename CT>
__global__ void csr_to_ellg_kernel(int R, int width, int rb, EllGroups eg, const int* __restrict__ rp,
                                   const int* __restrict__ ci, const VT* __restrict__ v, CT* __restrict__ ecol,
                                   VT* __restrict__ eval) {
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= R) return;
    const int b = rp[row], len = rp[row + 1] - b;
    for (int i = 0; i < width; ++i) {
        const long idx = ellg_index(eg, row, i);
        if (i < len) {
            ell_put(ecol + idx, ci[b + i], row + rb);
            eval[idx] = v[b + i];
        } else {
            ell_put(ecol + idx, -1, row + rb);
            VT z;
            __builtin_memset(&z, 0, sizeof(VT));
            eval[idx] = z;
        }
    }
}

template <typename VT>
spmv_status_t launch_csr_to_ellg(const DevCsr& A, int width, bool off16, int rb, void* d_col, VT* d_val,
                                 hipStream_t s) {
    if (A.R <= 0 || width <= 0) return SPMV_SUCCESS;
    if (off16)
        hipLaunchKernelGGL((csr_to_ellg_kernel<VT, short>), dim3(grid_for(A.R, 256)), dim3(256), 0, s, A.R, width,
                           rb, ell_groups(A.R, width), A.rp, A.ci, (const VT*)A.v, (short*)d_col, d_val);
    else
        hipLaunchKernelGGL((csr_to_ellg_kernel<VT, int>), dim3(grid_for(A.R, 256)), dim3(256), 0, s, A.R, width,
                           rb, ell_groups(A.R, width), A.rp, A.ci, (const VT*)A.v, (int*)d_col, d_val);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

// N consecutive elements with one non-temporal vector load.
template <typename T, int N>
struct Vec {
    T v[N];
};
template <typename T, int N>
__device__ __forceinline__ Vec<T, N> ldv(const T* p) {
    typedef T V __attribute__((ext_vector_type(N)));
    const V w = __builtin_nontemporal_load((const V*)p);
    Vec<T, N> q;
#pragma unroll
    for (int j = 0; j < N; ++j) q.v[j] = w[j];
    return q;
}
template <int N>
__device__ __forceinline__ Vec<bf16_t, N> ldv_bf16(const bf16_t* p) {
    typedef uint16_t V __attribute__((ext_vector_type(N)));
    const V w = __builtin_nontemporal_load((const V*)p);
    Vec<bf16_t, N> q;
#pragma unroll
    for (int j = 0; j < N; ++j) q.v[j].bits = w[j];
    return q;
}
template <typename T, int N>
__device__ __forceinline__ Vec<T, N> ldvec(const T* p) {
    if constexpr (std::is_same<T, bf16_t>::value) return ldv_bf16<N>(p);
    else return ldv<T, N>(p);
}

// One lane per row over the grouped layout: per round two groups of four
// slots (plus the pair and single in the last round), all column and value
// loads first, then all gathers, then the products. Padding (col < 0) gathers
// x[0] and adds nothing. CT is the stored column type (int or short offsets).
template <typename VT, typename XT, typename AT, bool COMP, typename CT, bool W>
__global__ __launch_bounds__(256) void ellg_spmv_kernel(int R, int C, int rb, int win, EllGroups eg,
                                                        const CT* __restrict__ ecol,
                                                        const VT* __restrict__ eval, const XT* __restrict__ x,
                                                        XT* __restrict__ y, AT alpha, AT beta) {
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // W: the block's rows read x only within win of their whole-matrix rows; that window goes to LDS first
    extern __shared__ unsigned char ell_lds[];
    XT* sxw = reinterpret_cast<XT*>(ell_lds);
    long w0 = 0;
    if constexpr (W) {
        const long first = (long)blockIdx.x * blockDim.x + rb;
        w0 = first - win > 0 ? first - win : 0;
        const long w1 = first + blockDim.x + win < C ? first + blockDim.x + win : C;
        for (long i = w0 + threadIdx.x; i < w1; i += blockDim.x) sxw[i - w0] = x[i];
        __syncthreads();
    }
    if (row >= R) return;
    auto gx = [&](int c) -> XT {
        if constexpr (W) return sxw[c >= w0 ? c - w0 : 0];
        else return x[c > 0 ? c : 0];
    };
    Acc<AT, COMP> a;
    a.zero();
    if (C > 0) {
        for (int g = 0;; g += 2) {
            const bool hasA = g < eg.n4, hasB = g + 1 < eg.n4, last = g + 2 >= eg.n4;
            const bool has2 = last && eg.two, has1 = last && eg.one;
            Vec<int, 4> cA{{-1, -1, -1, -1}}, cB{{-1, -1, -1, -1}};
            Vec<VT, 4> vA{}, vB{};
            Vec<int, 2> c2{{-1, -1}};
            Vec<VT, 2> v2{};
            int c1 = -1;
            VT v1{};
            if (hasA) {
                const Vec<CT, 4> w = ldvec<CT, 4>(ecol + 4L * eg.R * g + 4 * row);
#pragma unroll
                for (int j = 0; j < 4; ++j) cA.v[j] = ell_col(w.v[j], row + rb);
                vA = ldvec<VT, 4>(eval + 4L * eg.R * g + 4 * row);
            }
            if (hasB) {
                const Vec<CT, 4> w = ldvec<CT, 4>(ecol + 4L * eg.R * (g + 1) + 4 * row);
#pragma unroll
                for (int j = 0; j < 4; ++j) cB.v[j] = ell_col(w.v[j], row + rb);
                vB = ldvec<VT, 4>(eval + 4L * eg.R * (g + 1) + 4 * row);
            }
            if (has2) {
                const Vec<CT, 2> w = ldvec<CT, 2>(ecol + eg.base2 + 2 * row);
#pragma unroll
                for (int j = 0; j < 2; ++j) c2.v[j] = ell_col(w.v[j], row + rb);
                v2 = ldvec<VT, 2>(eval + eg.base2 + 2 * row);
            }
            if (has1) {
                c1 = ell_col(ntl(ecol + eg.base1 + row), row + rb);
                v1 = ntl(eval + eg.base1 + row);
            }
            // gathers only for the segments present (uniform branches: no gather instruction is wasted)
            XT xA[4] = {}, xB[4] = {}, x2[2] = {}, x1 = XT(0);
            if (hasA) {
#pragma unroll
                for (int j = 0; j < 4; ++j) xA[j] = gx(cA.v[j]);
            }
            if (hasB) {
#pragma unroll
                for (int j = 0; j < 4; ++j) xB[j] = gx(cB.v[j]);
            }
            if (has2) {
#pragma unroll
                for (int j = 0; j < 2; ++j) x2[j] = gx(c2.v[j]);
            }
            if (has1) x1 = gx(c1);
            if (hasA) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (cA.v[j] >= 0) a.fma(vA.v[j], xA[j]);
            }
            if (hasB) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (cB.v[j] >= 0) a.fma(vB.v[j], xB[j]);
            }
            if (has2) {
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (c2.v[j] >= 0) a.fma(v2.v[j], x2[j]);
            }
            if (has1 && c1 >= 0) a.fma(v1, x1);
            if (last) break;
        }
    }
    store_y(y, row, a.value(), alpha, beta);
}

// Rows of 3 or 4 groups (width 12 to 19): every column and value load of the
// row is issued at once, then every gather, then the products in slot order
// (the same order as the round-based kernel, so the sums are the same): one
// dependent load -> gather step instead of one per round, which is what a
// small, latency-bound matrix waits on.
template <typename VT, typename XT, typename AT, bool COMP, typename CT, bool W>
__global__ __launch_bounds__(256) void ellg_spmv_short_kernel(int R, int C, int rb, int win, EllGroups eg,
                                                              const CT* __restrict__ ecol,
                                                              const VT* __restrict__ eval, const XT* __restrict__ x,
                                                              XT* __restrict__ y, AT alpha, AT beta) {
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // W: the block's rows read x only within win of their whole-matrix rows; that window goes to LDS first
    extern __shared__ unsigned char ell_lds[];
    XT* sxw = reinterpret_cast<XT*>(ell_lds);
    long w0 = 0;
    if constexpr (W) {
        const long first = (long)blockIdx.x * blockDim.x + rb;
        w0 = first - win > 0 ? first - win : 0;
        const long w1 = first + blockDim.x + win < C ? first + blockDim.x + win : C;
        for (long i = w0 + threadIdx.x; i < w1; i += blockDim.x) sxw[i - w0] = x[i];
        __syncthreads();
    }
    if (row >= R) return;
    auto gx = [&](int c) -> XT {
        if constexpr (W) return sxw[c >= w0 ? c - w0 : 0];
        else return x[c > 0 ? c : 0];
    };
    Acc<AT, COMP> a;
    a.zero();
    if (C > 0) {
        Vec<int, 4> c[4];
        Vec<VT, 4> v[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            if (g < eg.n4) {
                const Vec<CT, 4> w = ldvec<CT, 4>(ecol + 4L * eg.R * g + 4 * row);
#pragma unroll
                for (int j = 0; j < 4; ++j) c[g].v[j] = ell_col(w.v[j], row + rb);
                v[g] = ldvec<VT, 4>(eval + 4L * eg.R * g + 4 * row);
            }
        }
        Vec<int, 2> c2{{-1, -1}};
        Vec<VT, 2> v2{};
        int c1 = -1;
        VT v1{};
        if (eg.two) {
            const Vec<CT, 2> w = ldvec<CT, 2>(ecol + eg.base2 + 2 * row);
#pragma unroll
            for (int j = 0; j < 2; ++j) c2.v[j] = ell_col(w.v[j], row + rb);
            v2 = ldvec<VT, 2>(eval + eg.base2 + 2 * row);
        }
        if (eg.one) {
            c1 = ell_col(ntl(ecol + eg.base1 + row), row + rb);
            v1 = ntl(eval + eg.base1 + row);
        }
        XT xs[4][4], x2[2] = {}, x1 = XT(0);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            if (g < eg.n4) {
#pragma unroll
                for (int j = 0; j < 4; ++j) xs[g][j] = gx(c[g].v[j]);
            }
        }
        if (eg.two) {
#pragma unroll
            for (int j = 0; j < 2; ++j) x2[j] = gx(c2.v[j]);
        }
        if (eg.one) x1 = gx(c1);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            if (g < eg.n4) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (c[g].v[j] >= 0) a.fma(v[g].v[j], xs[g][j]);
            }
        }
        if (eg.two) {
#pragma unroll
            for (int j = 0; j < 2; ++j)
                if (c2.v[j] >= 0) a.fma(v2.v[j], x2[j]);
        }
        if (eg.one && c1 >= 0) a.fma(v1, x1);
    }
    store_y(y, row, a.value(), alpha, beta);
}

template <typename VT, typename XT, typename AT, bool COMP, typename CT, bool W>
static void launch_ellg_t(int R, int C, int rb, int win, const EllGroups& eg, const CT* col, const VT* val,
                          const XT* x, XT* y, AT alpha, AT beta, hipStream_t s) {
    // every load of a row at once where the round-based kernel needs two rounds (3 or 4 groups); with fewer
    // groups it has one round and fewer registers (the 5-point stencil in FP64 runs 9 % slower in this one)
    const bool shortrow = eg.n4 == 3 || eg.n4 == 4;
    const size_t lds = W ? sizeof(XT) * (256 + 2 * (size_t)win) : 0;
    if (shortrow)
        hipLaunchKernelGGL((ellg_spmv_short_kernel<VT, XT, AT, COMP, CT, W>), dim3(grid_for(R, 256)), dim3(256), lds,
                           s, R, C, rb, win, eg, col, val, x, y, alpha, beta);
    else
        hipLaunchKernelGGL((ellg_spmv_kernel<VT, XT, AT, COMP, CT, W>), dim3(grid_for(R, 256)), dim3(256), lds, s, R,
                           C, rb, win, eg, col, val, x, y, alpha, beta);
}

template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_ellg(int R, int C, int width, bool off16, int rb, int win, const void* d_col, const VT* d_val,
                          const XT* x, XT* y, AT alpha, AT beta, hipStream_t s) {
    if (R <= 0) return SPMV_SUCCESS;
    const EllGroups eg = ell_groups(R, width);
    if (off16 && win >= 0)
        launch_ellg_t<VT, XT, AT, COMP, short, true>(R, C, rb, win, eg, (const short*)d_col, d_val, x, y, alpha,
                                                     beta, s);
    else if (off16)
        launch_ellg_t<VT, XT, AT, COMP, short, false>(R, C, rb, 0, eg, (const short*)d_col, d_val, x, y, alpha,
                                                      beta, s);
    else
        launch_ellg_t<VT, XT, AT, COMP, int, false>(R, C, rb, 0, eg, (const int*)d_col, d_val, x, y, alpha, beta,
                                                    s);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template spmv_status_t launch_csr_to_ellg<float>(const DevCsr&, int, bool, int, void*, float*, hipStream_t);
template spmv_status_t launch_csr_to_ellg<double>(const DevCsr&, int, bool, int, void*, double*, hipStream_t);
template spmv_status_t launch_csr_to_ellg<bf16_t>(const DevCsr&, int, bool, int, void*, bf16_t*, hipStream_t);

template spmv_status_t launch_csr_to_ell<float>(const DevCsr&, int, int*, float*, hipStream_t);
template spmv_status_t launch_csr_to_ell<double>(const DevCsr&, int, int*, double*, hipStream_t);
template spmv_status_t launch_csr_to_ell<bf16_t>(const DevCsr&, int, int*, bf16_t*, hipStream_t);

#define INST(VT, XT, AT, COMP)                                                                                       \
    template spmv_status_t launch_ell<VT, XT, AT, COMP>(int, int, int, const int*, const VT*, const XT*, XT*, AT, AT, \
                                                        hipStream_t);                                                    \
    template spmv_status_t launch_ellg<VT, XT, AT, COMP>(int, int, int, bool, int, int, const void*, const VT*,     \
                                                         const XT*, XT*, AT, AT, hipStream_t);
MUXI_FOR_EACH_COMBO(INST)
#undef INST

}  // namespace muxi
