// Device matrix generators, matrix features and the cache flush used by the
// autotune sweep. Randomness: SplitMix64 (Steele, Lea, Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014) of (seed, index).
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "muxi_spmv/muxi_spmv_tune.h"

namespace muxi {

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Uniform in [0, 1) with 53 random bits; `lane` separates independent streams.
__device__ __forceinline__ double rnd01(uint64_t seed, uint64_t lane, uint64_t i) {
    uint64_t h = splitmix64(seed ^ splitmix64(lane * 0x632be59bd9b4e019ull + i));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

struct GenParams {
    int family, R, C;
    double p0, p1, p2;
    int normal;
    uint64_t seed;
};

// Column range [lo, lo+span) of row r and its length.
__device__ __forceinline__ void row_range(const GenParams& g, long r, long& lo, long& span) {
    if (g.family == MUXI_GEN_BANDED) {
        long hb = (long)g.p0;
        long c = (g.R > 0) ? (long)((double)r * g.C / g.R) : 0;
        long a = c - hb < 0 ? 0 : c - hb;
        long b = c + hb > g.C - 1 ? g.C - 1 : c + hb;
        lo = a;
        span = b >= a ? b - a + 1 : 0;
    } else if (g.family == MUXI_GEN_BLOCKDIAG) {
        long bs = (long)g.p0;
        long a = (r / bs) * bs;
        long b = a + bs < g.C ? a + bs : g.C;
        lo = a;
        span = b > a ? b - a : 0;
    } else {
        lo = 0;
        span = g.C;
    }
}

__global__ void gen_len_kernel(GenParams g, int* __restrict__ len) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > g.R) return;
    if (r == g.R) {
        len[r] = 0;
        return;
    }
    long lo, span;
    row_range(g, r, lo, span);
    double u = rnd01(g.seed, 1, (uint64_t)r);
    long k = 0;
    switch (g.family) {
        case MUXI_GEN_UNIFORM: {
            long a = (long)g.p0, b = (long)g.p1;
            k = a + (long)(u * (double)(b - a + 1));
            break;
        }
        case MUXI_GEN_POWERLAW: {
            double uu = 1.0 - u;  // (0, 1]
            double v = floor(g.p1 * pow(uu, -1.0 / (g.p0 - 1.0)));
            if (!(v < g.p2)) v = g.p2;
            if (v < g.p1) v = g.p1;
            k = (long)v;
            break;
        }
        case MUXI_GEN_BANDED: k = (long)g.p1; break;
        case MUXI_GEN_BLOCKDIAG: k = span; break;
        case MUXI_GEN_POISSON: {
            // inversion by the CDF recurrence p_k = p_{k-1} * mean / k
            double mean = g.p0, p = exp(-mean), F = p;
            long kk = 0, cap = (long)(10.0 * mean + 100.0);
            while (u > F && kk < cap) {
                ++kk;
                p *= mean / (double)kk;
                F += p;
            }
            k = kk;
            break;
        }
        default: k = 0;
    }
    if (k > span) k = span;
    if (k < 0) k = 0;
    len[r] = (int)k;
}

// One wavefront per row; column j of a row of length k is drawn uniformly
// from its stratum [lo + floor(j*span/k), lo + floor((j+1)*span/k)).
template <typename VT>
__global__ void gen_fill_kernel(GenParams g, const int* __restrict__ rp, int* __restrict__ ci, VT* __restrict__ v) {
    const long r = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= g.R) return;
    long lo, span;
    row_range(g, r, lo, span);
    const int b = rp[r], k = rp[r + 1] - b;
    for (int j = lane; j < k; j += 64) {
        long s0 = (long)j * span / k;
        long s1 = (long)(j + 1) * span / k;
        long idx = (long)b + j;
        double u = rnd01(g.seed, 2, (uint64_t)idx);
        long c = lo + s0 + (long)(u * (double)(s1 - s0));
        if (c >= lo + s1) c = lo + s1 - 1;
        ci[idx] = (int)c;
        double val;
        if (g.normal) {
            double u1 = 1.0 - rnd01(g.seed, 3, (uint64_t)idx), u2 = rnd01(g.seed, 4, (uint64_t)idx);
            val = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        } else {
            val = 2.0 * rnd01(g.seed, 3, (uint64_t)idx) - 1.0;
        }
        v[idx] = (VT)val;
    }
}

template <typename VT>
__global__ void gen_vec_kernel(VT* __restrict__ x, long n, uint64_t seed) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = (VT)(2.0 * rnd01(seed, 5, (uint64_t)i) - 1.0);
}

__global__ void rand_keys_kernel(uint64_t* __restrict__ k, int* __restrict__ idx, int n, uint64_t seed, uint64_t lane) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    k[i] = splitmix64(seed ^ splitmix64(lane * 0x632be59bd9b4e019ull + (uint64_t)i));
    idx[i] = (int)i;
}

__global__ void perm_len_kernel(const int* __restrict__ rp, const int* __restrict__ perm, int* __restrict__ len, int R) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > R) return;
    len[i] = i < R ? rp[perm[i] + 1] - rp[perm[i]] : 0;
}

__global__ void invert_kernel(const int* __restrict__ perm, int* __restrict__ inv, int n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = (int)i;
}

template <typename VT>
__global__ void perm_rows_remap_kernel(int R, const int* __restrict__ rp, const int* __restrict__ ci,
                                       const VT* __restrict__ v, const int* __restrict__ rowperm,
                                       const int* __restrict__ colinv, const int* __restrict__ nrp,
                                       int* __restrict__ nci, VT* __restrict__ nv) {
    const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= R) return;
    const int o = rowperm[w], b = rp[o], len = rp[o + 1] - b, d = nrp[w];
    for (int i = lane; i < len; i += 64) {
        nci[d + i] = colinv[ci[b + i]];
        nv[d + i] = v[b + i];
    }
}

// ---- features: per-block partial sums, reduced on the host in block order
constexpr int kFeatBlock = 256;
struct FeatPart {
    double len_sum, len_sq, band_sum, diag_cnt, empty_cnt;
    int len_max, pad;
};

// Row r of the block is row r + row_base of the full matrix, whose diagonal runs through column
// (r + row_base) * cen_scale (cen_scale = C / rows of the full matrix).
__global__ __launch_bounds__(kFeatBlock) void features_kernel(int R, long row_base, double cen_scale,
                                                              const int* __restrict__ rp,
                                                              const int* __restrict__ ci, FeatPart* __restrict__ out) {
    __shared__ double s_d[5][kFeatBlock / 64];
    __shared__ int s_m[kFeatBlock / 64];
    const long r = (long)blockIdx.x * kFeatBlock + threadIdx.x;
    double ls = 0, lq = 0, bs = 0, dc = 0, ec = 0;
    int mx = 0;
    if (r < R) {
        const int b = rp[r], e = rp[r + 1], len = e - b;
        ls = len;
        lq = (double)len * len;
        ec = len == 0;
        mx = len;
        const double cen = (double)(r + row_base) * cen_scale;
        const long rb = (r + row_base) / 32;
        for (int i = b; i < e; ++i) {
            const int c = ci[i];
            bs += fabs((double)c - cen);
            dc += (c / 32) == rb;
        }
    }
    double vals[5] = {ls, lq, bs, dc, ec};
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        double t = vals[q];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) t += __shfl_xor(t, m);
        vals[q] = t;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) mx = max(mx, __shfl_xor(mx, m));
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        for (int q = 0; q < 5; ++q) s_d[q][w] = vals[q];
        s_m[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        FeatPart p{0, 0, 0, 0, 0, 0, 0};
        for (int u = 0; u < kFeatBlock / 64; ++u) {
            p.len_sum += s_d[0][u];
            p.len_sq += s_d[1][u];
            p.band_sum += s_d[2][u];
            p.diag_cnt += s_d[3][u];
            p.empty_cnt += s_d[4][u];
            p.len_max = max(p.len_max, s_m[u]);
        }
        out[blockIdx.x] = p;
    }
}

__global__ void flush_kernel(uint4* __restrict__ buf, size_t n, uint32_t tag) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        buf[i] = make_uint4(tag, (uint32_t)i, tag, (uint32_t)(i >> 32));
}

static spmv_status_t scan_lengths(int* d_len, int* d_rp, int n1, hipStream_t s) {
    size_t tmp = 0;
    MUXI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, d_len, d_rp, n1, s));
    void* d_tmp = nullptr;
    MUXI_HIP(hipMallocAsync(&d_tmp, tmp ? tmp : 16, s));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp, d_len, d_rp, n1, s);
    (void)hipFreeAsync(d_tmp, s);
    if (e != hipSuccess) {
        set_error(std::string("scan: ") + hipGetErrorString(e));
        return SPMV_ERROR_DEVICE;
    }
    return SPMV_SUCCESS;
}

static spmv_status_t random_perm(int n, uint64_t seed, uint64_t lane, int* d_perm, hipStream_t s) {
    uint64_t *k0 = nullptr, *k1 = nullptr;
    int* i0 = nullptr;
    MUXI_HIP(hipMallocAsync(&k0, sizeof(uint64_t) * n, s));
    MUXI_HIP(hipMallocAsync(&k1, sizeof(uint64_t) * n, s));
    MUXI_HIP(hipMallocAsync(&i0, sizeof(int) * n, s));
    hipLaunchKernelGGL(rand_keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, k0, i0, n, seed, lane);
    size_t tmp = 0;
    MUXI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k0, k1, i0, d_perm, n, 0, 64, s));
    void* d_tmp = nullptr;
    MUXI_HIP(hipMallocAsync(&d_tmp, tmp ? tmp : 16, s));
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp, k0, k1, i0, d_perm, n, 0, 64, s);
    (void)hipFreeAsync(d_tmp, s);
    (void)hipFreeAsync(k0, s);
    (void)hipFreeAsync(k1, s);
    (void)hipFreeAsync(i0, s);
    if (e != hipSuccess) {
        set_error(std::string("sort: ") + hipGetErrorString(e));
        return SPMV_ERROR_DEVICE;
    }
    return SPMV_SUCCESS;
}

}  // namespace muxi

using namespace muxi;

extern "C" spmv_status_t muxi_gen_matrix(muxi_gen_family_t family, int R, int C, double p0, double p1, double p2,
                                         int normal, spmv_type_t vt, uint64_t seed, void* stream, muxi_csr_t* out) {
    if (!out || R < 0 || C <= 0) {
        set_error("gen: invalid arguments");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (vt != SPMV_TYPE_FLOAT && vt != SPMV_TYPE_DOUBLE) return SPMV_ERROR_UNSUPPORTED_TYPE;
    if ((family == MUXI_GEN_BANDED || family == MUXI_GEN_BLOCKDIAG) && p0 < 1) return SPMV_ERROR_INVALID_MATRIX;
    if (family == MUXI_GEN_POWERLAW && !(p0 > 1.0 && p1 >= 1)) return SPMV_ERROR_INVALID_MATRIX;
    hipStream_t s = (hipStream_t)stream;
    GenParams g{(int)family, R, C, p0, p1, p2, normal, seed};
    *out = muxi_csr_t{};
    out->numRows = R;
    out->numCols = C;
    int* d_len = nullptr;
    MUXI_HIP(hipMalloc(&out->d_rowPtr, sizeof(int) * ((size_t)R + 1)));
    MUXI_HIP(hipMallocAsync(&d_len, sizeof(int) * ((size_t)R + 1), s));
    hipLaunchKernelGGL(gen_len_kernel, dim3(grid_for((long)R + 1, 256)), dim3(256), 0, s, g, d_len);
    MUXI_LAUNCH_CHECK();
    spmv_status_t st = scan_lengths(d_len, out->d_rowPtr, R + 1, s);
    (void)hipFreeAsync(d_len, s);
    if (st != SPMV_SUCCESS) return st;
    int N = 0;
    MUXI_HIP(hipMemcpyAsync(&N, out->d_rowPtr + R, sizeof(int), hipMemcpyDeviceToHost, s));
    MUXI_HIP(hipStreamSynchronize(s));
    if (N < 0) {
        set_error("gen: nnz overflows int32");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    out->nnz = N;
    const size_t es = vt == SPMV_TYPE_DOUBLE ? 8 : 4;
    MUXI_HIP(hipMalloc(&out->d_colIdx, sizeof(int) * (size_t)(N > 0 ? N : 1)));
    MUXI_HIP(hipMalloc(&out->d_values, es * (size_t)(N > 0 ? N : 1)));
    out->ownsDeviceMemory = true;
    if (R > 0) {
        dim3 grid(grid_for((long)R * 64, 256));
        if (vt == SPMV_TYPE_DOUBLE)
            hipLaunchKernelGGL(gen_fill_kernel<double>, grid, dim3(256), 0, s, g, out->d_rowPtr, out->d_colIdx,
                               (double*)out->d_values);
        else
            hipLaunchKernelGGL(gen_fill_kernel<float>, grid, dim3(256), 0, s, g, out->d_rowPtr, out->d_colIdx,
                               (float*)out->d_values);
        MUXI_LAUNCH_CHECK();
    }
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_gen_permute(muxi_csr_t* m, spmv_type_t vt, uint64_t seed, void* stream) {
    if (!m || !m->d_rowPtr || m->numRows != m->numCols) {
        set_error("gen_permute: needs a square generated matrix");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    hipStream_t s = (hipStream_t)stream;
    const int R = m->numRows, N = m->nnz;
    const size_t es = vt == SPMV_TYPE_DOUBLE ? 8 : 4;
    int *rperm, *cperm, *cinv, *len, *nrp, *nci;
    void* nv;
    MUXI_HIP(hipMallocAsync(&rperm, sizeof(int) * (R + 1), s));
    MUXI_HIP(hipMallocAsync(&cperm, sizeof(int) * (R + 1), s));
    MUXI_HIP(hipMallocAsync(&cinv, sizeof(int) * (R + 1), s));
    MUXI_HIP(hipMallocAsync(&len, sizeof(int) * (R + 1), s));
    MUXI_HIP(hipMalloc(&nrp, sizeof(int) * (R + 1)));
    MUXI_HIP(hipMalloc(&nci, sizeof(int) * (size_t)(N > 0 ? N : 1)));
    MUXI_HIP(hipMalloc(&nv, es * (size_t)(N > 0 ? N : 1)));
    spmv_status_t st = random_perm(R, seed, 11, rperm, s);
    if (st == SPMV_SUCCESS) st = random_perm(R, seed, 12, cperm, s);
    if (st != SPMV_SUCCESS) return st;
    hipLaunchKernelGGL(invert_kernel, dim3(grid_for(R, 256)), dim3(256), 0, s, cperm, cinv, R);
    hipLaunchKernelGGL(perm_len_kernel, dim3(grid_for((long)R + 1, 256)), dim3(256), 0, s, m->d_rowPtr, rperm, len, R);
    MUXI_LAUNCH_CHECK();
    if ((st = scan_lengths(len, nrp, R + 1, s)) != SPMV_SUCCESS) return st;
    dim3 grid(grid_for((long)R * 64, 256));
    if (vt == SPMV_TYPE_DOUBLE)
        hipLaunchKernelGGL(perm_rows_remap_kernel<double>, grid, dim3(256), 0, s, R, m->d_rowPtr, m->d_colIdx,
                           (const double*)m->d_values, rperm, cinv, nrp, nci, (double*)nv);
    else
        hipLaunchKernelGGL(perm_rows_remap_kernel<float>, grid, dim3(256), 0, s, R, m->d_rowPtr, m->d_colIdx,
                           (const float*)m->d_values, rperm, cinv, nrp, nci, (float*)nv);
    MUXI_LAUNCH_CHECK();
    MUXI_HIP(hipStreamSynchronize(s));
    (void)hipFreeAsync(rperm, s);
    (void)hipFreeAsync(cperm, s);
    (void)hipFreeAsync(cinv, s);
    (void)hipFreeAsync(len, s);
    MUXI_HIP(hipStreamSynchronize(s));
    if (m->ownsDeviceMemory) {
        (void)hipFree(m->d_rowPtr);
        (void)hipFree(m->d_colIdx);
        (void)hipFree(m->d_values);
    }
    m->d_rowPtr = nrp;
    m->d_colIdx = nci;
    m->d_values = nv;
    m->ownsDeviceMemory = true;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_gen_free(muxi_csr_t* m) {
    if (!m) return SPMV_ERROR_INVALID_MATRIX;
    if (m->ownsDeviceMemory) {
        (void)hipFree(m->d_rowPtr);
        (void)hipFree(m->d_colIdx);
        (void)hipFree(m->d_values);
    }
    m->d_rowPtr = m->d_colIdx = nullptr;
    m->d_values = nullptr;
    m->ownsDeviceMemory = false;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_gen_vector(void* x, spmv_type_t t, long n, uint64_t seed, void* stream) {
    if (n <= 0) return SPMV_SUCCESS;
    if (!x) return SPMV_ERROR_INVALID_VECTOR;
    unsigned g = grid_for(n, 256);
    if (g > 65536) g = 65536;
    if (t == SPMV_TYPE_DOUBLE)
        hipLaunchKernelGGL(gen_vec_kernel<double>, dim3(g), dim3(256), 0, (hipStream_t)stream, (double*)x, n, seed);
    else if (t == SPMV_TYPE_FLOAT)
        hipLaunchKernelGGL(gen_vec_kernel<float>, dim3(g), dim3(256), 0, (hipStream_t)stream, (float*)x, n, seed);
    else
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

spmv_status_t muxi::matrix_features_rows(const muxi_csr_t* A, void* stream, long row_base, long rows_full,
                                         double* out) {
    if (!A || !out || A->numRows < 0 || (A->numRows > 0 && !A->d_rowPtr)) return SPMV_ERROR_INVALID_MATRIX;
    if (row_base < 0 || rows_full < row_base + A->numRows) return SPMV_ERROR_INVALID_MATRIX;
    const int R = A->numRows;
    for (int i = 0; i < 8; ++i) out[i] = 0;
    out[0] = R;
    out[1] = A->nnz;
    if (R == 0) return SPMV_SUCCESS;
    hipStream_t s = (hipStream_t)stream;
    const int nb = (int)grid_for(R, kFeatBlock);
    FeatPart* d_part = nullptr;
    MUXI_HIP(hipMallocAsync(&d_part, sizeof(FeatPart) * nb, s));
    hipLaunchKernelGGL(features_kernel, dim3(nb), dim3(kFeatBlock), 0, s, R, row_base,
                       (double)A->numCols / (double)rows_full, A->d_rowPtr, A->d_colIdx, d_part);
    MUXI_LAUNCH_CHECK();
    std::vector<FeatPart> h(nb);
    MUXI_HIP(hipMemcpyAsync(h.data(), d_part, sizeof(FeatPart) * nb, hipMemcpyDeviceToHost, s));
    (void)hipFreeAsync(d_part, s);
    MUXI_HIP(hipStreamSynchronize(s));
    double ls = 0, lq = 0, bs = 0, dc = 0, ec = 0;
    int mx = 0;
    for (const FeatPart& p : h) {
        ls += p.len_sum;
        lq += p.len_sq;
        bs += p.band_sum;
        dc += p.diag_cnt;
        ec += p.empty_cnt;
        mx = p.len_max > mx ? p.len_max : mx;
    }
    const double mean = ls / R;
    out[2] = mean;
    out[3] = lq / R - mean * mean;
    out[4] = mx;
    out[5] = A->nnz > 0 ? bs / A->nnz / (A->numCols > 0 ? A->numCols : 1) : 0.0;
    if (row_base != 0 || R != rows_full) {
        // A row block: scale the band by (1/3) / E, E = the band uniformly random columns would give for rows
        // at t in [t0, t1) of the full diagonal (mean over t of (t^2 + (1 - t)^2) / 2), so that a random block
        // reads 1/3 like a random square matrix wherever it sits, and a banded block stays near 0.
        const double t0 = (double)row_base / rows_full, t1 = (double)(row_base + R) / rows_full;
        const double u0 = 1.0 - t0, u1 = 1.0 - t1;
        const double E = ((t1 * t1 * t1 - t0 * t0 * t0) - (u1 * u1 * u1 - u0 * u0 * u0)) / (6.0 * (t1 - t0));
        if (E > 0) out[5] *= (1.0 / 3.0) / E;
    }
    out[6] = A->nnz > 0 ? dc / A->nnz : 0.0;
    out[7] = ec / R;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_matrix_features(const muxi_csr_t* A, void* stream, double* out) {
    if (!A) return SPMV_ERROR_INVALID_MATRIX;
    return matrix_features_rows(A, stream, 0, A->numRows, out);
}

extern "C" spmv_status_t muxi_flush_caches(void* buf, size_t bytes, void* stream) {
    if (!buf || bytes < 16) return SPMV_ERROR_INVALID_VECTOR;
    static uint32_t tag = 1;
    hipLaunchKernelGGL(flush_kernel, dim3(4096), dim3(256), 0, (hipStream_t)stream, (uint4*)buf, bytes / 16, tag++);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}
