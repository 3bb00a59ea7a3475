// Batched SpMV over many independent CSR matrices in one persistent launch.
//
// Every matrix is cut into merge-path tiles of at most 256*IPT path items
// (the same coordinates as the single-matrix merge kernel). A tile table
// maps each global tile to its matrix; a persistent grid sized to the device
// (a multiple of 8) strides through it, the workgroups of each XCD through
// their own contiguous eighth, running the shared tile body. Carries are
// keyed by a virtual row (each matrix's rows plus one dummy row, laid end to
// end) so that runs never cross matrices, and one fix-up kernel folds them.
#include <algorithm>
#include <climits>
#include <vector>

#include "merge_tile.hpp"
#include "muxi_spmv/muxi_spmv_batch.h"
#include "sched.hpp"

namespace muxi {

struct BatchDesc {
    const int* rp;
    const int* ci;
    const void* v;
    const void* x;
    void* y;
    int R, C, N;
    int tile_base;   // first global tile of this matrix
    int coord_base;  // first of its P+1 coordinates in gdiag/grow
    int vrow_base;   // first virtual row
};

template <typename VT, typename XT, typename AT, int IPT>
__global__ __launch_bounds__(256) void batch_merge_kernel(const BatchDesc* __restrict__ desc,
                                                          const int* __restrict__ tile_mat,
                                                          const int* __restrict__ gdiag, const int* __restrict__ grow,
                                                          int T, AT alpha, AT beta, Acc<AT, false>* __restrict__ carry) {
    __shared__ MergeSmem<AT, false, IPT> sm;
    // Workgroup b runs on XCD b mod 8. Each XCD walks its own contiguous eighth of the tile table (tiles are in
    // matrix order), so at any time an XCD's L2 holds the x of one or two matrices instead of a dozen.
    const int xcd = blockIdx.x % 8, per = gridDim.x / 8;
    const int t1 = (int)((long)T * (xcd + 1) / 8);
    for (int t = (int)((long)T * xcd / 8) + (int)(blockIdx.x / 8); t < t1; t += per) {
        const int m = tile_mat[t];
        const BatchDesc d = desc[m];
        const int c = d.coord_base + (t - d.tile_base);
        Acc<AT, false> total = merge_tile<VT, XT, AT, false, IPT>(sm, d.rp, d.ci, (const VT*)d.v, (const XT*)d.x,
                                                                 (XT*)d.y, gdiag[c], grow[c], gdiag[c + 1],
                                                                 grow[c + 1], alpha, beta);
        if (threadIdx.x == 0) carry[t] = total;
        __syncthreads();  // LDS is reused by the next tile
    }
}

__global__ void batch_keys_kernel(const BatchDesc* __restrict__ desc, const int* __restrict__ tile_mat,
                                  const int* __restrict__ grow, int T, int* __restrict__ key) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const BatchDesc d = desc[tile_mat[t]];
    key[t] = d.vrow_base + grow[d.coord_base + (t - d.tile_base) + 1];
}

// Same run logic as carry_fixup_kernel, with virtual-row keys; a key that is
// a matrix's dummy row (r == R) carries nothing.
template <typename XT, typename AT>
__global__ __launch_bounds__(256) void batch_fixup_kernel(const int* __restrict__ tkey,
                                                          const int* __restrict__ tile_mat,
                                                          const BatchDesc* __restrict__ desc, int count,
                                                          const Acc<AT, false>* __restrict__ carry, AT alpha) {
    using A = Acc<AT, false>;
    const int lane = threadIdx.x & 63;
    const long base = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64 * 64;
    if (base >= count) return;
    auto key_of = [&](long t) -> int { return t < count ? tkey[t] : INT_MAX; };
    const long t = base + lane;
    const int key = key_of(t);
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
    const bool owned = !((base > 0) && (key_of(base - 1) == key0) && key == key0);
    auto add_to = [&](long tt, int k, AT val) {
        const BatchDesc d = desc[tile_mat[tt]];
        const int r = k - d.vrow_base;
        if (r < d.R) {
            XT* y = (XT*)d.y;
            y[r] = (XT)((AT)y[r] + alpha * val);
        }
    };
    if (t < count && knext != key && owned) add_to(t, key, v.value());
    const int k63 = __shfl(key, 63), kn63 = __shfl(knext, 63);
    const bool own63 = __shfl((int)owned, 63);
    if (kn63 == k63 && own63 && k63 != INT_MAX) {
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
        if (lane == 0) add_to(base + 63, k63, sum.value());
    }
}

}  // namespace muxi

using namespace muxi;

struct muxi_batch_plan {
    int count = 0, T = 0, grid = 0, ipt = 8, wg_per_cu = 8;
    spmv_type_t vtype = SPMV_TYPE_FLOAT, xtype = SPMV_TYPE_FLOAT;
    std::vector<BatchDesc> host;
    BatchDesc* d_desc = nullptr;
    int* d_tile_mat = nullptr;
    int* d_tile_key = nullptr;
    int* d_gdiag = nullptr;
    int* d_grow = nullptr;
    void* d_carry = nullptr;
    size_t scratch = 0;
};

struct muxi_graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
};

static void batch_free(muxi_batch_plan* p) {
    if (!p) return;
    (void)hipFree(p->d_desc);
    (void)hipFree(p->d_tile_mat);
    (void)hipFree(p->d_tile_key);
    (void)hipFree(p->d_gdiag);
    (void)hipFree(p->d_carry);
    delete p;
}

#define BT_HIP(expr)                                                      \
    do {                                                                  \
        hipError_t e_ = (expr);                                           \
        if (e_ != hipSuccess) {                                           \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            batch_free(p);                                                \
            return SPMV_ERROR_DEVICE;                                     \
        }                                                                 \
    } while (0)

static spmv_status_t batch_plan_create_impl(const muxi_csr_t* mats, int count, spmv_type_t vt, spmv_type_t xt,
                                                const void* const* xs, void* const* ys, int ipt, void* stream,
                                                muxi_batch_plan** out) {
    if (!out) return SPMV_ERROR_INVALID_HANDLE;
    *out = nullptr;
    if (!mats || count <= 0 || !xs || !ys) {
        set_error("batch: no matrices or vectors");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (!((vt == SPMV_TYPE_FLOAT && xt == SPMV_TYPE_FLOAT) || (vt == SPMV_TYPE_DOUBLE && xt == SPMV_TYPE_DOUBLE))) {
        set_error("batch: FLOAT/FLOAT or DOUBLE/DOUBLE only");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    // ipt: merge items per lane (bits 0-7: 4, 8 or 16) | workgroups per CU << 8 (1 to 32; 0 = 8); ipt <= 0
    // takes both from the tuned schedule table
    int wg_per_cu = ipt > 0 ? (ipt >> 8) & 0xff : 0;
    if (ipt > 0) ipt &= 0xff;
    if (ipt <= 0) {
        long rows = 0, nnz = 0;
        for (int m = 0; m < count; ++m) {
            rows += std::max(0, mats[m].numRows);
            nnz += std::max(0, mats[m].nnz);
        }
        batch_schedule(rows > 0 ? (double)nnz / (double)rows : 0.0, (double)rows / count, &ipt, &wg_per_cu);
    }
    if (wg_per_cu == 0) wg_per_cu = 8;
    if (!(ipt == 4 || ipt == 8 || ipt == 16) || wg_per_cu < 1 || wg_per_cu > 32) {
        set_error("batch: items per lane must be 4, 8 or 16 and workgroups per CU 1 to 32");
        return SPMV_ERROR_INTERNAL;
    }
    hipStream_t s = (hipStream_t)stream;
    muxi_batch_plan* p = new muxi_batch_plan();
    p->count = count;
    p->vtype = vt;
    p->xtype = xt;
    p->ipt = ipt;
    p->wg_per_cu = wg_per_cu;
    p->host.resize(count);
    const long CH = 256L * ipt;
    long T = 0, ncoord = 0, vrows = 0;
    std::vector<int> P(count);
    for (int m = 0; m < count; ++m) {
        const muxi_csr_t& A = mats[m];
        if (A.numRows < 0 || A.nnz < 0 || (long)A.numRows + A.nnz > 2147483647L ||
            (A.numRows > 0 && !A.d_rowPtr) || (A.nnz > 0 && (!A.d_colIdx || !A.d_values)) ||
            (A.numRows > 0 && (!xs[m] || !ys[m]))) {
            set_error("batch: invalid matrix " + std::to_string(m));
            batch_free(p);
            return SPMV_ERROR_INVALID_MATRIX;
        }
        P[m] = A.numRows > 0 ? (int)(((long)A.numRows + A.nnz + CH - 1) / CH) : 0;
        BatchDesc& d = p->host[m];
        d = BatchDesc{A.d_rowPtr, A.d_colIdx, A.d_values, xs[m], ys[m], A.numRows, A.numCols, A.nnz,
                      (int)T, (int)ncoord, (int)vrows};
        T += P[m];
        ncoord += P[m] + 1;
        vrows += (long)A.numRows + 1;
        if (T > INT_MAX / 2 || vrows > INT_MAX / 2) {
            set_error("batch: too many tiles or rows");
            batch_free(p);
            return SPMV_ERROR_INVALID_MATRIX;
        }
    }
    p->T = (int)T;
    int dev = 0, cus = 256;
    BT_HIP(hipGetDevice(&dev));
    BT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // a multiple of 8 workgroups: an equal share for each XCD's part of the tile table
    p->grid = (int)std::max<long>(8, std::min<long>(T, (long)cus * wg_per_cu) / 8 * 8);
    const size_t acc = xt == SPMV_TYPE_DOUBLE ? 8 : 4;
    BT_HIP(hipMalloc(&p->d_desc, sizeof(BatchDesc) * count));
    BT_HIP(hipMalloc(&p->d_tile_mat, sizeof(int) * (size_t)(T > 0 ? T : 1)));
    BT_HIP(hipMalloc(&p->d_tile_key, sizeof(int) * (size_t)(T > 0 ? T : 1)));
    BT_HIP(hipMalloc(&p->d_gdiag, 2 * sizeof(int) * (size_t)ncoord));
    p->d_grow = p->d_gdiag + ncoord;
    BT_HIP(hipMalloc(&p->d_carry, acc * (size_t)(T > 0 ? T : 1)));
    p->scratch = sizeof(BatchDesc) * count + 2 * sizeof(int) * T + 2 * sizeof(int) * ncoord + acc * T;
    BT_HIP(hipMemcpyAsync(p->d_desc, p->host.data(), sizeof(BatchDesc) * count, hipMemcpyHostToDevice, s));
    std::vector<int> tm((size_t)(T > 0 ? T : 1));
    for (int m = 0; m < count; ++m) {
        const BatchDesc& d = p->host[m];
        for (int q = 0; q < P[m]; ++q) tm[d.tile_base + q] = m;
        if (P[m] == 0) continue;
        spmv_status_t st = launch_merge_partitions(d.R, d.N, P[m], p->d_gdiag + d.coord_base, s);
        if (st == SPMV_SUCCESS)
            st = launch_merge_coords(d.rp, d.R, d.N, p->d_gdiag + d.coord_base, P[m] + 1, p->d_grow + d.coord_base,
                                     nullptr, s);
        if (st != SPMV_SUCCESS) {
            batch_free(p);
            return st;
        }
    }
    BT_HIP(hipMemcpyAsync(p->d_tile_mat, tm.data(), sizeof(int) * tm.size(), hipMemcpyHostToDevice, s));
    if (T > 0) hipLaunchKernelGGL(batch_keys_kernel, dim3(grid_for(T, 256)), dim3(256), 0, s, p->d_desc,
                                  p->d_tile_mat, p->d_grow, (int)T, p->d_tile_key);
    BT_HIP(hipGetLastError());
    BT_HIP(hipStreamSynchronize(s));  // host vectors above go out of scope
    *out = p;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_batch_plan_destroy(muxi_batch_plan* p) {
    if (!p) return SPMV_ERROR_INVALID_HANDLE;
    batch_free(p);
    return SPMV_SUCCESS;
}

template <typename VT, typename XT, typename AT>
static spmv_status_t batch_go(const muxi_batch_plan* p, AT alpha, AT beta, hipStream_t s, hipEvent_t ev1) {
    if (p->T > 0) {
        switch (p->ipt) {
            case 4:
                hipLaunchKernelGGL((batch_merge_kernel<VT, XT, AT, 4>), dim3(p->grid), dim3(256), 0, s, p->d_desc,
                                   p->d_tile_mat, p->d_gdiag, p->d_grow, p->T, alpha, beta,
                                   (Acc<AT, false>*)p->d_carry);
                break;
            case 16:
                hipLaunchKernelGGL((batch_merge_kernel<VT, XT, AT, 16>), dim3(p->grid), dim3(256), 0, s, p->d_desc,
                                   p->d_tile_mat, p->d_gdiag, p->d_grow, p->T, alpha, beta,
                                   (Acc<AT, false>*)p->d_carry);
                break;
            default:
                hipLaunchKernelGGL((batch_merge_kernel<VT, XT, AT, 8>), dim3(p->grid), dim3(256), 0, s, p->d_desc,
                                   p->d_tile_mat, p->d_gdiag, p->d_grow, p->T, alpha, beta,
                                   (Acc<AT, false>*)p->d_carry);
        }
        MUXI_LAUNCH_CHECK();
        hipLaunchKernelGGL((batch_fixup_kernel<XT, AT>), dim3(grid_for(p->T, 256)), dim3(256), 0, s, p->d_tile_key,
                           p->d_tile_mat, p->d_desc, p->T, (const Acc<AT, false>*)p->d_carry, alpha);
        MUXI_LAUNCH_CHECK();
    }
    if (ev1) MUXI_HIP(hipEventRecord(ev1, s));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_batch_exec_ev(const muxi_batch_plan* p, double alpha, double beta, void* stream,
                                            void* ev0, void* ev1) {
    if (!p) return SPMV_ERROR_INVALID_HANDLE;
    hipStream_t s = (hipStream_t)stream;
    if (ev0) MUXI_HIP(hipEventRecord((hipEvent_t)ev0, s));
    if (p->xtype == SPMV_TYPE_DOUBLE) return batch_go<double, double, double>(p, alpha, beta, s, (hipEvent_t)ev1);
    return batch_go<float, float, float>(p, (float)alpha, (float)beta, s, (hipEvent_t)ev1);
}

extern "C" spmv_status_t muxi_batch_exec(const muxi_batch_plan* p, double alpha, double beta, void* stream) {
    return muxi_batch_exec_ev(p, alpha, beta, stream, nullptr, nullptr);
}

// One launch per member matrix (the baseline the batched launch is measured against): y_m = alpha A_m x_m +
// beta y_m with the scalar, vector or virtual-wavefront kernel.
template <typename T>
static spmv_status_t batch_per_matrix(const muxi_batch_plan* p, muxi_kernel_t kernel, int param0, T alpha, T beta,
                                      hipStream_t s) {
    for (const BatchDesc& d : p->host) {
        if (d.R == 0) continue;
        DevCsr A{d.R, d.C, d.N, d.rp, d.ci, d.v};
        const T* x = (const T*)d.x;
        T* y = (T*)d.y;
        spmv_status_t st;
        if (kernel == MUXI_KERNEL_SCALAR) st = launch_scalar<T, T, T, false>(A, x, y, alpha, beta, s);
        else if (kernel == MUXI_KERNEL_VWAVE) st = launch_vwave<T, T, T, false>(A, param0, x, y, alpha, beta, s);
        else st = launch_vector<T, T, T, false>(A, x, y, alpha, beta, s);
        if (st != SPMV_SUCCESS) return st;
    }
    return SPMV_SUCCESS;
}

static spmv_status_t batch_per_matrix_ab(const muxi_batch_plan* p, muxi_kernel_t kernel, int param0, double alpha,
                                         double beta, hipStream_t s) {
    if (!p) return SPMV_ERROR_INVALID_HANDLE;
    if (!(kernel == MUXI_KERNEL_SCALAR || kernel == MUXI_KERNEL_VECTOR || kernel == MUXI_KERNEL_VWAVE)) {
        set_error("batch per-matrix launches: kernel must be SCALAR, VECTOR or VWAVE");
        return SPMV_ERROR_INTERNAL;
    }
    if (kernel == MUXI_KERNEL_VWAVE && !(param0 >= 1 && param0 <= 64 && (param0 & (param0 - 1)) == 0)) {
        set_error("batch per-matrix launches: VWAVE lanes per row must be a power of two in [1, 64]");
        return SPMV_ERROR_INTERNAL;
    }
    if (p->xtype == SPMV_TYPE_DOUBLE) return batch_per_matrix<double>(p, kernel, param0, alpha, beta, s);
    return batch_per_matrix<float>(p, kernel, param0, (float)alpha, (float)beta, s);
}

extern "C" spmv_status_t muxi_batch_exec_per_matrix(const muxi_batch_plan* p, muxi_kernel_t kernel, int param0,
                                                    void* stream) {
    return batch_per_matrix_ab(p, kernel, param0, 1.0, 0.0, (hipStream_t)stream);
}

extern "C" spmv_status_t muxi_batch_info(const muxi_batch_plan* p, int* tiles, int* grid, size_t* scratch) {
    if (!p) return SPMV_ERROR_INVALID_HANDLE;
    if (tiles) *tiles = p->T;
    if (grid) *grid = p->grid;
    if (scratch) *scratch = p->scratch;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_batch_schedule(const muxi_batch_plan* p, int* ipt, int* wg_per_cu) {
    if (!p) return SPMV_ERROR_INVALID_HANDLE;
    if (ipt) *ipt = p->ipt;
    if (wg_per_cu) *wg_per_cu = p->wg_per_cu;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_batch_graph_create(const muxi_batch_plan* p, double alpha, double beta,
                                                 int per_matrix_kernel, int param0, void* stream, muxi_graph** out) {
    if (!p || !out) return SPMV_ERROR_INVALID_HANDLE;
    hipStream_t s = (hipStream_t)stream;
    if (!s) {
        set_error("graph capture needs a non-null stream");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    MUXI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    muxi_graph* g = new muxi_graph();
    spmv_status_t st = per_matrix_kernel
                           ? batch_per_matrix_ab(p, (muxi_kernel_t)per_matrix_kernel, param0, alpha, beta, s)
                           : muxi_batch_exec(p, alpha, beta, s);
    hipError_t e = hipStreamEndCapture(s, &g->g);
    if (st != SPMV_SUCCESS || e != hipSuccess) {
        if (g->g) (void)hipGraphDestroy(g->g);
        delete g;
        if (e != hipSuccess) set_error(std::string("capture: ") + hipGetErrorString(e));
        return st != SPMV_SUCCESS ? st : SPMV_ERROR_DEVICE;
    }
    e = hipGraphInstantiate(&g->e, g->g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g->g);
        delete g;
        set_error(std::string("instantiate: ") + hipGetErrorString(e));
        return SPMV_ERROR_DEVICE;
    }
    *out = g;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_graph_launch(const muxi_graph* g, void* stream) {
    if (!g) return SPMV_ERROR_INVALID_HANDLE;
    MUXI_HIP(hipGraphLaunch(g->e, (hipStream_t)stream));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_graph_destroy(muxi_graph* g) {
    if (!g) return SPMV_ERROR_INVALID_HANDLE;
    (void)hipGraphExecDestroy(g->e);
    (void)hipGraphDestroy(g->g);
    delete g;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_batch_plan_create(const muxi_csr_t* mats, int count, spmv_type_t vt, spmv_type_t xt,
                                                const void* const* xs, void* const* ys, int ipt, void* stream,
                                                muxi_batch_plan** out) {
    return no_throw("muxi_batch_plan_create",
                    [&] { return batch_plan_create_impl(mats, count, vt, xt, xs, ys, ipt, stream, out); });
}
