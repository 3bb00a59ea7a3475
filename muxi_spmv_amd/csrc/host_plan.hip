// SpMV with x and y in host memory: row chunks with their own device plans,
// x sent in column pieces, y returned chunk by chunk while the next chunk
// computes (three streams, default priority). Reference experiments it
// replaces: tests/test_pinned_memory.cu:191-330, tests/test_multi_stream.cu:121-290.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "common.hpp"
#include "muxi_spmv/muxi_spmv_host.h"
#include "plan.hpp"
#include "muxi_spmv/muxi_spmv_prep.h"

using namespace muxi;

namespace {

// out[i] = rp[i] - rp[0], i in [0, n]: a row range's rowPtr rebased to 0.
__global__ void rebase_rowptr_kernel(const int* __restrict__ rp, int n, int* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) out[i] = rp[i] - rp[0];
}

// Largest column of ci[0, n) into *m (initialised to -1).
__global__ void max_col_kernel(const int* __restrict__ ci, long n, int* __restrict__ m) {
    int v = -1;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        v = max(v, ci[k]);
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0 && v >= 0) atomicMax(m, v);
}

}  // namespace

struct muxi_host_plan {
    int R = 0, C = 0, K = 0;
    spmv_type_t vt = SPMV_TYPE_FLOAT, xt = SPMV_TYPE_FLOAT;
    size_t xs = 4;
    std::vector<int> r0, r1, lo, hi;  // rows of chunk k; x piece [lo, hi) sent before it
    std::vector<muxi_plan*> plans;
    int* d_rp = nullptr;
    int* d_colmax = nullptr;
    void* d_x = nullptr;
    void* d_y = nullptr;
    hipStream_t s_in = nullptr, s_c = nullptr, s_out = nullptr;
    std::vector<hipEvent_t> ev_in, ev_c;
    hipEvent_t t0 = nullptr, tk0 = nullptr, tkend = nullptr, tend = nullptr;
    bool ran = false;
};

static void host_plan_free(muxi_host_plan* h) {
    if (!h) return;
    if (h->s_c) (void)hipStreamSynchronize(h->s_c);
    for (muxi_plan* p : h->plans)
        if (p) (void)muxi_plan_destroy(p);
    (void)hipFree(h->d_rp);
    (void)hipFree(h->d_colmax);
    (void)hipFree(h->d_x);
    (void)hipFree(h->d_y);
    for (hipEvent_t e : h->ev_in)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : h->ev_c)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {h->t0, h->tk0, h->tkend, h->tend})
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t s : {h->s_in, h->s_c, h->s_out})
        if (s) (void)hipStreamDestroy(s);
    delete h;
}

#define HP_HIP(expr)                                                        \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
            host_plan_free(h);                                              \
            return SPMV_ERROR_DEVICE;                                       \
        }                                                                   \
    } while (0)

#define HP_TRY(expr)                   \
    do {                               \
        spmv_status_t st_ = (expr);    \
        if (st_ != SPMV_SUCCESS) {     \
            host_plan_free(h);         \
            return st_;                \
        }                              \
    } while (0)

static spmv_status_t host_plan_create_impl(const muxi_csr_t* A, spmv_type_t vt, spmv_type_t xt, int chunks,
                                               muxi_kernel_t kernel, int param0, muxi_host_plan** out) {
    if (!out) {
        set_error("host plan output pointer is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    *out = nullptr;
    if (!A || A->numRows < 0 || A->numCols < 0 || A->nnz < 0 || (A->numRows > 0 && !A->d_rowPtr) ||
        (A->nnz > 0 && (!A->d_colIdx || !A->d_values))) {
        set_error("invalid matrix (dimensions or device arrays)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (chunks < 1 || chunks > 64) {
        set_error("host plan: chunks must be in [1, 64]");
        return SPMV_ERROR_INTERNAL;
    }
    if (xt != SPMV_TYPE_FLOAT && xt != SPMV_TYPE_DOUBLE) {
        set_error("host plan: x and y must be FLOAT or DOUBLE");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    const size_t vs = vt == SPMV_TYPE_DOUBLE ? 8 : (vt == SPMV_TYPE_BF16 ? 2 : 4);
    muxi_host_plan* h = new (std::nothrow) muxi_host_plan();
    if (!h) return SPMV_ERROR_MEMORY_ALLOC;
    h->R = A->numRows;
    h->C = A->numCols;
    h->K = chunks;
    h->vt = vt;
    h->xt = xt;
    h->xs = xt == SPMV_TYPE_DOUBLE ? 8 : 4;
    const int R = h->R, K = chunks;
    HP_HIP(hipStreamCreate(&h->s_in));
    HP_HIP(hipStreamCreate(&h->s_c));
    HP_HIP(hipStreamCreate(&h->s_out));
    for (hipEvent_t* e : {&h->t0, &h->tk0, &h->tkend, &h->tend}) HP_HIP(hipEventCreate(e));
    h->ev_in.assign(K, nullptr);
    h->ev_c.assign(K, nullptr);
    for (int k = 0; k < K; ++k) {
        HP_HIP(hipEventCreateWithFlags(&h->ev_in[k], hipEventDisableTiming));
        HP_HIP(hipEventCreateWithFlags(&h->ev_c[k], hipEventDisableTiming));
    }
    HP_HIP(hipMalloc(&h->d_x, h->xs * (size_t)(h->C > 0 ? h->C : 1)));
    HP_HIP(hipMalloc(&h->d_y, h->xs * (size_t)(R > 0 ? R : 1)));
    // chunk bounds: equal merge-path work, cut at row starts
    std::vector<int> hrp((size_t)R + 1, 0), bounds((size_t)K + 1, 0);
    if (R > 0) {
        HP_HIP(hipMemcpyAsync(hrp.data(), A->d_rowPtr, sizeof(int) * ((size_t)R + 1), hipMemcpyDeviceToHost, h->s_c));
        HP_HIP(hipStreamSynchronize(h->s_c));
    }
    HP_TRY(muxi_shard_rows(hrp.data(), R, K, bounds.data()));
    h->r0.assign(K, 0);
    h->r1.assign(K, 0);
    h->lo.assign(K, 0);
    h->hi.assign(K, 0);
    h->plans.assign(K, nullptr);
    HP_HIP(hipMalloc(&h->d_rp, sizeof(int) * ((size_t)R + K)));
    HP_HIP(hipMalloc(&h->d_colmax, sizeof(int) * K));
    HP_HIP(hipMemsetAsync(h->d_colmax, 0xff, sizeof(int) * K, h->s_c));  // -1
    long rp_off = 0;
    for (int k = 0; k < K; ++k) {
        const int a = bounds[k], b = bounds[k + 1];
        h->r0[k] = a;
        h->r1[k] = b;
        if (b <= a) continue;  // an empty chunk (more chunks than rows)
        const int n0 = hrp[a], n1 = hrp[b];
        int* rpk = h->d_rp + rp_off;
        rp_off += (long)(b - a) + 1;
        hipLaunchKernelGGL(rebase_rowptr_kernel, dim3(grid_for((long)(b - a) + 1, 256)), dim3(256), 0, h->s_c,
                           A->d_rowPtr + a, b - a, rpk);
        HP_HIP(hipGetLastError());
        if (n1 > n0) {
            unsigned g = grid_for((long)(n1 - n0), 256);
            hipLaunchKernelGGL(max_col_kernel, dim3(g > 2048 ? 2048 : g), dim3(256), 0, h->s_c, A->d_colIdx + n0,
                               (long)(n1 - n0), h->d_colmax + k);
            HP_HIP(hipGetLastError());
        }
        muxi_csr_t sub{};
        sub.numRows = b - a;
        sub.numCols = h->C;
        sub.nnz = n1 - n0;
        sub.d_rowPtr = rpk;
        sub.d_colIdx = A->d_colIdx + n0;
        sub.d_values = (char*)A->d_values + vs * (size_t)n0;
        // a row block of A: AUTO sees its band against A's diagonal, ELL offsets come from A's rows
        HP_TRY(plan_create_rows(&sub, vt, xt, kernel, param0, 0, h->s_c, a, R, &h->plans[k]));
    }
    std::vector<int> cmax(K, -1);
    HP_HIP(hipMemcpyAsync(cmax.data(), h->d_colmax, sizeof(int) * K, hipMemcpyDeviceToHost, h->s_c));
    HP_HIP(hipStreamSynchronize(h->s_c));
    int sent = 0;
    for (int k = 0; k < K; ++k) {
        const int need = std::min(h->C, cmax[k] + 1);  // columns [0, need) cover chunk k
        h->lo[k] = sent;
        h->hi[k] = std::max(sent, need);
        sent = h->hi[k];
    }
    *out = h;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_host_plan_destroy(muxi_host_plan* h) {
    if (!h) {
        set_error("host plan is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    host_plan_free(h);
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_host_plan_exec(muxi_host_plan* h, const void* h_x, void* h_y, double alpha,
                                             double beta) {
    if (!h) {
        set_error("host plan is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    if ((h->C > 0 && !h_x) || (h->R > 0 && !h_y)) {
        set_error("host x or y is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    const size_t xs = h->xs;
    const char* hx = (const char*)h_x;
    char* hy = (char*)h_y;
    char* dx = (char*)h->d_x;
    char* dy = (char*)h->d_y;
    MUXI_HIP(hipEventRecord(h->t0, h->s_in));
    for (int k = 0; k < h->K; ++k) {
        if (h->hi[k] > h->lo[k])
            MUXI_HIP(hipMemcpyAsync(dx + xs * h->lo[k], hx + xs * h->lo[k], xs * (size_t)(h->hi[k] - h->lo[k]),
                                    hipMemcpyHostToDevice, h->s_in));
        const int rows = h->r1[k] - h->r0[k];
        if (beta != 0.0 && rows > 0)
            MUXI_HIP(hipMemcpyAsync(dy + xs * h->r0[k], hy + xs * h->r0[k], xs * (size_t)rows, hipMemcpyHostToDevice,
                                    h->s_in));
        MUXI_HIP(hipEventRecord(h->ev_in[k], h->s_in));
    }
    bool first = true;
    for (int k = 0; k < h->K; ++k) {
        MUXI_HIP(hipStreamWaitEvent(h->s_c, h->ev_in[k], 0));
        if (first) {
            MUXI_HIP(hipEventRecord(h->tk0, h->s_c));
            first = false;
        }
        const int rows = h->r1[k] - h->r0[k];
        if (rows <= 0) continue;
        spmv_status_t st = muxi_plan_exec(h->plans[k], dx, dy + xs * h->r0[k], alpha, beta, 1, h->s_c);
        if (st != SPMV_SUCCESS) return st;
        MUXI_HIP(hipEventRecord(h->ev_c[k], h->s_c));
        MUXI_HIP(hipStreamWaitEvent(h->s_out, h->ev_c[k], 0));
        MUXI_HIP(hipMemcpyAsync(hy + xs * h->r0[k], dy + xs * h->r0[k], xs * (size_t)rows, hipMemcpyDeviceToHost,
                                h->s_out));
    }
    MUXI_HIP(hipEventRecord(h->tkend, h->s_c));
    MUXI_HIP(hipStreamWaitEvent(h->s_out, h->tkend, 0));
    MUXI_HIP(hipEventRecord(h->tend, h->s_out));
    MUXI_HIP(hipStreamSynchronize(h->s_out));
    h->ran = true;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_host_plan_timing(const muxi_host_plan* h, float* total, float* first_in,
                                               float* kernels, float* tail) {
    if (!h) {
        set_error("host plan is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    if (!h->ran) {
        set_error("host plan has not run");
        return SPMV_ERROR_INTERNAL;
    }
    float t[4] = {0, 0, 0, 0};
    MUXI_HIP(hipEventElapsedTime(&t[0], h->t0, h->tend));
    MUXI_HIP(hipEventElapsedTime(&t[1], h->t0, h->tk0));
    MUXI_HIP(hipEventElapsedTime(&t[2], h->tk0, h->tkend));
    MUXI_HIP(hipEventElapsedTime(&t[3], h->tkend, h->tend));
    if (total) *total = t[0];
    if (first_in) *first_in = t[1];
    if (kernels) *kernels = t[2];
    if (tail) *tail = t[3];
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_host_plan_chunk(const muxi_host_plan* h, int k, int* row0, int* row1, int* kernel,
                                              int* param0) {
    if (!h) {
        set_error("host plan is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    if (k < 0 || k >= h->K) {
        set_error("host plan: chunk index out of range");
        return SPMV_ERROR_INTERNAL;
    }
    if (row0) *row0 = h->r0[k];
    if (row1) *row1 = h->r1[k];
    int kk = 0, pp = 0;
    if (h->plans[k]) {
        spmv_status_t st = muxi_plan_info(h->plans[k], &kk, &pp, nullptr);
        if (st != SPMV_SUCCESS) return st;
    }
    if (kernel) *kernel = kk;
    if (param0) *param0 = pp;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_host_alloc(void** ptr, size_t bytes) {
    if (!ptr) return SPMV_ERROR_INVALID_VECTOR;
    *ptr = nullptr;
    MUXI_HIP(hipHostMalloc(ptr, bytes > 0 ? bytes : 1, hipHostMallocDefault));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_host_free(void* ptr) {
    if (ptr) MUXI_HIP(hipHostFree(ptr));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_host_plan_create(const muxi_csr_t* A, spmv_type_t vt, spmv_type_t xt, int chunks,
                                               muxi_kernel_t kernel, int param0, muxi_host_plan** out) {
    return no_throw("muxi_host_plan_create", [&] { return host_plan_create_impl(A, vt, xt, chunks, kernel, param0, out); });
}
