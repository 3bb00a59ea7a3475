// Learned kernel selection for MUXI_KERNEL_AUTO.
//
// A selector is an MLP (weights, input normalisation and the candidate list
// it scores). For an AUTO plan the library computes the matrix features on
// the device, forms the model input, runs the MLP with the HIP forward
// kernels, masks candidates that the autotune sweep would not run on this
// matrix, and takes the candidate with the smallest predicted log-time.
// When the build found trained weights (selector_weights.inc, generated from
// the package's selector_weights.json) they are the selector until the caller
// installs another one or clears it; with none, a fixed rule decides.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "muxi_spmv/muxi_spmv_tune.h"
#include "plan.hpp"

#if __has_include("selector_weights.inc")
#include "selector_weights.inc"
#define MUXI_BUILTIN_SELECTOR 1
#else
#define MUXI_BUILTIN_SELECTOR 0
#endif

using namespace muxi;

namespace {

struct Selector {
    int n_in = 0, h = 0, n_out = 0;
    std::vector<float> mu, sd;
    std::vector<int> kernel, param;
    float* d_params = nullptr;
    float* d_in = nullptr;
    float* d_out = nullptr;
    void* d_work = nullptr;
};

std::mutex g_mu;
Selector* g_sel = nullptr;
// which selector decides: the built-in one (installed on first use), one the
// caller set, or none (fixed rule)
int g_state = MUXI_BUILTIN_SELECTOR ? 1 : 0;

void free_sel(Selector* s) {
    if (!s) return;
    (void)hipFree(s->d_params);
    (void)hipFree(s->d_in);
    (void)hipFree(s->d_out);
    (void)hipFree(s->d_work);
    delete s;
}

}  // namespace

// Model inputs from the eight features and the x / y precision (same transform
// as tune.feature_vector): ten size and shape inputs, then 1 for FP64 x and y
// (read only by 11-input models).
static void feature_input(const double* f, bool fp64, float* out) {
    const double rows = f[0], nnz = f[1], mean = f[2], var = f[3], mx = f[4];
    out[0] = (float)std::log1p(rows);
    out[1] = (float)std::log1p(nnz);
    out[2] = (float)std::log1p(mean);
    out[3] = (float)std::log1p(std::sqrt(var > 0 ? var : 0.0));
    out[4] = (float)std::log1p(mx);
    out[5] = (float)f[5];
    out[6] = (float)f[6];
    out[7] = (float)f[7];
    out[8] = (float)std::log1p(mx / (mean > 1e-9 ? mean : 1e-9));
    out[9] = (float)(std::log1p(nnz) - std::log1p(rows));
    out[10] = fp64 ? 1.0f : 0.0f;
}

// The candidates the sweep times on a matrix (muxi_spmv_amd/tune.py, feasible): ELL within 4x padding, one lane or
// one wavefront per row only up to 20,000 nonzeros a row, LDS column slabs only with FP32 x and 65,536 to
// 8,388,608 columns.
static bool candidate_allowed(int kernel, const double* f, spmv_type_t xtype, long cols) {
    const double rows = f[0], nnz = f[1], mx = f[4];
    if (kernel == MUXI_KERNEL_ELL) return nnz > 0 && rows * mx <= 4.0 * nnz && rows * mx < 2147483648.0;
    if (kernel == MUXI_KERNEL_SCALAR || kernel == MUXI_KERNEL_VECTOR) return mx <= 20000;
    if (kernel == MUXI_KERNEL_LSLAB)
        return xtype == SPMV_TYPE_FLOAT && cols >= (1L << 16) && cols <= kLsSelectorMaxCols;
    return true;
}

static spmv_status_t make_selector(const float* params, int n_in, int hidden, int n_out, const float* mu,
                                   const float* sd, const int* kernels, const int* params0, Selector** out) {
    if (!params || !mu || !sd || !kernels || !params0 || (n_in != 10 && n_in != 11) || hidden <= 0 || n_out <= 0) {
        set_error("selector: invalid model (expects 10 inputs, or 11 with the FP64 flag)");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    Selector* s = new Selector();
    s->n_in = n_in;
    s->h = hidden;
    s->n_out = n_out;
    s->mu.assign(mu, mu + n_in);
    s->sd.assign(sd, sd + n_in);
    s->kernel.assign(kernels, kernels + n_out);
    s->param.assign(params0, params0 + n_out);
    const long np = muxi_mlp_param_count(n_in, hidden, n_out);
    if (hipMalloc(&s->d_params, sizeof(float) * np) != hipSuccess ||
        hipMalloc(&s->d_in, sizeof(float) * n_in) != hipSuccess ||
        hipMalloc(&s->d_out, sizeof(float) * n_out) != hipSuccess ||
        hipMalloc(&s->d_work, muxi_mlp_workspace_bytes(1, n_in, hidden, n_out)) != hipSuccess ||
        hipMemcpy(s->d_params, params, sizeof(float) * np, hipMemcpyHostToDevice) != hipSuccess) {
        free_sel(s);
        set_error("selector: device allocation failed");
        return SPMV_ERROR_MEMORY_ALLOC;
    }
    *out = s;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_selector_set(const float* params, int n_in, int hidden, int n_out, const float* mu,
                                           const float* sd, const int* kernels, const int* params0) {
    Selector* s = nullptr;
    spmv_status_t st = make_selector(params, n_in, hidden, n_out, mu, sd, kernels, params0, &s);
    if (st != SPMV_SUCCESS) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    free_sel(g_sel);
    g_sel = s;
    g_state = 2;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_selector_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    free_sel(g_sel);
    g_sel = nullptr;
    g_state = 0;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_selector_reset(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    free_sel(g_sel);
    g_sel = nullptr;
    g_state = MUXI_BUILTIN_SELECTOR ? 1 : 0;
    return SPMV_SUCCESS;
}

extern "C" int muxi_selector_state(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_state;
}

// Fixed rule without a model: merge-path, or column slabs (default slab count,
// 4 items per lane) when x does not fit one XCD's L2 (4 MiB) and the rows are
// short on average.
static void fallback_rule(const double* f, double x_bytes, int* kernel, int* param0) {
    if (x_bytes > 4.0 * 1024 * 1024 && f[2] < 64) {
        *kernel = MUXI_KERNEL_XSLAB;
        *param0 = 4;
    } else {
        *kernel = MUXI_KERNEL_MERGE;
        *param0 = 4;
    }
}

spmv_status_t muxi::selector_pick_rows(const muxi_csr_t* A, spmv_type_t xtype, void* stream, long row_base,
                                       long rows_full, int* kernel, int* param0) {
    if (!A || !kernel || !param0) return SPMV_ERROR_INVALID_MATRIX;
    double f[8];
    spmv_status_t st = matrix_features_rows(A, stream, row_base, rows_full, f);
    if (st != SPMV_SUCCESS) return st;
    std::lock_guard<std::mutex> lk(g_mu);
#if MUXI_BUILTIN_SELECTOR
    if (g_state == 1 && !g_sel) {  // first use of the built-in selector on this process
        st = make_selector(kSelParams, kSelIn, kSelHidden, kSelOut, kSelMu, kSelSd, kSelKernel, kSelParam0, &g_sel);
        if (st != SPMV_SUCCESS) return st;
    }
#endif
    if (!g_sel) {
        fallback_rule(f, (double)A->numCols * (xtype == SPMV_TYPE_DOUBLE ? 8 : 4), kernel, param0);
        return SPMV_SUCCESS;
    }
    Selector& s = *g_sel;
    float in[11];
    feature_input(f, xtype == SPMV_TYPE_DOUBLE, in);
    for (int i = 0; i < s.n_in; ++i) in[i] = (in[i] - s.mu[i]) / s.sd[i];
    hipStream_t hs = (hipStream_t)stream;
    MUXI_HIP(hipMemcpyAsync(s.d_in, in, sizeof(float) * s.n_in, hipMemcpyHostToDevice, hs));
    st = muxi_mlp_forward(s.d_params, s.n_in, s.h, s.n_out, s.d_in, 1, s.d_out, s.d_work, stream);
    if (st != SPMV_SUCCESS) return st;
    std::vector<float> out(s.n_out);
    MUXI_HIP(hipMemcpyAsync(out.data(), s.d_out, sizeof(float) * s.n_out, hipMemcpyDeviceToHost, hs));
    MUXI_HIP(hipStreamSynchronize(hs));
    int best = -1;
    for (int i = 0; i < s.n_out; ++i) {
        if (!candidate_allowed(s.kernel[i], f, xtype, A->numCols)) continue;
        if (best < 0 || out[i] < out[best]) best = i;
    }
    if (best < 0) {
        fallback_rule(f, (double)A->numCols * (xtype == SPMV_TYPE_DOUBLE ? 8 : 4), kernel, param0);
        return SPMV_SUCCESS;
    }
    *kernel = s.kernel[best];
    *param0 = s.param[best];
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_selector_pick(const muxi_csr_t* A, spmv_type_t xtype, void* stream, int* kernel,
                                            int* param0) {
    if (!A) return SPMV_ERROR_INVALID_MATRIX;
    return selector_pick_rows(A, xtype, stream, 0, A->numRows, kernel, param0);
}
