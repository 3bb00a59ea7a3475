// Kernel-selector MLP: in -> h -> h -> out with ReLU, FP32, forward,
// backward of a masked mean squared error, and Adam (Kingma & Ba 2015).
// The layers are tiny (h = 64), so each output element is one lane's dot
// product; sums run over a fixed index order, so training is reproducible.
#include "common.hpp"
#include "muxi_spmv/muxi_spmv_tune.h"

namespace muxi {

struct MlpLayout {
    long w1, b1, w2, b2, w3, b3, total;
};

static MlpLayout mlp_layout(int in, int h, int out) {
    MlpLayout L;
    L.w1 = 0;
    L.b1 = L.w1 + (long)h * in;
    L.w2 = L.b1 + h;
    L.b2 = L.w2 + (long)h * h;
    L.w3 = L.b2 + h;
    L.b3 = L.w3 + (long)out * h;
    L.total = L.b3 + out;
    return L;
}

// Y[b][o] = act(b[o] + sum_i X[b][i] W[o][i])
__global__ void linear_fwd_kernel(const float* __restrict__ X, const float* __restrict__ W,
                                  const float* __restrict__ bias, float* __restrict__ Y, int B, int I, int O, int relu) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * O) return;
    const int b = (int)(t / O), o = (int)(t % O);
    float s = bias[o];
    for (int i = 0; i < I; ++i) s = fmaf(X[(long)b * I + i], W[(long)o * I + i], s);
    Y[t] = relu ? fmaxf(s, 0.f) : s;
}

// dX[b][i] = (sum_o dY[b][o] W[o][i]) * (H[b][i] > 0)
__global__ void linear_bwd_x_kernel(const float* __restrict__ dY, const float* __restrict__ W,
                                    const float* __restrict__ H, float* __restrict__ dX, int B, int I, int O) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * I) return;
    const int b = (int)(t / I), i = (int)(t % I);
    float s = 0.f;
    for (int o = 0; o < O; ++o) s = fmaf(dY[(long)b * O + o], W[(long)o * I + i], s);
    dX[t] = H[t] > 0.f ? s : 0.f;
}

// dW[o][i] = sum_b dY[b][o] Xin[b][i];  db[o] = sum_b dY[b][o]  (thread (o, i); i == I -> bias)
__global__ void linear_bwd_w_kernel(const float* __restrict__ dY, const float* __restrict__ Xin,
                                    float* __restrict__ dW, float* __restrict__ db, int B, int I, int O) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)O * (I + 1)) return;
    const int o = (int)(t / (I + 1)), i = (int)(t % (I + 1));
    float s = 0.f;
    if (i < I) {
        for (int b = 0; b < B; ++b) s = fmaf(dY[(long)b * O + o], Xin[(long)b * I + i], s);
        dW[(long)o * I + i] = s;
    } else {
        for (int b = 0; b < B; ++b) s += dY[(long)b * O + o];
        db[o] = s;
    }
}

// Masked MSE: L = sum M (Y-T)^2 / max(1, sum M);  dY = 2 M (Y-T) / max(1, sum M).
// One workgroup; fixed-order tree reduction.
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                                  const float* __restrict__ M, float* __restrict__ dY,
                                                  float* __restrict__ loss, long n) {
    __shared__ float s_l[256], s_m[256];
    float l = 0.f, m = 0.f;
    for (long i = threadIdx.x; i < n; i += 256) {
        float d = Y[i] - T[i];
        l += M[i] * d * d;
        m += M[i];
    }
    s_l[threadIdx.x] = l;
    s_m[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            s_l[threadIdx.x] += s_l[threadIdx.x + st];
            s_m[threadIdx.x] += s_m[threadIdx.x + st];
        }
        __syncthreads();
    }
    const float cnt = fmaxf(s_m[0], 1.f);
    if (threadIdx.x == 0) loss[0] = s_l[0] / cnt;
    for (long i = threadIdx.x; i < n; i += 256) dY[i] = 2.f * M[i] * (Y[i] - T[i]) / cnt;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

struct MlpWork {
    float *h1, *h2, *y, *dy, *dh2, *dh1;
};

static MlpWork work_of(void* w, int B, int h, int out) {
    float* f = (float*)w;
    MlpWork W;
    W.h1 = f;
    W.h2 = W.h1 + (long)B * h;
    W.y = W.h2 + (long)B * h;
    W.dy = W.y + (long)B * out;
    W.dh2 = W.dy + (long)B * out;
    W.dh1 = W.dh2 + (long)B * h;
    return W;
}

static spmv_status_t fwd_all(const float* P, int in, int h, int out, const float* X, int B, float* Y, MlpWork& W,
                             hipStream_t s) {
    MlpLayout L = mlp_layout(in, h, out);
    hipLaunchKernelGGL(linear_fwd_kernel, dim3(grid_for((long)B * h, 256)), dim3(256), 0, s, X, P + L.w1, P + L.b1,
                       W.h1, B, in, h, 1);
    hipLaunchKernelGGL(linear_fwd_kernel, dim3(grid_for((long)B * h, 256)), dim3(256), 0, s, W.h1, P + L.w2,
                       P + L.b2, W.h2, B, h, h, 1);
    hipLaunchKernelGGL(linear_fwd_kernel, dim3(grid_for((long)B * out, 256)), dim3(256), 0, s, W.h2, P + L.w3,
                       P + L.b3, Y, B, h, out, 0);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

}  // namespace muxi

using namespace muxi;

extern "C" long muxi_mlp_param_count(int in, int hidden, int out) { return mlp_layout(in, hidden, out).total; }

extern "C" size_t muxi_mlp_workspace_bytes(int B, int in, int h, int out) {
    (void)in;
    return sizeof(float) * ((size_t)B * h * 4 + (size_t)B * out * 2);
}

extern "C" spmv_status_t muxi_mlp_forward(const float* P, int in, int h, int out, const float* X, int B, float* Y,
                                          void* work, void* stream) {
    if (!P || !X || !Y || !work || B <= 0 || in <= 0 || h <= 0 || out <= 0) return SPMV_ERROR_INVALID_VECTOR;
    MlpWork W = work_of(work, B, h, out);
    return fwd_all(P, in, h, out, X, B, Y, W, (hipStream_t)stream);
}

extern "C" spmv_status_t muxi_mlp_backward(const float* P, int in, int h, int out, const float* X, int B,
                                           const float* T, const float* M, float* G, float* loss, void* work,
                                           void* stream) {
    if (!P || !X || !T || !M || !G || !loss || !work || B <= 0) return SPMV_ERROR_INVALID_VECTOR;
    hipStream_t s = (hipStream_t)stream;
    MlpWork W = work_of(work, B, h, out);
    MlpLayout L = mlp_layout(in, h, out);
    spmv_status_t st = fwd_all(P, in, h, out, X, B, W.y, W, s);
    if (st != SPMV_SUCCESS) return st;
    hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(256), 0, s, W.y, T, M, W.dy, loss, (long)B * out);
    // layer 3
    hipLaunchKernelGGL(linear_bwd_w_kernel, dim3(grid_for((long)out * (h + 1), 256)), dim3(256), 0, s, W.dy, W.h2,
                       G + L.w3, G + L.b3, B, h, out);
    hipLaunchKernelGGL(linear_bwd_x_kernel, dim3(grid_for((long)B * h, 256)), dim3(256), 0, s, W.dy, P + L.w3, W.h2,
                       W.dh2, B, h, out);
    // layer 2
    hipLaunchKernelGGL(linear_bwd_w_kernel, dim3(grid_for((long)h * (h + 1), 256)), dim3(256), 0, s, W.dh2, W.h1,
                       G + L.w2, G + L.b2, B, h, h);
    hipLaunchKernelGGL(linear_bwd_x_kernel, dim3(grid_for((long)B * h, 256)), dim3(256), 0, s, W.dh2, P + L.w2, W.h1,
                       W.dh1, B, h, h);
    // layer 1
    hipLaunchKernelGGL(linear_bwd_w_kernel, dim3(grid_for((long)h * (in + 1), 256)), dim3(256), 0, s, W.dh1, X,
                       G + L.w1, G + L.b1, B, in, h);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_mlp_adam(float* P, const float* G, float* m, float* v, long n, float lr, float b1,
                                       float b2, float eps, int t, void* stream) {
    if (!P || !G || !m || !v || n <= 0 || t < 1) return SPMV_ERROR_INVALID_VECTOR;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, P, G, m, v, n, lr, b1,
                       b2, eps, c1, c2);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}
