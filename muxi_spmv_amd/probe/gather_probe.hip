// Random-gather throughput probe: how many x[idx] gathers per second the
// memory system of one GPU serves, by the span of x that the indices cover
// and by whether each XCD (workgroup b runs on XCD b mod 8) gathers from its
// own slice of x. This is the ceiling an SpMV with uniformly scattered
// columns meets: every nonzero costs one such gather.
//
//   gather_probe                full table (hash-only gathers, then gathers
//                               with an 8-byte index+value stream per gather)
//   gather_probe MODE SPAN K    one streamed configuration (for counter runs):
//                               MODE 0 global, 1 XCD slice, 2 sequential;
//                               SPAN a power of two (elements); K 4 or 8
//
// Sizes: x holds 64M floats (256 MiB), the streams 64M elements each (512 MiB
// together, more than the Infinity Cache), so only x can stay cached.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

#define CK(x)                                                                          \
    do {                                                                               \
        hipError_t e = (x);                                                            \
        if (e != hipSuccess) {                                                         \
            printf("HIP error %s at %s:%d\n", hipGetErrorString(e), __FILE__, __LINE__); \
            exit(1);                                                                   \
        }                                                                              \
    } while (0)

__device__ __host__ inline uint32_t hsh(uint32_t a) {
    a ^= a >> 16;
    a *= 0x7feb352du;
    a ^= a >> 15;
    a *= 0x846ca68bu;
    a ^= a >> 16;
    return a;
}

// idx[i] in [base, base+span), base = ((i / per_block) & 7) * span in slab mode.
// mode 0 random global, 1 random XCD slab, 2 sequential (i mod span).
__global__ void mk_idx(int* idx, float* val, long n, int span, int per_block, int mode) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    long b = i / per_block;
    int base = mode == 1 ? (int)(b & 7) * span : 0;
    int off = mode == 2 ? (int)(i & (span - 1)) : (int)(hsh((uint32_t)i * 2654435761u + 12345u) & (span - 1));
    idx[i] = base + off;
    val[i] = 1.0f;
}

template <int K>
__global__ __launch_bounds__(256) void stream_gather(const int* __restrict__ idx, const float* __restrict__ val,
                                                     const float* __restrict__ x, float* __restrict__ out) {
    const long base = (long)blockIdx.x * 256 * K + threadIdx.x;
    int c[K];
    float v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        c[k] = __builtin_nontemporal_load(idx + base + k * 256);
        v[k] = __builtin_nontemporal_load(val + base + k * 256);
    }
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < K; ++k) acc += v[k] * x[c[k]];
    out[blockIdx.x * 256 + threadIdx.x] = acc;
}

// Gathers only: indices from a hash, no index stream.
template <int K>
__global__ __launch_bounds__(256) void hash_gather(const float* __restrict__ x, int span, int slab, int iters,
                                                   float* __restrict__ out) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const int base = slab ? (int)(blockIdx.x & 7) * span : 0;
    float acc = 0.f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < K; ++k) acc += x[base + (int)(hsh(t * 1024u + (uint32_t)(it * K + k)) & (span - 1))];
    }
    out[t] = acc;
}

template <typename F>
float time_ms(F f, int reps = 7) {
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    f();
    CK(hipDeviceSynchronize());
    std::vector<float> ts;
    for (int r = 0; r < reps; ++r) {
        CK(hipEventRecord(a, 0));
        f();
        CK(hipEventRecord(b, 0));
        CK(hipEventSynchronize(b));
        float ms;
        CK(hipEventElapsedTime(&ms, a, b));
        ts.push_back(ms);
    }
    std::sort(ts.begin(), ts.end());
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
    return ts[reps / 2];
}

int one(int mode, int span, int K);
int main(int argc, char** argv) {
    if (argc == 4) return one(atoi(argv[1]), atoi(argv[2]), atoi(argv[3]));
    const long XN = 64L << 20;  // 64M floats = 256 MiB
    const long N = 64L << 20;   // 64M gathers; idx + val = 512 MiB (beyond the Infinity Cache)
    float *x, *val, *out;
    int* idx;
    CK(hipMalloc(&x, XN * 4));
    CK(hipMalloc(&idx, N * 4));
    CK(hipMalloc(&val, N * 4));
    CK(hipMalloc(&out, N / 1 * 4 / 1));  // grid*256 <= N
    CK(hipMemset(x, 0, XN * 4));
    const int spans[] = {1 << 18, 1 << 19, 1 << 20, 1 << 21, 1 << 22, 1 << 24};
    printf("# hash_gather (no index stream): grid 16384x256, K=8, iters 4 -> 134M gathers\n");
    for (int span : spans)
        for (int slab = 0; slab < 2; ++slab) {
            if (slab && (long)span * 8 > XN) continue;
            const int grid = 16384, iters = 4;
            float ms = time_ms([&] { hipLaunchKernelGGL(hash_gather<8>, dim3(grid), dim3(256), 0, 0, x, span, slab, iters, out); });
            double g = (double)grid * 256 * 8 * iters;
            printf("hash span=%8d (%6.1f MiB) slab=%d  %.3f ms  %.1f Ggather/s\n", span, span * 4.0 / (1 << 20), slab, ms,
                   g / ms / 1e6);
        }
    printf("# stream_gather: idx+val streamed (8 B) + x gather, 64M elements\n");
    for (int mode = 0; mode < 3; ++mode)
        for (int span : spans) {
            if (mode == 1 && (long)span * 8 > XN) continue;
            if (mode == 2 && span != (1 << 24)) continue;
            for (int K : {1, 4, 8, 16}) {
                const int per_block = 256 * K;
                const int grid = (int)(N / per_block);
                hipLaunchKernelGGL(mk_idx, dim3((unsigned)(N / 256)), dim3(256), 0, 0, idx, val, N, span, per_block, mode);
                CK(hipGetLastError());
                auto launch = [&] {
                    switch (K) {
                        case 1: hipLaunchKernelGGL(stream_gather<1>, dim3(grid), dim3(256), 0, 0, idx, val, x, out); break;
                        case 4: hipLaunchKernelGGL(stream_gather<4>, dim3(grid), dim3(256), 0, 0, idx, val, x, out); break;
                        case 8: hipLaunchKernelGGL(stream_gather<8>, dim3(grid), dim3(256), 0, 0, idx, val, x, out); break;
                        default: hipLaunchKernelGGL(stream_gather<16>, dim3(grid), dim3(256), 0, 0, idx, val, x, out);
                    }
                };
                float ms = time_ms(launch);
                printf("stream mode=%d span=%8d (%6.1f MiB) K=%2d  %.3f ms  %.1f Ggather/s  %.0f GB/s (8 B stream + 4 B x)\n",
                       mode, span, span * 4.0 / (1 << 20), K, ms, N / ms / 1e6, N * 12.0 / ms / 1e6);
            }
        }
    CK(hipFree(x));
    CK(hipFree(idx));
    CK(hipFree(val));
    CK(hipFree(out));
    return 0;
}

// a single stream_gather configuration, 5 timed launches (for counter runs)
int one(int mode, int span, int K) {
    const long XN = 64L << 20, N = 64L << 20;
    if (span <= 0 || (span & (span - 1)) || (mode == 1 && (long)span * 8 > XN) || span > XN || !(K == 4 || K == 8)) {
        printf("bad args\n");
        return 2;
    }
    float *x, *val, *out;
    int* idx;
    CK(hipMalloc(&x, XN * 4));
    CK(hipMalloc(&idx, N * 4));
    CK(hipMalloc(&val, N * 4));
    CK(hipMalloc(&out, N * 4));
    CK(hipMemset(x, 0, XN * 4));
    const int per_block = 256 * K;
    const int grid = (int)(N / per_block);
    hipLaunchKernelGGL(mk_idx, dim3((unsigned)(N / 256)), dim3(256), 0, 0, idx, val, N, span, per_block, mode);
    CK(hipGetLastError());
    auto launch = [&] {
        if (K == 4) hipLaunchKernelGGL(stream_gather<4>, dim3(grid), dim3(256), 0, 0, idx, val, x, out);
        else hipLaunchKernelGGL(stream_gather<8>, dim3(grid), dim3(256), 0, 0, idx, val, x, out);
    };
    float ms = time_ms(launch, 5);
    printf("stream mode=%d span=%d K=%d  %.3f ms  %.1f Ggather/s\n", mode, span, K, ms, N / ms / 1e6);
    CK(hipFree(x));
    CK(hipFree(idx));
    CK(hipFree(val));
    CK(hipFree(out));
    return 0;
}
