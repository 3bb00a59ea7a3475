// HBM streaming probe: the bandwidth a kernel reaches reading (and copying)
// 2 GiB, far beyond the Infinity Cache, with 16-byte loads per lane. This is
// the practical ceiling of a bandwidth-bound SpMV such as ELL on a stencil,
// next to the nominal 8 TB/s and to hipMemcpy's device-to-device rate.
//
//   stream_probe     table: read-only kernels (non-temporal or plain loads,
//                    4 or 8 loads in flight per lane) and a non-temporal copy
//                    kernel over grids of 1024 .. 16384 workgroups, then
//                    hipMemcpy; median of 9 launches each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#define CK(x)                                                                            \
    do {                                                                                 \
        hipError_t e = (x);                                                              \
        if (e != hipSuccess) {                                                           \
            printf("HIP error %s at %s:%d\n", hipGetErrorString(e), __FILE__, __LINE__); \
            return 1;                                                                    \
        }                                                                                \
    } while (0)

typedef float f4 __attribute__((ext_vector_type(4)));

// Sums U 16-byte loads per lane per step (grid-stride); the result is kept live.
template <int U, bool NT>
__global__ __launch_bounds__(256) void read_kernel(const f4* __restrict__ a, long n4, float* out) {
    f4 acc = {0, 0, 0, 0};
    const long stride = (long)gridDim.x * 256 * U;
    for (long i = (long)blockIdx.x * 256 * U + threadIdx.x; i < n4; i += stride) {
        f4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long j = i + u * 256;
            v[u] = j < n4 ? (NT ? __builtin_nontemporal_load(a + j) : a[j]) : f4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u];
    }
    if (acc.x + acc.y + acc.z + acc.w == 12345.f) out[0] = 1.f;
}

template <int U>
__global__ __launch_bounds__(256) void copy_kernel(const f4* __restrict__ a, f4* __restrict__ b, long n4) {
    const long stride = (long)gridDim.x * 256 * U;
    for (long i = (long)blockIdx.x * 256 * U + threadIdx.x; i < n4; i += stride) {
        f4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long j = i + u * 256;
            v[u] = j < n4 ? __builtin_nontemporal_load(a + j) : f4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long j = i + u * 256;
            if (j < n4) __builtin_nontemporal_store(v[u], b + j);
        }
    }
}

template <typename F>
static float median_ms(F f) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    f();
    (void)hipDeviceSynchronize();
    std::vector<float> t;
    for (int r = 0; r < 9; ++r) {
        (void)hipEventRecord(a);
        f();
        (void)hipEventRecord(b);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return t[4];
}

int main() {
    const long bytes = 2L << 30, n4 = bytes / 16;
    f4 *a, *b;
    float* o;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMalloc(&o, 4));
    CK(hipMemset(a, 0, bytes));
    CK(hipMemset(b, 0, bytes));
    printf("# GB/s over 2 GiB (copy: read + write bytes)\n");
    for (int grid : {1024, 2048, 4096, 8192, 16384}) {
        const float r4nt = median_ms([&] { hipLaunchKernelGGL((read_kernel<4, true>), dim3(grid), dim3(256), 0, 0, a, n4, o); });
        const float r8nt = median_ms([&] { hipLaunchKernelGGL((read_kernel<8, true>), dim3(grid), dim3(256), 0, 0, a, n4, o); });
        const float r4 = median_ms([&] { hipLaunchKernelGGL((read_kernel<4, false>), dim3(grid), dim3(256), 0, 0, a, n4, o); });
        const float c4 = median_ms([&] { hipLaunchKernelGGL((copy_kernel<4>), dim3(grid), dim3(256), 0, 0, a, b, n4); });
        printf("grid %5d: read 4/lane nt %.0f, read 8/lane nt %.0f, read 4/lane %.0f, copy nt %.0f\n", grid,
               bytes / r4nt / 1e6, bytes / r8nt / 1e6, bytes / r4 / 1e6, 2.0 * bytes / c4 / 1e6);
    }
    const float m = median_ms([&] { (void)hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, 0); });
    printf("hipMemcpy device to device: %.0f\n", 2.0 * bytes / m / 1e6);
    CK(hipFree(a));
    CK(hipFree(b));
    CK(hipFree(o));
    return 0;
}
