// Shared device/host helpers for the gfx950 SpMV kernels.
//
// Element types: values are stored as float, double or bf16 (uint16 bit
// pattern); x and y are float or double. Products and sums are formed in an
// accumulator type AT (float or double). With COMP=true the FP32 accumulator
// carries an error term: products are split exactly with an FMA (TwoProduct)
// and sums with Knuth's TwoSum, i.e. the compensated dot product "Dot2" of
// Ogita, Rump and Oishi, "Accurate sum and dot product", SIAM J. Sci.
// Comput. 26(6), 2005, Algorithm 5.3.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <exception>
#include <new>
#include <string>

#include "muxi_spmv/muxi_spmv.h"

namespace muxi {

constexpr int kWave = 64;

// bf16 storage: the upper 16 bits of an IEEE binary32.
struct bf16_t {
    uint16_t bits;
};

template <typename T> struct TypeOf;
template <> struct TypeOf<float> { static constexpr spmv_type_t v = SPMV_TYPE_FLOAT; };
template <> struct TypeOf<double> { static constexpr spmv_type_t v = SPMV_TYPE_DOUBLE; };
template <> struct TypeOf<bf16_t> { static constexpr spmv_type_t v = SPMV_TYPE_BF16; };

__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(double v) { return (float)v; }
__device__ __forceinline__ float to_f32(bf16_t v) { return __uint_as_float(((uint32_t)v.bits) << 16); }
__device__ __forceinline__ double to_f64(float v) { return (double)v; }
__device__ __forceinline__ double to_f64(double v) { return v; }
__device__ __forceinline__ double to_f64(bf16_t v) { return (double)to_f32(v); }

template <typename AT> struct Conv;
template <> struct Conv<float> {
    template <typename T> __device__ __forceinline__ static float f(T v) { return to_f32(v); }
};
template <> struct Conv<double> {
    template <typename T> __device__ __forceinline__ static double f(T v) { return to_f64(v); }
};

// Round-to-nearest-even binary32 -> bf16 (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
    uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
    u = __float_as_uint(f);
#else
    __builtin_memcpy(&u, &f, 4);
#endif
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------- accumulators
// A cross-lane DPP move of one value (GFX9 controls: row_shr:n = 0x110 | n inside a row of 16 lanes,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143, wave_shr:1 = 0x138; ROWS masks the rows written); a lane without a
// source reads 0. Runs in the ALUs, without the LDS round trip of a bpermute.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(dpp_move<CTRL, ROWS>(__float_as_int(v)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_move<CTRL, ROWS>((int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_move<CTRL, ROWS>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename AT, bool COMP> struct Acc;

template <typename AT> struct Acc<AT, false> {
    AT s;
    __device__ __forceinline__ void zero() { s = AT(0); }
    template <typename VT, typename XT>
    __device__ __forceinline__ void fma(VT v, XT xv) { s += Conv<AT>::f(v) * Conv<AT>::f(xv); }
    __device__ __forceinline__ void add(const Acc& o) { s += o.s; }
    __device__ __forceinline__ AT value() const { return s; }
    __device__ __forceinline__ Acc shfl_xor(int m) const { Acc r; r.s = __shfl_xor(s, m); return r; }
    __device__ __forceinline__ Acc shfl_up(int d) const { Acc r; r.s = __shfl_up(s, d); return r; }
    __device__ __forceinline__ Acc shfl(int src) const { Acc r; r.s = __shfl(s, src); return r; }
    template <int CTRL, int ROWS>
    __device__ __forceinline__ Acc dpp() const { Acc r; r.s = dpp_move<CTRL, ROWS>(s); return r; }
};

// Dot2: s + c carries the sum to about twice the working precision. The arithmetic is written out here with
// contraction off: a product fused into the following TwoSum addition would make TwoSum compute the error of a
// different sum.
template <> struct Acc<float, true> {
    float s, c;
    __device__ __forceinline__ void zero() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void two_sum_add(float a) {
#pragma clang fp contract(off)
        const float t = s + a;
        const float bp = t - s;
        const float e = (s - (t - bp)) + (a - bp);
        s = t;
        c = c + e;
    }
    template <typename VT, typename XT>
    __device__ __forceinline__ void fma(VT v, XT xv) {
#pragma clang fp contract(off)
        const float a = to_f32(v), b = to_f32(xv);
        const float p = a * b;
        const float ep = __builtin_fmaf(a, b, -p);  // exact product error
        two_sum_add(p);
        c = c + ep;
    }
    __device__ __forceinline__ void add(const Acc& o) {
#pragma clang fp contract(off)
        two_sum_add(o.s);
        c = c + o.c;
    }
    __device__ __forceinline__ float value() const {
#pragma clang fp contract(off)
        return s + c;
    }
    __device__ __forceinline__ Acc shfl_xor(int m) const { Acc r; r.s = __shfl_xor(s, m); r.c = __shfl_xor(c, m); return r; }
    __device__ __forceinline__ Acc shfl_up(int d) const { Acc r; r.s = __shfl_up(s, d); r.c = __shfl_up(c, d); return r; }
    __device__ __forceinline__ Acc shfl(int src) const { Acc r; r.s = __shfl(s, src); r.c = __shfl(c, src); return r; }
    template <int CTRL, int ROWS>
    __device__ __forceinline__ Acc dpp() const {
        Acc r;
        r.s = dpp_move<CTRL, ROWS>(s);
        r.c = dpp_move<CTRL, ROWS>(c);
        return r;
    }
};

// Product term stored between phases (LDS staging): one AT, or (p, err) when compensated.
template <typename AT, bool COMP> struct Prod;
template <typename AT> struct Prod<AT, false> {
    AT p;
    template <typename VT, typename XT>
    __device__ __forceinline__ static Prod make(VT v, XT xv) { Prod r; r.p = Conv<AT>::f(v) * Conv<AT>::f(xv); return r; }
    __device__ __forceinline__ static Prod zero() { Prod r; r.p = AT(0); return r; }
    __device__ __forceinline__ void into(Acc<AT, false>& a) const { a.s += p; }
};
template <> struct Prod<float, true> {
    float p, e;
    template <typename VT, typename XT>
    __device__ __forceinline__ static Prod make(VT v, XT xv) {
#pragma clang fp contract(off)
        const float a = to_f32(v), b = to_f32(xv);
        Prod r;
        r.p = a * b;
        r.e = __builtin_fmaf(a, b, -r.p);
        return r;
    }
    __device__ __forceinline__ static Prod zero() { Prod r; r.p = 0.f; r.e = 0.f; return r; }
    __device__ __forceinline__ void into(Acc<float, true>& a) const {
#pragma clang fp contract(off)
        a.two_sum_add(p);
        a.c = a.c + e;
    }
};

// Butterfly sum over groups of W lanes (W a power of two <= 64); every lane of
// the group ends with the same total, combined in a fixed order.
template <int W, typename A>
__device__ __forceinline__ A group_reduce(A a) {
#pragma unroll
    for (int m = W / 2; m >= 1; m >>= 1) {
        A o = a.shfl_xor(m);
        a.add(o);
    }
    return a;
}

// Non-temporal load for matrix arrays that one SpMV reads once: they pass
// through L2 without displacing the x lines that the gathers reuse.
template <typename T>
__device__ __forceinline__ T ntl(const T* p) {
    return __builtin_nontemporal_load(p);
}
__device__ __forceinline__ bf16_t ntl(const bf16_t* p) {
    bf16_t b;
    b.bits = __builtin_nontemporal_load(&p->bits);
    return b;
}

// y[i] = alpha*v + beta*y[i]; beta == 0 never reads y (so y may hold NaN).
template <typename YT, typename AT>
__device__ __forceinline__ void store_y(YT* y, long i, AT v, AT alpha, AT beta) {
    AT r = alpha * v;
    if (beta != AT(0)) r += beta * (AT)y[i];
    __builtin_nontemporal_store((YT)r, y + i);  // y is written once per SpMV: stream it
}

}  // namespace muxi

// ------------------------------------------------------------------ host side
namespace muxi {

void set_error(const std::string& msg);
const char* get_error();

#define MUXI_HIP(call)                                                           \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::muxi::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return SPMV_ERROR_DEVICE;                                            \
        }                                                                        \
    } while (0)

#define MUXI_LAUNCH_CHECK()                                                        \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess) {                                                    \
            ::muxi::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
            return SPMV_ERROR_DEVICE;                                              \
        }                                                                          \
    } while (0)

// Runs f() and turns a C++ exception (host containers on absurd sizes) into a
// status, so that none crosses the C ABI. An object half built when the
// exception came is leaked rather than freed on an unknown path.
template <typename F>
spmv_status_t no_throw(const char* where, F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        set_error(std::string(where) + ": out of host memory");
        return SPMV_ERROR_MEMORY_ALLOC;
    } catch (const std::exception& e) {
        set_error(std::string(where) + ": " + e.what());
        return SPMV_ERROR_INTERNAL;
    }
}

inline unsigned grid_for(long work, int per_block) {
    return (unsigned)((work + per_block - 1) / per_block);
}

}  // namespace muxi

namespace muxi {
// A row block of a larger matrix: rows [row_base, row_base + numRows) of a matrix of rows_full rows, numbered
// from 0 within the block, with the full matrix's columns (the host path's chunks). The features measure the
// band against the full matrix's diagonal, so a block of a banded matrix looks banded to the selector.
spmv_status_t matrix_features_rows(const muxi_csr_t* A, void* stream, long row_base, long rows_full, double* out);
spmv_status_t selector_pick_rows(const muxi_csr_t* A, spmv_type_t xtype, void* stream, long row_base,
                                 long rows_full, int* kernel, int* param0);
}  // namespace muxi
