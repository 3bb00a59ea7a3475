// Plan API: build once (partitions, tiles, ELL slots, carries), execute many
// times without allocating, so an exec may be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <new>
#include <type_traits>
#include <vector>

#include "muxi_spmv/muxi_spmv_tune.h"
#include "plan.hpp"
#include "sched.hpp"

using namespace muxi;

static void plan_free(muxi_plan* p) {
    if (!p) return;
    (void)hipFree(p->d_pos);
    (void)hipFree(p->d_tile_row);
    (void)hipFree(p->d_carry);
    (void)hipFree(p->d_ecol);
    (void)hipFree(p->d_eval);
    (void)hipFree(p->d_blocks);
    (void)hipFree(p->d_slab);
    (void)hipFree(p->d_slab_rpp);
    (void)hipFree(p->d_slab_ckey);
    (void)hipFree(p->d_slab_rmask);
    (void)hipFree(p->d_slab_wbase);
    (void)hipFree(p->d_slab_pc);
    (void)hipFree(p->d_slab_ct);
    (void)hipFree(p->d_slab_heavy);
    for (void* q : {(void*)p->d_ls_lc, p->d_ls_lv, (void*)p->d_ls_tp, (void*)p->d_ls_ch, (void*)p->d_ls_blk,
                    (void*)p->d_ls_F, (void*)p->d_ls_map, p->d_ls_P, (void*)p->d_ls_chk})
        (void)hipFree(q);
    delete p;
}

#define PLAN_TRY(expr)                     \
    do {                                   \
        spmv_status_t st_ = (expr);        \
        if (st_ != SPMV_SUCCESS) {         \
            plan_free(p);                  \
            return st_;                    \
        }                                  \
    } while (0)

#define PLAN_HIP(expr)                                                         \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
            plan_free(p);                                                      \
            return SPMV_ERROR_DEVICE;                                          \
        }                                                                      \
    } while (0)

static bool pow2_in(int v, int lo, int hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; }

static spmv_status_t plan_create_impl(const muxi_csr_t* Ain, spmv_type_t vt, spmv_type_t xt, muxi_kernel_t kernel,
                                          int param0, int compensated, void* stream, muxi_plan** out,
                                          long row_base = 0, long rows_full = -1, double check_rtol = 0.0) {
    if (!out) {
        set_error("plan output pointer is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    *out = nullptr;
    if (!Ain) {
        set_error("matrix is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (Ain->numRows < 0 || Ain->numCols < 0 || Ain->nnz < 0 ||
        (long)Ain->numRows + (long)Ain->nnz > 2147483647L) {
        set_error("invalid matrix dimensions");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (Ain->numRows > 0 && !Ain->d_rowPtr) {
        set_error("rowPtr device array is NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (Ain->nnz > 0 && (!Ain->d_colIdx || !Ain->d_values)) {
        set_error("matrix device arrays are NULL");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    // type check up front
    spmv_status_t tst = with_types(vt, xt, compensated != 0, [](auto) { return SPMV_SUCCESS; });
    if (tst != SPMV_SUCCESS) return tst;

    hipStream_t s = (hipStream_t)stream;
    muxi_plan* p = new (std::nothrow) muxi_plan();
    if (!p) return SPMV_ERROR_MEMORY_ALLOC;
    p->A = DevCsr{Ain->numRows, Ain->numCols, Ain->nnz, Ain->d_rowPtr, Ain->d_colIdx, Ain->d_values};
    p->vtype = vt;
    p->xtype = xt;
    p->comp = compensated != 0;
    p->check = check_rtol > 0.0;
    p->rtol = (float)check_rtol;
    if (rows_full < 0) rows_full = Ain->numRows;
    if (row_base < 0 || row_base + Ain->numRows > rows_full || rows_full > 2147483647L) {
        set_error("invalid row block");
        plan_free(p);
        return SPMV_ERROR_INVALID_MATRIX;
    }
    p->row_base = (int)row_base;
    PLAN_HIP(hipGetDevice(&p->device));
    const int R = Ain->numRows, N = Ain->nnz;

    if (kernel == MUXI_KERNEL_AUTO) {
        int k = MUXI_KERNEL_MERGE, q = 4;
        PLAN_TRY(selector_pick_rows(Ain, xt, stream, row_base, rows_full, &k, &q));
        // Compensated sums: the L2 column-slab kernel has no Dot2 variant; its place goes to the LDS column slabs
        // (Dot2 with 8-byte pair sums in narrower slabs, at most kLsMaxSlabs of them, with FP32 x), else to
        // merge-path.
        if (p->comp && (k == MUXI_KERNEL_XSLAB || k == MUXI_KERNEL_LSLAB)) {
            const bool ls_ok = xt == SPMV_TYPE_FLOAT &&
                               (long)Ain->numCols <= (long)kLsMaxSlabs * lslab_slab_cols(4, 8);
            k = ls_ok ? MUXI_KERNEL_LSLAB : MUXI_KERNEL_MERGE;
            q = 4;
        }
        kernel = (muxi_kernel_t)k;
        param0 = q;
    }
    p->kernel = kernel;
    p->param0 = param0;
    const size_t acc_bytes = xt == SPMV_TYPE_DOUBLE ? 8 : (p->comp ? 8 : 4);

    switch (kernel) {
        case MUXI_KERNEL_SCALAR:
        case MUXI_KERNEL_VECTOR:
            break;
        case MUXI_KERNEL_VWAVE:
            if (param0 <= 0) param0 = 8;
            if (!pow2_in(param0, 1, 64)) {
                set_error("vwave: lanes per row must be a power of two in [1, 64]");
                plan_free(p);
                return SPMV_ERROR_INTERNAL;
            }
            p->param0 = param0;
            break;
        case MUXI_KERNEL_MERGE: {
            if (param0 <= 0) param0 = 8;
            if (!(param0 == 2 || param0 == 4 || param0 == 8 || param0 == 16)) {
                set_error("merge: items per thread must be 2, 4, 8 or 16");
                plan_free(p);
                return SPMV_ERROR_INTERNAL;
            }
            p->param0 = param0;
            const int P = merge_default_partitions(R + N, param0);
            p->P = P;
            PLAN_HIP(hipMalloc(&p->d_pos, 2 * sizeof(int) * (size_t)(P + 1)));
            p->d_row = p->d_pos + (P + 1);
            PLAN_HIP(hipMalloc(&p->d_carry, acc_bytes * (size_t)P));
            p->scratch_bytes = 2 * sizeof(int) * (size_t)(P + 1) + acc_bytes * (size_t)P;
            PLAN_TRY(launch_merge_partitions(R, N, P, p->d_pos, s));
            PLAN_TRY(launch_merge_coords(p->A.rp, R, N, p->d_pos, P + 1, p->d_row, nullptr, s));
            break;
        }
        case MUXI_KERNEL_CSR5: {
            if (param0 <= 0) param0 = 256;
            if (!csr5_sigma_supported(param0)) {
                set_error("csr5: sigma must be 64, 128, 256, 512 or 1024");
                plan_free(p);
                return SPMV_ERROR_INTERNAL;
            }
            p->param0 = param0;
            const int T = (int)(((long)N + param0 - 1) / param0);
            p->T = T;
            if (T > 0) {
                PLAN_HIP(hipMalloc(&p->d_tile_row, 2 * sizeof(int) * (size_t)T));
                p->d_tile_off = p->d_tile_row + T;
                PLAN_HIP(hipMalloc(&p->d_carry, acc_bytes * (size_t)T));
                p->scratch_bytes = 2 * sizeof(int) * (size_t)T + acc_bytes * (size_t)T;
                PLAN_TRY(launch_csr5_tiles(p->A.rp, R, N, param0, T, p->d_tile_row, p->d_tile_off, s));
            }
            break;
        }
        case MUXI_KERNEL_ADAPTIVE: {
            if (param0 <= 0) param0 = 1024;
            if (!(param0 == 256 || param0 == 512 || param0 == 1024 || param0 == 2048)) {
                set_error("adaptive: block nonzeros must be 256, 512, 1024 or 2048");
                plan_free(p);
                return SPMV_ERROR_INTERNAL;
            }
            p->param0 = param0;
            if (R > 0) {
                std::vector<int> br, bc, lr, lf;
                PLAN_TRY(adaptive_build(p->A.rp, R, param0, s, br, bc, lr, lf));
                p->nblocks = (int)bc.size();
                p->nlong = (int)lr.size();
                const size_t ints = br.size() + bc.size() + lr.size() + lf.size();
                PLAN_HIP(hipMalloc(&p->d_blocks, sizeof(int) * ints));
                p->d_blk_chunk = p->d_blocks + br.size();
                p->d_long_row = p->d_blk_chunk + bc.size();
                p->d_long_first = p->d_long_row + lr.size();
                PLAN_HIP(hipMemcpy(p->d_blocks, br.data(), sizeof(int) * br.size(), hipMemcpyHostToDevice));
                PLAN_HIP(hipMemcpy(p->d_blk_chunk, bc.data(), sizeof(int) * bc.size(), hipMemcpyHostToDevice));
                if (!lr.empty()) {
                    PLAN_HIP(hipMemcpy(p->d_long_row, lr.data(), sizeof(int) * lr.size(), hipMemcpyHostToDevice));
                    PLAN_HIP(hipMemcpy(p->d_long_first, lf.data(), sizeof(int) * lf.size(), hipMemcpyHostToDevice));
                }
                PLAN_HIP(hipMalloc(&p->d_carry, acc_bytes * (size_t)(p->nblocks > 0 ? p->nblocks : 1)));
                p->scratch_bytes = sizeof(int) * ints + acc_bytes * (size_t)p->nblocks;
            }
            break;
        }
        case MUXI_KERNEL_XSLAB: {
            int S = param0 >> 8, ipt = param0 & 0xff;  // ipt: nonzeros per lane
            if (S <= 0) S = xslab_default_slabs(Ain->numCols, xt == SPMV_TYPE_DOUBLE ? 8 : 4);
            if (ipt <= 0) ipt = 4;
            if (!(ipt == 4 || ipt == 8 || ipt == 16) || S > kMaxSlabs) {
                set_error("xslab: param0 = (slabs << 8) | nonzeros per lane, slabs in [1, 16], per lane 4, 8 or 16");
                plan_free(p);
                return SPMV_ERROR_INTERNAL;
            }
            if (p->comp) {
                set_error("xslab: compensated accumulation is not available");
                plan_free(p);
                return SPMV_ERROR_UNSUPPORTED_TYPE;
            }
            p->param0 = (S << 8) | ipt;
            if (R > 0) PLAN_TRY(xslab_build(p, S, ipt, vt, s));
            break;
        }
        case MUXI_KERNEL_LSLAB: {
            if (param0 <= 0) param0 = 4;
            if (!((param0 & 0xff) == 4 || (param0 & 0xff) == 8) || (param0 & ~(0xff | kLsHalfSlabs))) {
                set_error("lslab: param0 = nonzeros per lane (4 or 8) | 0x100 (slabs of 16384 columns)");
                plan_free(p);
                return SPMV_ERROR_INTERNAL;
            }
            if (xt != SPMV_TYPE_FLOAT) {
                set_error("lslab: x and y must be FLOAT (the slab of x in LDS is FP32)");
                plan_free(p);
                return SPMV_ERROR_UNSUPPORTED_TYPE;
            }
            p->param0 = param0;
            if (R > 0) PLAN_TRY(lslab_build(p, param0, vt, s));
            break;
        }
        case MUXI_KERNEL_ELL: {
            int width = param0;
            if (width <= 0) {
                int* d_w = nullptr;
                PLAN_HIP(hipMalloc(&d_w, sizeof(int)));
                spmv_status_t st = launch_max_row_len(p->A.rp, R, d_w, s);
                if (st == SPMV_SUCCESS) {
                    if (hipMemcpyAsync(&width, d_w, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess)
                        st = SPMV_ERROR_MEMORY_COPY;
                }
                (void)hipFree(d_w);
                PLAN_TRY(st);
            }
            if ((long)R * (long)width > 2147483647L) {
                set_error("ELLPACK numRows*width exceeds INT32_MAX");
                plan_free(p);
                return SPMV_ERROR_INVALID_MATRIX;
            }
            p->width = width;
            p->param0 = width;
            const size_t slots = (size_t)R * (size_t)width;
            if (slots > 0) {
                // 16-bit column offsets when every column is within 32767 of its row
                int fit[2] = {1, 32768};  // needs more than 16 bits, largest |column - row|
                {
                    int* d_bad = nullptr;
                    PLAN_HIP(hipMalloc(&d_bad, 2 * sizeof(int)));
                    spmv_status_t st = launch_ell_offsets_fit(p->A, width, p->row_base, d_bad, s);
                    if (st == SPMV_SUCCESS &&
                        (hipMemcpyAsync(fit, d_bad, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                         hipStreamSynchronize(s) != hipSuccess))
                        st = SPMV_ERROR_MEMORY_COPY;
                    (void)hipFree(d_bad);
                    PLAN_TRY(st);
                }
                p->ecol16 = fit[0] == 0;
                // narrow bands (every column within 512 of its row) stage each block's x window in LDS
                p->ell_win = p->ecol16 && fit[1] <= 512 ? fit[1] : -1;
                const size_t csize = p->ecol16 ? sizeof(short) : sizeof(int);
                PLAN_HIP(hipMalloc(&p->d_ecol, csize * slots));
                PLAN_HIP(hipMalloc(&p->d_eval, vsize(vt) * slots));
                p->scratch_bytes = (csize + vsize(vt)) * slots;
                PLAN_TRY(with_types(vt, xt, p->comp, [&](auto tag) {
                    using T = decltype(tag);
                    return launch_csr_to_ellg<typename T::VT>(p->A, width, p->ecol16, p->row_base, p->d_ecol,
                                                              (typename T::VT*)p->d_eval, s);
                }));
            }
            break;
        }
        default:
            set_error("unknown kernel id");
            plan_free(p);
            return SPMV_ERROR_INTERNAL;
    }
    *out = p;
    return SPMV_SUCCESS;
}

static spmv_status_t plan_create_spmm_impl(const muxi_csr_t* Ain, spmv_type_t vt, spmv_type_t xt, int nrhs,
                                               int ipt, void* stream, muxi_plan** out) {
    if (!out) {
        set_error("plan output pointer is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    *out = nullptr;
    if (!Ain || Ain->numRows < 0 || Ain->numCols < 0 || Ain->nnz < 0 ||
        (long)Ain->numRows + (long)Ain->nnz > 2147483647L || (Ain->numRows > 0 && !Ain->d_rowPtr) ||
        (Ain->nnz > 0 && (!Ain->d_colIdx || !Ain->d_values))) {
        set_error("invalid matrix");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    if (nrhs < 1 || nrhs > 64) {
        set_error("spmm: nrhs must be in [1, 64]");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    if (ipt <= 0)  // the tuned schedule table: by k and the merge-path length
        ipt = spmm_schedule(nrhs, (double)Ain->numRows + (double)Ain->nnz);
    if (!(ipt == 4 || ipt == 8 || ipt == 16 || ipt == 32)) {
        set_error("spmm: items per group must be 4, 8, 16 or 32");
        return SPMV_ERROR_INTERNAL;
    }
    const bool ok_types = (xt == SPMV_TYPE_FLOAT && (vt == SPMV_TYPE_FLOAT || vt == SPMV_TYPE_BF16 ||
                                                     vt == SPMV_TYPE_DOUBLE)) ||
                          (xt == SPMV_TYPE_DOUBLE && vt == SPMV_TYPE_DOUBLE);
    if (!ok_types) {
        set_error("spmm: unsupported type combination");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    hipStream_t s = (hipStream_t)stream;
    muxi_plan* p = new (std::nothrow) muxi_plan();
    if (!p) return SPMV_ERROR_MEMORY_ALLOC;
    p->A = DevCsr{Ain->numRows, Ain->numCols, Ain->nnz, Ain->d_rowPtr, Ain->d_colIdx, Ain->d_values};
    p->vtype = vt;
    p->xtype = xt;
    p->comp = false;
    p->kernel = MUXI_KERNEL_MERGE;
    p->param0 = ipt;
    p->nrhs = nrhs;
    PLAN_HIP(hipGetDevice(&p->device));
    const int R = Ain->numRows, N = Ain->nnz;
    const long ch = spmm_chunk_items(nrhs, ipt);
    const int P = (int)std::max(1L, ((long)R + N + ch - 1) / ch);
    p->P = P;
    const size_t es = xt == SPMV_TYPE_DOUBLE ? 8 : 4;
    PLAN_HIP(hipMalloc(&p->d_pos, 2 * sizeof(int) * (size_t)(P + 1)));
    p->d_row = p->d_pos + (P + 1);
    PLAN_HIP(hipMalloc(&p->d_carry, es * (size_t)P * nrhs));
    p->scratch_bytes = 2 * sizeof(int) * (size_t)(P + 1) + es * (size_t)P * nrhs;
    PLAN_TRY(launch_merge_partitions(R, N, P, p->d_pos, s));
    PLAN_TRY(launch_merge_coords(p->A.rp, R, N, p->d_pos, P + 1, p->d_row, nullptr, s));
    *out = p;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_plan_destroy(muxi_plan* p) {
    if (!p) {
        set_error("plan is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    plan_free(p);
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_plan_info(const muxi_plan* p, int* kernel, int* param0, size_t* scratch) {
    if (!p) {
        set_error("plan is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    if (kernel) *kernel = (int)p->kernel;
    if (param0) *param0 = p->param0;
    if (scratch) *scratch = p->scratch_bytes;
    return SPMV_SUCCESS;
}

static spmv_status_t plan_exec(const muxi_plan* p, const void* x, void* y, double alpha, double beta, int nrhs,
                               hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (!p) {
        set_error("plan is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    if (!x || !y) {
        set_error("x or y is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    if (p->nrhs > 0) {
        if (nrhs != p->nrhs) {
            set_error("nrhs differs from the SpMM plan's");
            return SPMV_ERROR_INVALID_VECTOR;
        }
        if (ev0) MUXI_HIP(hipEventRecord(ev0, s));
        if (p->xtype == SPMV_TYPE_DOUBLE)
            return launch_spmm<double, double>(p->A, nrhs, p->param0, p->d_pos, p->d_row, p->P, p->d_carry,
                                               (const double*)x, (double*)y, alpha, beta, s, ev1);
        if (p->vtype == SPMV_TYPE_BF16)
            return launch_spmm<bf16_t, float>(p->A, nrhs, p->param0, p->d_pos, p->d_row, p->P, p->d_carry,
                                              (const float*)x, (float*)y, (float)alpha, (float)beta, s, ev1);
        if (p->vtype == SPMV_TYPE_DOUBLE)
            return launch_spmm<double, float>(p->A, nrhs, p->param0, p->d_pos, p->d_row, p->P, p->d_carry,
                                              (const float*)x, (float*)y, (float)alpha, (float)beta, s, ev1);
        return launch_spmm<float, float>(p->A, nrhs, p->param0, p->d_pos, p->d_row, p->P, p->d_carry,
                                         (const float*)x, (float*)y, (float)alpha, (float)beta, s, ev1);
    }
    if (nrhs != 1) {
        set_error("nrhs > 1 needs a plan from muxi_plan_create_spmm");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    if (ev0) MUXI_HIP(hipEventRecord(ev0, s));
    spmv_status_t st = with_types(p->vtype, p->xtype, p->comp, [&](auto tag) -> spmv_status_t {
        using T = decltype(tag);
        using VT = typename T::VT;
        using XT = typename T::XT;
        using AT = typename T::AT;
        constexpr bool C = T::COMP;
        const XT* xx = (const XT*)x;
        XT* yy = (XT*)y;
        AT a = (AT)alpha, b = (AT)beta;
        switch (p->kernel) {
            case MUXI_KERNEL_SCALAR: return launch_scalar<VT, XT, AT, C>(p->A, xx, yy, a, b, s);
            case MUXI_KERNEL_VECTOR: return launch_vector<VT, XT, AT, C>(p->A, xx, yy, a, b, s);
            case MUXI_KERNEL_VWAVE: return launch_vwave<VT, XT, AT, C>(p->A, p->param0, xx, yy, a, b, s);
            case MUXI_KERNEL_MERGE:
                return launch_merge<VT, XT, AT, C>(p->A, p->param0, p->d_pos, p->d_row, p->P, p->d_carry, xx, yy, a,
                                                   b, s, ev1);
            case MUXI_KERNEL_CSR5:
                if (p->T == 0) {
                    // no nonzeros: every row is empty
                    return launch_scalar<VT, XT, AT, C>(p->A, xx, yy, a, b, s);
                }
                return launch_csr5<VT, XT, AT, C>(p->A, p->param0, p->d_tile_row, p->T, p->d_carry, xx, yy, a, b, s,
                                                  ev1);
            case MUXI_KERNEL_ELL:
                return launch_ellg<VT, XT, AT, C>(p->A.R, p->A.C, p->width, p->ecol16, p->row_base, p->ell_win,
                                                  p->d_ecol, (const VT*)p->d_eval, xx, yy, a, b, s);
            case MUXI_KERNEL_XSLAB:
                if constexpr (C) {
                    set_error("xslab: no compensated variant");
                    return SPMV_ERROR_UNSUPPORTED_TYPE;
                } else {
                    return launch_xslab<VT, XT, AT>(p, xx, yy, a, b, s, ev1);
                }
            case MUXI_KERNEL_LSLAB:
                if constexpr (std::is_same<XT, float>::value && std::is_same<AT, float>::value) {
                    return launch_lslab<VT, C>(p, xx, yy, a, b, s, ev1);
                } else {
                    set_error("lslab: x and y must be FLOAT");
                    return SPMV_ERROR_UNSUPPORTED_TYPE;
                }
            case MUXI_KERNEL_ADAPTIVE:
                return launch_adaptive<VT, XT, AT, C>(p->A, p->param0, p->nblocks, p->d_blocks, p->d_blk_chunk,
                                                      p->nlong, p->d_long_row, p->d_long_first, p->d_carry, xx, yy, a,
                                                      b, s, ev1);
            default: set_error("plan kernel not executable"); return SPMV_ERROR_INTERNAL;
        }
    });
    if (st != SPMV_SUCCESS) return st;
    // single-kernel paths: the main kernel ends with the exec
    if (ev1 && !((p->kernel == MUXI_KERNEL_MERGE && p->A.R > 0) || (p->kernel == MUXI_KERNEL_CSR5 && p->T > 0) ||
                 (p->kernel == MUXI_KERNEL_ADAPTIVE && p->A.R > 0 && p->nblocks > 0) ||
                 (p->kernel == MUXI_KERNEL_XSLAB && p->A.R > 0) || p->kernel == MUXI_KERNEL_LSLAB))
        MUXI_HIP(hipEventRecord(ev1, s));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_plan_exec(const muxi_plan* p, const void* x, void* y, double alpha, double beta,
                                        int nrhs, void* stream) {
    return plan_exec(p, x, y, alpha, beta, nrhs, (hipStream_t)stream, nullptr, nullptr);
}

extern "C" spmv_status_t muxi_plan_exec_ev(const muxi_plan* p, const void* x, void* y, double alpha, double beta,
                                           int nrhs, void* stream, void* ev_before, void* ev_after) {
    return plan_exec(p, x, y, alpha, beta, nrhs, (hipStream_t)stream, (hipEvent_t)ev_before, (hipEvent_t)ev_after);
}

extern "C" spmv_status_t muxi_event_create(void** ev) {
    if (!ev) return SPMV_ERROR_INVALID_HANDLE;
    MUXI_HIP(hipEventCreate((hipEvent_t*)ev));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_event_destroy(void* ev) {
    if (!ev) return SPMV_ERROR_INVALID_HANDLE;
    MUXI_HIP(hipEventDestroy((hipEvent_t)ev));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_event_record(void* ev, void* stream) {
    if (!ev) return SPMV_ERROR_INVALID_HANDLE;
    MUXI_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_event_elapsed_ms(void* start, void* end, float* ms) {
    if (!start || !end || !ms) return SPMV_ERROR_INVALID_HANDLE;
    MUXI_HIP(hipEventSynchronize((hipEvent_t)end));
    MUXI_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_plan_create(const muxi_csr_t* Ain, spmv_type_t vt, spmv_type_t xt, muxi_kernel_t kernel, int param0,
                                          int compensated, void* stream, muxi_plan** out) {
    return no_throw("muxi_plan_create", [&] { return plan_create_impl(Ain, vt, xt, kernel, param0, compensated, stream, out); });
}

spmv_status_t muxi::plan_create_rows(const muxi_csr_t* A, spmv_type_t vt, spmv_type_t xt, muxi_kernel_t kernel,
                                     int param0, int compensated, void* stream, long row_base, long rows_full,
                                     muxi_plan** out) {
    return no_throw("plan_create_rows", [&] {
        return plan_create_impl(A, vt, xt, kernel, param0, compensated, stream, out, row_base, rows_full);
    });
}

extern "C" spmv_status_t muxi_plan_create_rows(const muxi_csr_t* A, spmv_type_t vt, spmv_type_t xt,
                                               muxi_kernel_t kernel, int param0, int compensated, void* stream,
                                               long long row_base, long long rows_full, muxi_plan** out) {
    return plan_create_rows(A, vt, xt, kernel, param0, compensated, stream, (long)row_base, (long)rows_full, out);
}

extern "C" spmv_status_t muxi_batch_schedule_for(double mean_row_len, double mean_rows, int* ipt, int* wg_per_cu) {
    if (!ipt || !wg_per_cu) return SPMV_ERROR_INVALID_HANDLE;
    batch_schedule(mean_row_len, mean_rows, ipt, wg_per_cu);
    return SPMV_SUCCESS;
}

extern "C" int muxi_spmm_schedule_for(int k, double path_len) { return spmm_schedule(k, path_len); }

extern "C" spmv_status_t muxi_plan_create_checked(const muxi_csr_t* A, spmv_type_t vt, int compensated, double rtol,
                                                  void* stream, muxi_plan** out) {
    return no_throw("muxi_plan_create_checked", [&] {
        if (!(rtol > 0.0 && rtol < 1.0)) {
            if (out) *out = nullptr;
            set_error("checked plan: rtol must lie in (0, 1)");
            return SPMV_ERROR_INTERNAL;
        }
        return plan_create_impl(A, vt, SPMV_TYPE_FLOAT, MUXI_KERNEL_LSLAB, 4, compensated, stream, out, 0, -1, rtol);
    });
}

extern "C" spmv_status_t muxi_plan_checked_rows(const muxi_plan* p, int* count, int* rows, int cap, void* stream) {
    if (!p) {
        set_error("plan is NULL");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    if (!p->check) {
        set_error("plan is not a checked plan (muxi_plan_create_checked)");
        return SPMV_ERROR_INVALID_HANDLE;
    }
    if (!count || (cap > 0 && !rows)) {
        set_error("count or rows is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    *count = 0;
    if (p->A.R <= 0 || !p->d_ls_chk) return SPMV_SUCCESS;
    hipStream_t s = (hipStream_t)stream;
    MUXI_HIP(hipMemcpyAsync(count, p->d_ls_chk, sizeof(int), hipMemcpyDeviceToHost, s));
    MUXI_HIP(hipStreamSynchronize(s));
    const int n = std::min(std::min(*count, cap), p->A.R);
    if (n > 0) {
        MUXI_HIP(hipMemcpyAsync(rows, p->d_ls_chk + 1, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
        MUXI_HIP(hipStreamSynchronize(s));
    }
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_plan_create_spmm(const muxi_csr_t* Ain, spmv_type_t vt, spmv_type_t xt, int nrhs, int ipt,
                                               void* stream, muxi_plan** out) {
    return no_throw("muxi_plan_create_spmm", [&] { return plan_create_spmm_impl(Ain, vt, xt, nrhs, ipt, stream, out); });
}
