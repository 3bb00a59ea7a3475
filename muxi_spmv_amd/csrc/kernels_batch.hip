// Batched SpMV over many independent CSR matrices in one persistent launch.
//
// Every matrix is cut into merge-path tiles of at most 256*IPT path items
// (the same coordinates as the single-matrix merge kernel). A tile table
// maps each global tile to its matrix; a persistent grid sized to the device
// (a multiple of 8) strides through it, the workgroups of each XCD through
// their own contiguous eighth, running the shared tile body. Carries are
// keyed by a virtual row (each matrix's rows plus one dummy row, laid end to
// end) so that runs never cross matrices, and one fix-up kernel folds them.
//
// Slab mode (FP32, large batches): the batch is one SpMV with the block-diagonal matrix of its members, x and y
// laid end to end. The plan builds that matrix once and an LDS column-slab plan over it (x from LDS, no gather
// goes to L2 at random); an exec gathers the members' x into one vector, runs the slab plan into one y and
// scatters y back to the members with alpha and beta.
#include <algorithm>
#include <climits>
#include <vector>

#include "merge_tile.hpp"
#include "muxi_spmv/muxi_spmv_batch.h"
#include "plan.hpp"
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

// ---- slab mode: members' vectors to and from one vector laid end to end
// A workgroup takes kBatchCopyPer consecutive elements of the concatenation, kBatchCopyThreads apart per lane.
// seg[b] describes the member holding workgroup b's first element -- its index, its first element in the
// concatenation, the first element of the next member with any, its vector -- so that a lane's element costs one
// dependent load; a lane steps forward (off[m] = a member's first element, off[count] = n) only past a member that
// ends inside the workgroup's range.
constexpr int kBatchCopyThreads = 256, kBatchCopyEpt = 4, kBatchCopyPer = kBatchCopyThreads * kBatchCopyEpt;

struct CopySeg {
    long start, next;
    void* v;
    int m, pad;
};

__global__ __launch_bounds__(kBatchCopyThreads) void batch_gather_x_kernel(const CopySeg* __restrict__ seg,
                                                                           const long* __restrict__ coff, int count,
                                                                           const float* const* __restrict__ xs,
                                                                           float* __restrict__ xcat, long n) {
    const CopySeg d = seg[blockIdx.x];  // uniform: scalar loads
    int m = d.m;
    long start = d.start, next = d.next;
    const float* x = (const float*)d.v;
    // every element's member first (a walk only past a member that ends inside the range), then every load at
    // once, then the stores
    const float* src[kBatchCopyEpt];
    const long i0 = (long)blockIdx.x * kBatchCopyPer + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kBatchCopyEpt; ++k) {
        const long i = i0 + k * kBatchCopyThreads;
        src[k] = nullptr;
        if (i >= n) continue;
        if (i >= next) {
            while (i >= next && m + 1 < count) {
                start = coff[++m];
                next = coff[m + 1];
            }
            x = xs[m];
        }
        src[k] = x ? x + (i - start) : nullptr;  // a member without rows may come without x: never read
    }
    float v[kBatchCopyEpt];
#pragma unroll
    for (int k = 0; k < kBatchCopyEpt; ++k) v[k] = src[k] ? *src[k] : 0.f;
#pragma unroll
    for (int k = 0; k < kBatchCopyEpt; ++k)
        if (i0 + k * kBatchCopyThreads < n) xcat[i0 + k * kBatchCopyThreads] = v[k];
}

__global__ __launch_bounds__(kBatchCopyThreads) void batch_scatter_y_kernel(const CopySeg* __restrict__ seg,
                                                                            const long* __restrict__ roff, int count,
                                                                            float* const* __restrict__ ys,
                                                                            const float* __restrict__ ycat, long n,
                                                                            float alpha, float beta) {
    const CopySeg d = seg[blockIdx.x];
    int m = d.m;
    long start = d.start, next = d.next;
    float* y = (float*)d.v;
    float* dst[kBatchCopyEpt];
    const long i0 = (long)blockIdx.x * kBatchCopyPer + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kBatchCopyEpt; ++k) {
        const long i = i0 + k * kBatchCopyThreads;
        dst[k] = nullptr;
        if (i >= n) continue;
        if (i >= next) {
            while (i >= next && m + 1 < count) {
                start = roff[++m];
                next = roff[m + 1];
            }
            y = ys[m];
        }
        dst[k] = y + (i - start);
    }
    float v[kBatchCopyEpt];
#pragma unroll
    for (int k = 0; k < kBatchCopyEpt; ++k) v[k] = dst[k] ? ycat[i0 + k * kBatchCopyThreads] : 0.f;
#pragma unroll
    for (int k = 0; k < kBatchCopyEpt; ++k)
        if (dst[k]) store_y(dst[k], 0, v[k], alpha, beta);
}

// The descriptor of each copy workgroup's first member: the last m with off[m] <= b * kBatchCopyPer.
static std::vector<CopySeg> batch_copy_segs(const std::vector<long>& off, const std::vector<void*>& vec, long n) {
    const int count = (int)off.size() - 1;
    std::vector<CopySeg> f((size_t)std::max(1L, (n + kBatchCopyPer - 1) / kBatchCopyPer));
    int m = 0;
    for (size_t b = 0; b < f.size(); ++b) {
        const long i = (long)b * kBatchCopyPer;
        while (m + 1 < count && off[m + 1] <= i) ++m;
        f[b] = CopySeg{off[m], off[m + 1], vec[m], m, 0};
    }
    return f;
}

// A member's row pointers must run from 0 to its nonzero count for its block of the block-diagonal matrix to line
// up with the column indices and values copied for it.
__global__ void batch_rowptr_ends_kernel(const int* __restrict__ rp, int R, int nnz, int* __restrict__ bad) {
    if (rp[0] != 0 || rp[R] != nnz) *bad = 1;
}

// dst[i] = src[i] + add for i < n (a member's row pointers or column indices moved into the block-diagonal matrix)
__global__ __launch_bounds__(256) void batch_offset_copy_kernel(const int* __restrict__ src, int* __restrict__ dst,
                                                                long n, int add) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + add;
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
    // slab mode: the LDS column-slab plan over the block-diagonal matrix (whose row pointers it keeps), the
    // concatenated x and y, the members' first element in each (count + 1 entries) and their vector pointers
    bool slabs = false;
    muxi_plan* lp = nullptr;
    int* d_bd_rp = nullptr;
    float* d_xcat = nullptr;
    float* d_ycat = nullptr;
    long* d_coff = nullptr;
    long* d_roff = nullptr;
    CopySeg* d_xseg = nullptr;  // the first member of each gather / scatter workgroup
    CopySeg* d_yseg = nullptr;
    const float** d_xptr = nullptr;
    float** d_yptr = nullptr;
    long Rtot = 0, Ctot = 0;
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
    if (p->lp) (void)muxi_plan_destroy(p->lp);
    for (void* q : {(void*)p->d_bd_rp, (void*)p->d_xcat, (void*)p->d_ycat, (void*)p->d_coff, (void*)p->d_roff,
                    (void*)p->d_xseg, (void*)p->d_yseg, (void*)p->d_xptr, (void*)p->d_yptr})
        (void)hipFree(q);
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

// Slab mode when ipt == 0: FP32 values and vectors, at least 64 Ki columns and 8 Mi nonzeros in all, columns
// within kLsMaxSlabs slabs and rows + nonzeros within an int. Measured against the tuned merge schedule (cold,
// MI355X): 1.3 to 2.0x faster from 9 M nonzeros up, level near 5 M, slower below and on 7.5 to 8 M nonzeros of
// very short rows or of hundreds of tiny matrices (the merge tiles take one launch, the slab mode four).
static constexpr long kBatchSlabMinCols = 1L << 16, kBatchSlabMinNnz = 1L << 23;

namespace {
struct DevBuf {  // device scratch of the plan build, freed on every return path
    void* p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
};
}  // namespace

// The block-diagonal matrix of the members (rows, columns and nonzeros laid end to end), its LDS column-slab plan,
// and the concatenation buffers. The plan keeps the row pointers; the column indices and values are read only while
// it is built.
static spmv_status_t batch_slabs_build(muxi_batch_plan* p, const muxi_csr_t* mats, int count,
                                       const void* const* xs, void* const* ys, hipStream_t s) {
    std::vector<long> roff((size_t)count + 1, 0), coff((size_t)count + 1, 0), noff((size_t)count + 1, 0);
    for (int m = 0; m < count; ++m) {
        roff[m + 1] = roff[m] + mats[m].numRows;
        coff[m + 1] = coff[m] + mats[m].numCols;
        noff[m + 1] = noff[m] + mats[m].nnz;
    }
    const long R = roff[count], C = coff[count], N = noff[count];
    p->Rtot = R;
    p->Ctot = C;
    DevBuf ci, va, bad;
    MUXI_HIP(hipMalloc(&bad.p, sizeof(int)));
    MUXI_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    MUXI_HIP(hipMalloc(&p->d_bd_rp, sizeof(int) * (size_t)(R + 1)));
    MUXI_HIP(hipMalloc(&ci.p, sizeof(int) * (size_t)std::max(N, 1L)));
    MUXI_HIP(hipMalloc(&va.p, sizeof(float) * (size_t)std::max(N, 1L)));
    for (int m = 0; m < count; ++m) {
        const muxi_csr_t& A = mats[m];
        if (A.numRows > 0) {
            hipLaunchKernelGGL(batch_rowptr_ends_kernel, dim3(1), dim3(1), 0, s, A.d_rowPtr, A.numRows, A.nnz,
                               (int*)bad.p);
            hipLaunchKernelGGL(batch_offset_copy_kernel, dim3(grid_for(A.numRows, 256)), dim3(256), 0, s, A.d_rowPtr,
                               p->d_bd_rp + roff[m], (long)A.numRows, (int)noff[m]);
        }
        if (A.nnz > 0) {
            hipLaunchKernelGGL(batch_offset_copy_kernel, dim3(grid_for(A.nnz, 256)), dim3(256), 0, s, A.d_colIdx,
                               (int*)ci.p + noff[m], (long)A.nnz, (int)coff[m]);
            MUXI_HIP(hipMemcpyAsync((float*)va.p + noff[m], A.d_values, sizeof(float) * (size_t)A.nnz,
                                    hipMemcpyDeviceToDevice, s));
        }
        MUXI_LAUNCH_CHECK();
    }
    int hbad = 0;
    MUXI_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MUXI_HIP(hipStreamSynchronize(s));
    if (hbad) {
        set_error("batch: a member's rowPtr must start at 0 and end at its nnz");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const int nend = (int)N;
    MUXI_HIP(hipMemcpyAsync(p->d_bd_rp + R, &nend, sizeof(int), hipMemcpyHostToDevice, s));
    muxi_csr_t bd{};
    bd.numRows = (int)R;
    bd.numCols = (int)C;
    bd.nnz = (int)N;
    bd.d_rowPtr = p->d_bd_rp;
    bd.d_colIdx = (int*)ci.p;
    bd.d_values = va.p;
    const spmv_status_t st = muxi_plan_create(&bd, SPMV_TYPE_FLOAT, SPMV_TYPE_FLOAT, MUXI_KERNEL_LSLAB, 4, 0, s, &p->lp);
    if (st != SPMV_SUCCESS) return st;
    MUXI_HIP(hipStreamSynchronize(s));
    p->lp->A.ci = nullptr;  // freed on return: a plain slab plan reads neither after its build
    p->lp->A.v = nullptr;
    std::vector<const float*> xp((size_t)count);
    std::vector<float*> yp((size_t)count);
    for (int m = 0; m < count; ++m) {
        xp[m] = (const float*)xs[m];
        yp[m] = (float*)ys[m];
    }
    MUXI_HIP(hipMalloc(&p->d_xcat, sizeof(float) * (size_t)std::max(C, 1L)));
    MUXI_HIP(hipMalloc(&p->d_ycat, sizeof(float) * (size_t)std::max(R, 1L)));
    MUXI_HIP(hipMalloc(&p->d_coff, sizeof(long) * (size_t)(count + 1)));
    MUXI_HIP(hipMalloc(&p->d_roff, sizeof(long) * (size_t)(count + 1)));
    MUXI_HIP(hipMalloc(&p->d_xptr, sizeof(float*) * (size_t)count));
    MUXI_HIP(hipMalloc(&p->d_yptr, sizeof(float*) * (size_t)count));
    MUXI_HIP(hipMemcpyAsync(p->d_coff, coff.data(), sizeof(long) * coff.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_roff, roff.data(), sizeof(long) * roff.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_xptr, xp.data(), sizeof(float*) * xp.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_yptr, yp.data(), sizeof(float*) * yp.size(), hipMemcpyHostToDevice, s));
    std::vector<void*> xv((size_t)count), yv((size_t)count);
    for (int m = 0; m < count; ++m) {
        xv[m] = const_cast<float*>(xp[m]);
        yv[m] = yp[m];
    }
    const std::vector<CopySeg> xf = batch_copy_segs(coff, xv, C), yf = batch_copy_segs(roff, yv, R);
    MUXI_HIP(hipMalloc(&p->d_xseg, sizeof(CopySeg) * xf.size()));
    MUXI_HIP(hipMalloc(&p->d_yseg, sizeof(CopySeg) * yf.size()));
    MUXI_HIP(hipMemcpyAsync(p->d_xseg, xf.data(), sizeof(CopySeg) * xf.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_yseg, yf.data(), sizeof(CopySeg) * yf.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipStreamSynchronize(s));  // the host vectors above go out of scope
    p->scratch = p->lp->scratch_bytes + sizeof(int) * (size_t)(R + 1) + sizeof(float) * (size_t)(C + R) +
                 2 * sizeof(long) * (size_t)(count + 1) + 2 * sizeof(void*) * (size_t)count +
                 sizeof(CopySeg) * (xf.size() + yf.size());
    return SPMV_SUCCESS;
}

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
    long rows_all = 0, cols_all = 0, nnz_all = 0;
    for (int m = 0; m < count; ++m) {
        const muxi_csr_t& A = mats[m];
        if (A.numRows < 0 || A.numCols < 0 || A.nnz < 0 || (long)A.numRows + A.nnz > 2147483647L ||
            (A.numRows > 0 && !A.d_rowPtr) || (A.nnz > 0 && (!A.d_colIdx || !A.d_values)) ||
            (A.nnz > 0 && !xs[m]) || (A.numRows > 0 && !ys[m])) {  // x is read only through nonzeros
            set_error("batch: invalid matrix " + std::to_string(m));
            return SPMV_ERROR_INVALID_MATRIX;
        }
        rows_all += A.numRows;
        cols_all += A.numCols;
        nnz_all += A.nnz;
    }
    // mode: ipt with MUXI_BATCH_SLABS set takes the slab mode; ipt == 0 takes it when the batch qualifies, and the
    // tuned merge schedule otherwise; ipt < 0 the tuned merge schedule; other ipt > 0 the merge schedule given
    const bool slabs_fit = vt == SPMV_TYPE_FLOAT && rows_all + nnz_all <= 2147483647L && cols_all <= 2147483647L &&
                           cols_all <= (long)kLsMaxSlabs * kLsSlabCols;
    const bool want_slabs = ipt > 0 && (ipt & MUXI_BATCH_SLABS) != 0;
    if (want_slabs && !slabs_fit) {
        set_error("batch: the slab mode takes FLOAT values and vectors, at most " +
                  std::to_string((long)kLsMaxSlabs * kLsSlabCols) +
                  " columns in all and rows + nonzeros within an int");
        return SPMV_ERROR_UNSUPPORTED_FORMAT;
    }
    if (want_slabs ||
        (ipt == 0 && slabs_fit && cols_all >= kBatchSlabMinCols && nnz_all >= kBatchSlabMinNnz)) {
        muxi_batch_plan* p = new muxi_batch_plan();
        p->count = count;
        p->vtype = vt;
        p->xtype = xt;
        p->ipt = 0;
        p->wg_per_cu = 0;
        p->slabs = true;
        p->host.resize(count);
        for (int m = 0; m < count; ++m) {  // the members themselves, for the per-matrix launches
            const muxi_csr_t& A = mats[m];
            p->host[m] = BatchDesc{A.d_rowPtr, A.d_colIdx, A.d_values, xs[m], ys[m], A.numRows, A.numCols, A.nnz,
                                   0, 0, 0};
        }
        const spmv_status_t st = batch_slabs_build(p, mats, count, xs, ys, (hipStream_t)stream);
        if (st != SPMV_SUCCESS) {
            batch_free(p);
            return st;
        }
        p->grid = p->lp->ls_chunks;
        *out = p;
        return SPMV_SUCCESS;
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
    for (int m = 0; m < count; ++m) {  // (every member was checked above)
        const muxi_csr_t& A = mats[m];
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

// Slab mode: x gathered into one vector, the block-diagonal slab plan into one y, y scattered with alpha and beta.
static spmv_status_t batch_slabs_go(const muxi_batch_plan* p, float alpha, float beta, hipStream_t s,
                                    hipEvent_t ev1) {
    if (p->Ctot > 0)
        hipLaunchKernelGGL(batch_gather_x_kernel, dim3(grid_for(p->Ctot, kBatchCopyPer)), dim3(kBatchCopyThreads), 0,
                           s, p->d_xseg, p->d_coff, p->count, p->d_xptr, p->d_xcat, p->Ctot);
    MUXI_LAUNCH_CHECK();
    const spmv_status_t st = muxi_plan_exec(p->lp, p->d_xcat, p->d_ycat, 1.0, 0.0, 1, s);
    if (st != SPMV_SUCCESS) return st;
    if (p->Rtot > 0)
        hipLaunchKernelGGL(batch_scatter_y_kernel, dim3(grid_for(p->Rtot, kBatchCopyPer)), dim3(kBatchCopyThreads), 0,
                           s, p->d_yseg, p->d_roff, p->count, p->d_yptr, p->d_ycat, p->Rtot, alpha, beta);
    MUXI_LAUNCH_CHECK();
    if (ev1) MUXI_HIP(hipEventRecord(ev1, s));
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_batch_exec_ev(const muxi_batch_plan* p, double alpha, double beta, void* stream,
                                            void* ev0, void* ev1) {
    if (!p) return SPMV_ERROR_INVALID_HANDLE;
    hipStream_t s = (hipStream_t)stream;
    if (ev0) MUXI_HIP(hipEventRecord((hipEvent_t)ev0, s));
    if (p->slabs) return batch_slabs_go(p, (float)alpha, (float)beta, s, (hipEvent_t)ev1);
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
