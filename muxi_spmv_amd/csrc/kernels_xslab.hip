// Column-slab merge-path SpMV ("xslab"): keeps each XCD's x gathers in its
// own L2.
//
// MI355X has 8 XCDs, each with a private 4 MiB L2, and hands workgroup i to
// XCD i mod 8. When x is larger than one L2 (the 2M-column power-law matrix
// has an 8 MB x) and the columns are spread, every XCD pulls all of x through
// its L2: on that matrix the plain merge kernel moves ~1.1 GB through L2 for
// 139 MB of matrix and vectors (TCC_MISS x 128 B). The plan splits the
// columns into S slabs (column c is in slab floor(c*S/C)) and stores, per
// slab, a compressed CSR of the nonzeros that fall in it: slab-local row
// pointers over the rows that have any, columns and values. Workgroup b works
// on slab b mod S, so with S = 8 an XCD gathers from one 1/8 slice of x that
// stays in its L2; the matrix stream uses non-temporal loads so it does not
// evict that slice.
//
// A (row, slab) pair with nonzeros is a "pair"; pairs are numbered row-major,
// so a row's pairs are contiguous. Each slab runs the merge-path algorithm
// over its compressed rows and stores a row's partial sum to P[pair] (plain
// stores, no atomics); the partial of the row open at a tile end goes to a
// carry keyed by its pair, and a keyed fix-up adds carries in tile order. A
// combine kernel then forms y[r] = alpha * (sum of the row's pairs, in slab
// order) + beta * y[r]. Every sum has a fixed order: results are reproducible.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <vector>

#include "merge_tile.hpp"
#include "plan.hpp"

namespace muxi {

// Registers of one tile in flight: columns, values, row ends and gathered x.
template <typename VT, typename XT, int IPT>
struct TileRegs {
    int c[IPT], e[IPT];
    VT v[IPT];
    XT x[IPT];
};

struct TileMeta {
    int n0, r0, nn, nr;
};

// Persistent, software-pipelined column-slab kernel. Workgroup b works on
// slab b mod S (so on one XCD's x slice) and walks a contiguous run of that
// slab's merge tiles. While it reduces tile t out of LDS, the x gathers of
// tile t+1 and the column/value/row-end loads of tile t+2 are in flight, so
// the memory system never waits for the reduction. Clamped indices keep all
// loads unconditional. A row's partial goes to P[row_base[s] + slab row]
// (slab-local order: consecutive rows store to consecutive slots); the row
// left open at a tile end goes to the tile's carry, keyed by its slot.
template <typename VT, typename XT, typename AT, int IPT>
__global__ __launch_bounds__(256) void xslab_kernel(SlabArgs sa, int W, const int* __restrict__ srp,
                                                    const int* __restrict__ sci, const VT* __restrict__ svals,
                                                    const XT* __restrict__ x, const int* __restrict__ gdiag,
                                                    const int* __restrict__ grow, AT* __restrict__ P,
                                                    AT* __restrict__ carry, int* __restrict__ carry_key) {
    constexpr int B = 256, CH = B * IPT;
    using A = Acc<AT, false>;
    __shared__ int s_end[CH];
    __shared__ AT s_prod[CH];
    __shared__ A s_agg[4];
    __shared__ int s_flag[4];
    const int s = blockIdx.x % sa.S, j = blockIdx.x / sa.S;
    const int T = sa.tiles[s];
    const int tb = (int)((long)j * T / W), te = (int)((long)(j + 1) * T / W);
    if (tb >= te) return;  // uniform over the workgroup
    const int tid = threadIdx.x;
    const int* rp = srp + sa.row_base[s] + s;
    const int* ci = sci + sa.nnz_base[s];
    const VT* vals = svals + sa.nnz_base[s];
    const int* gd = gdiag + sa.coord_base[s];
    const int* gr = grow + sa.coord_base[s];
    const int rows = sa.rows[s], pbase = sa.row_base[s], cbase = sa.tile_base[s];

    auto meta = [&](int t) {
        TileMeta m;
        const int d0 = gd[t], r0 = gr[t], d1 = gd[t + 1], r1 = gr[t + 1];
        m.n0 = d0 - r0;
        m.r0 = r0;
        m.nn = (d1 - r1) - m.n0;
        m.nr = r1 - r0;
        return m;
    };
    auto loads = [&](const TileMeta& m, TileRegs<VT, XT, IPT>& g) {
        if (m.nn > 0) {
#pragma unroll
            for (int k = 0; k < IPT; ++k) g.c[k] = ntl(ci + m.n0 + min(tid + k * B, m.nn - 1));
#pragma unroll
            for (int k = 0; k < IPT; ++k) g.v[k] = ntl(vals + m.n0 + min(tid + k * B, m.nn - 1));
        }
        if (m.nr > 0) {
#pragma unroll
            for (int k = 0; k < IPT; ++k) g.e[k] = ntl(rp + m.r0 + 1 + min(tid + k * B, m.nr - 1));
        }
    };
    auto gathers = [&](const TileMeta& m, TileRegs<VT, XT, IPT>& g) {
        if (m.nn > 0) {
#pragma unroll
            for (int k = 0; k < IPT; ++k) g.x[k] = x[g.c[k]];
        }
    };
    auto stage = [&](const TileMeta& m, const TileRegs<VT, XT, IPT>& g) {
        if (m.nn > 0) {
#pragma unroll
            for (int k = 0; k < IPT; ++k)
                if (tid + k * B < m.nn) s_prod[tid + k * B] = Conv<AT>::f(g.v[k]) * Conv<AT>::f(g.x[k]);
        }
        if (m.nr > 0) {
#pragma unroll
            for (int k = 0; k < IPT; ++k)
                if (tid + k * B < m.nr) s_end[tid + k * B] = g.e[k] - m.n0;
        }
    };

    // Reduces tile t (staged in LDS) while tile t+1's gathers (into G1) and
    // tile t+2's loads (into G2) are in flight, then stages tile t+1. The
    // loop below alternates the two register sets so no set is copied (a copy
    // would wait for its loads).
    TileMeta cur = meta(tb), nb{0, 0, 0, 0}, nc{0, 0, 0, 0};
    auto step = [&](int t, TileRegs<VT, XT, IPT>& G1, TileRegs<VT, XT, IPT>& G2) {
        if (t + 1 < te) gathers(nb, G1);
        if (t + 2 < te) {
            nc = meta(t + 2);
            loads(nc, G2);
        }
        const int nr = cur.nr, nn = cur.nn, tot = nr + nn, r0 = cur.r0;
        const int di = min(tid * IPT, tot), de = min(di + IPT, tot);
        int lo = max(0, di - nn), hi = min(di, nr);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s_end[mid] <= di - mid - 1)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int rs = lo, ns = di - lo;
        A acc;
        acc.zero();
        int fl = 0;
        {
            int r = rs, n = ns;
#pragma unroll
            for (int k = 0; k < IPT; ++k)
                if (di + k < de) {
                    if (r < nr && s_end[r] <= n) {
                        acc.zero();
                        fl = 1;
                        ++r;
                    } else {
                        acc.s += s_prod[n++];
                    }
                }
        }
        A total;
        A ex = block_seg_scan_excl<A, B / 64>(acc, fl, s_agg, s_flag, &total);
        {
            A a = ex;
            int r = rs, n = ns;
#pragma unroll
            for (int k = 0; k < IPT; ++k)
                if (di + k < de) {
                    if (r < nr && s_end[r] <= n) {
                        P[pbase + r0 + r] = a.s;
                        a.zero();
                        ++r;
                    } else {
                        a.s += s_prod[n++];
                    }
                }
        }
        if (tid == 0) {
            const int r1 = r0 + nr, n1 = cur.n0 + nn;
            const bool open = r1 < rows && n1 > rp[r1];
            carry[cbase + t] = open ? total.s : AT(0);
            carry_key[cbase + t] = open ? pbase + r1 : -1;
        }
        __syncthreads();  // every lane is done with LDS (and with s_agg)
        if (t + 1 < te) {
            stage(nb, G1);
            cur = nb;
            nb = nc;
        }
        __syncthreads();
    };

    TileRegs<VT, XT, IPT> ga, gb;
    loads(cur, ga);
    gathers(cur, ga);
    stage(cur, ga);
    if (tb + 1 < te) {
        nb = meta(tb + 1);
        loads(nb, gb);
    }
    __syncthreads();
    for (int t = tb; t < te; t += 2) {
        step(t, gb, ga);
        if (t + 1 < te) step(t + 1, ga, gb);
    }
}

// Same run logic as carry_fixup_kernel with explicit keys (-1 = nothing):
// P[key] += sum of the run's carries, in tile order.
template <typename AT>
__global__ __launch_bounds__(256) void keyed_fixup_kernel(const int* __restrict__ keys, int count,
                                                          const AT* __restrict__ carry, AT* __restrict__ P) {
    const int lane = threadIdx.x & 63;
    const long base = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64 * 64;
    if (base >= count) return;
    auto key_of = [&](long t) -> int { return t < count ? keys[t] : -2; };
    const long t = base + lane;
    const int key = key_of(t);
    AT v = t < count ? carry[t] : AT(0);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        AT vu = __shfl_up(v, d);
        int ku = __shfl_up(key, d);
        if (lane >= d && ku == key) v = vu + v;
    }
    int knext = __shfl_down(key, 1);
    if (lane == 63) knext = key_of(base + 64);
    const int key0 = __shfl(key, 0);
    const bool owned = !((base > 0) && (key_of(base - 1) == key0) && key == key0);
    if (knext != key && owned && key >= 0) P[key] += v;
    const int k63 = __shfl(key, 63), kn63 = __shfl(knext, 63);
    const bool own63 = __shfl((int)owned, 63);
    if (kn63 == k63 && own63 && k63 >= 0) {
        AT sum = __shfl(v, 63);
        for (long nb = base + 64; nb < count; nb += 64) {
            const long tt = nb + lane;
            const int kk = key_of(tt);
            AT c = (kk == k63) ? carry[tt] : AT(0);
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
            sum += c;
            if (__shfl(kk, 63) != k63) break;
        }
        if (lane == 0) P[k63] += sum;
    }
}

// y[r] = alpha * (sum of the row's pair partials, in slab order) + beta * y[r].
// The row's pairs are ploc[rpp[r] .. rpp[r+1]): slots of P in slab order.
template <typename XT, typename AT>
__global__ __launch_bounds__(256) void xslab_combine_kernel(int R, const int* __restrict__ rpp,
                                                            const int* __restrict__ ploc, const AT* __restrict__ P,
                                                            XT* __restrict__ y, AT alpha, AT beta) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    AT s = AT(0);
    const int b = rpp[r], n = rpp[r + 1] - b;
    if (n > 0) {
        // the first four pairs with clamped (unconditional) loads, all in flight together
        int q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = ntl(ploc + b + (k < n ? k : n - 1));
        AT v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = ntl(P + q[k]);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < n) s += v[k];
        for (int k = 4; k < n; ++k) s += ntl(P + ntl(ploc + b + k));
    }
    store_y(y, r, s, alpha, beta);
}

// ---- plan construction
__device__ __forceinline__ int slab_of(int col, int S, int C) { return (int)(((long)col * S) / C); }

// One wavefront per row (rows reach 10^5 nonzeros): the row is read in
// 64-wide chunks and lane q < S counts slab q's nonzeros from ballots.
__global__ __launch_bounds__(256) void xslab_count_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                          int R, int C, int S, int* __restrict__ cnt,
                                                          int* __restrict__ flag, int* __restrict__ npair) {
    const long r = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= R) return;  // whole wavefront
    const int b = rp[r], e = rp[r + 1];
    int c = 0;
    for (int i = b; i < e; i += 64) {
        const int s = i + lane < e ? slab_of(ci[i + lane], S, C) : -1;
        for (int q = 0; q < S; ++q) {
            const unsigned long long m = __ballot(s == q);
            if (lane == q) c += __popcll(m);
        }
    }
    if (lane < S) {
        cnt[(long)lane * R + r] = c;
        flag[(long)lane * R + r] = c > 0;
    }
    const int np = __popcll(__ballot(lane < S && c > 0));
    if (lane == 0) npair[r] = np;
}

// One wavefront per row. Lane q < S holds slab q's write offset; each 64-wide
// chunk of the row is scattered with ballot ranks, which keeps the row's
// order inside every slab.
template <typename VT>
__global__ __launch_bounds__(256) void xslab_fill_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                         const VT* __restrict__ v, int R, int C, int S,
                                                         const int* __restrict__ cnt, const int* __restrict__ pos,
                                                         const int* __restrict__ lrow,
                                                         const int* __restrict__ nnz_base,
                                                         const int* __restrict__ rpp, int* __restrict__ srp,
                                                         int* __restrict__ ploc, int* __restrict__ sci,
                                                         VT* __restrict__ sv) {
    const long r = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= R) return;  // whole wavefront
    int off = 0;
    {
        const long k = (long)(lane < S ? lane : 0) * R + r;
        const bool has = lane < S && cnt[k] > 0;  // row r has nonzeros in slab `lane`
        const unsigned long long hm = __ballot(has);
        if (lane < S) off = pos[k];
        if (has) {
            const int lr = lrow[k];
            ploc[rpp[r] + __popcll(hm & __lanemask_lt())] = lr;  // the row's pairs in slab order
            srp[lr + lane] = off - nnz_base[lane];
        }
    }
    const int b = rp[r], e = rp[r + 1];
    for (int i = b; i < e; i += 64) {
        const bool in = i + lane < e;
        const int col = in ? ci[i + lane] : 0;
        const int s = in ? slab_of(col, S, C) : -1;
        int o = 0;
        for (int q = 0; q < S; ++q) {
            const unsigned long long m = __ballot(s == q);
            const int base = __shfl(off, q);
            if (s == q) o = base + __popcll(m & __lanemask_lt());
            if (lane == q) off += __popcll(m);
        }
        if (in) {
            sci[o] = col;
            sv[o] = v[i + lane];
        }
    }
}

static spmv_status_t exscan(const int* in, int* out, long n, hipStream_t s) {
    size_t tmp = 0;
    MUXI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, s));
    void* d_tmp = nullptr;
    MUXI_HIP(hipMallocAsync(&d_tmp, tmp ? tmp : 16, s));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp, in, out, (int)n, s);
    (void)hipFreeAsync(d_tmp, s);
    if (e != hipSuccess) {
        set_error(std::string("scan: ") + hipGetErrorString(e));
        return SPMV_ERROR_DEVICE;
    }
    return SPMV_SUCCESS;
}

template <typename VT, typename XT, typename AT>
using XslabFn = void (*)(SlabArgs, int, const int*, const int*, const VT*, const XT*, const int*, const int*, AT*, AT*,
                         int*);

template <typename VT, typename XT, typename AT>
static XslabFn<VT, XT, AT> xslab_fn(int ipt) {
    switch (ipt) {
        case 4: return xslab_kernel<VT, XT, AT, 4>;
        case 16: return xslab_kernel<VT, XT, AT, 16>;
        default: return xslab_kernel<VT, XT, AT, 8>;
    }
}

// Workgroups per slab for the persistent kernel: the slab's share of what
// the device holds resident at once, and no more than the slab's tiles.
template <typename VT, typename XT, typename AT>
static spmv_status_t xslab_workgroups(muxi_plan* p) {
    int cus = 0, per_cu = 0;
    MUXI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
    MUXI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, xslab_fn<VT, XT, AT>(p->param0 & 0xff), 256, 0));
    const int S = p->slab.S;
    int W = cus * (per_cu > 0 ? per_cu : 1) / S;
    W = W > p->P ? p->P : W;  // p->P: the most tiles any slab has
    p->slab_W = W > 0 ? W : 1;
    return SPMV_SUCCESS;
}

spmv_status_t xslab_build(muxi_plan* p, int S, int ipt, spmv_type_t vt, hipStream_t s) {
    const int R = p->A.R, C = p->A.C, N = p->A.N;
    if (S < 1 || S > kMaxSlabs) {
        set_error("xslab: slabs must be in [1, 16]");
        return SPMV_ERROR_INTERNAL;
    }
    if ((long)S * R + 1 > INT_MAX) {
        set_error("xslab: rows * slabs exceeds INT32_MAX");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const long SR = (long)S * R;
    int *cnt, *flag, *pos, *lrow, *npair;
    MUXI_HIP(hipMallocAsync(&cnt, sizeof(int) * (SR + 1), s));
    MUXI_HIP(hipMallocAsync(&flag, sizeof(int) * (SR + 1), s));
    MUXI_HIP(hipMallocAsync(&pos, sizeof(int) * (SR + 1), s));
    MUXI_HIP(hipMallocAsync(&lrow, sizeof(int) * (SR + 1), s));
    MUXI_HIP(hipMallocAsync(&npair, sizeof(int) * ((size_t)R + 1), s));
    MUXI_HIP(hipMemsetAsync(cnt + SR, 0, sizeof(int), s));
    MUXI_HIP(hipMemsetAsync(flag + SR, 0, sizeof(int), s));
    MUXI_HIP(hipMemsetAsync(npair + R, 0, sizeof(int), s));
    hipLaunchKernelGGL(xslab_count_kernel, dim3(grid_for((long)R * 64, 256)), dim3(256), 0, s, p->A.rp, p->A.ci, R, C,
                       S, cnt, flag, npair);
    MUXI_LAUNCH_CHECK();
    // row-major pair offsets rpp[R+1] live in the plan
    MUXI_HIP(hipMalloc(&p->d_slab_rpp, sizeof(int) * ((size_t)R + 1)));
    spmv_status_t st = exscan(cnt, pos, SR + 1, s);
    if (st == SPMV_SUCCESS) st = exscan(flag, lrow, SR + 1, s);
    if (st == SPMV_SUCCESS) st = exscan(npair, p->d_slab_rpp, (long)R + 1, s);
    if (st != SPMV_SUCCESS) return st;
    std::vector<int> hpos(S + 1), hrow(S + 1);
    for (int q = 0; q <= S; ++q) {
        MUXI_HIP(hipMemcpyAsync(&hpos[q], pos + (long)q * R, sizeof(int), hipMemcpyDeviceToHost, s));
        MUXI_HIP(hipMemcpyAsync(&hrow[q], lrow + (long)q * R, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    MUXI_HIP(hipStreamSynchronize(s));
    const int total_rows = hrow[S];  // number of pairs
    SlabArgs& sa = p->slab;
    sa.S = S;
    for (int q = 0; q < S; ++q) {
        sa.rows[q] = hrow[q + 1] - hrow[q];
        sa.row_base[q] = hrow[q];
        sa.nnz_base[q] = hpos[q];
    }
    int* d_nb = nullptr;
    MUXI_HIP(hipMallocAsync(&d_nb, sizeof(int) * S, s));
    MUXI_HIP(hipMemcpyAsync(d_nb, hpos.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
    const size_t vs = vsize(vt);
    const size_t ints = (size_t)total_rows * 2 + S + (size_t)N;
    MUXI_HIP(hipMalloc(&p->d_slab, sizeof(int) * ints + vs * (size_t)(N > 0 ? N : 1) + 256));
    p->d_slab_rp = p->d_slab;
    p->d_slab_ploc = p->d_slab_rp + total_rows + S;
    p->d_slab_ci = p->d_slab_ploc + total_rows;
    uintptr_t va = (uintptr_t)(p->d_slab_ci + N);
    va = (va + 255) & ~(uintptr_t)255;
    p->d_slab_v = (void*)va;
    std::vector<int> term(S);
    for (int q = 0; q < S; ++q) term[q] = hpos[q + 1] - hpos[q];
    for (int q = 0; q < S; ++q)
        MUXI_HIP(hipMemcpyAsync(p->d_slab_rp + sa.row_base[q] + q + sa.rows[q], &term[q], sizeof(int),
                                hipMemcpyHostToDevice, s));
    st = with_types(vt, p->xtype, p->comp, [&](auto tag) -> spmv_status_t {
        using T = decltype(tag);
        using VT = typename T::VT;
        hipLaunchKernelGGL(xslab_fill_kernel<VT>, dim3(grid_for((long)R * 64, 256)), dim3(256), 0, s, p->A.rp, p->A.ci,
                           (const VT*)p->A.v, R, C, S, cnt, pos, lrow, d_nb, p->d_slab_rpp, p->d_slab_rp,
                           p->d_slab_ploc, p->d_slab_ci, (VT*)p->d_slab_v);
        MUXI_LAUNCH_CHECK();
        return SPMV_SUCCESS;
    });
    if (st != SPMV_SUCCESS) return st;
    const long CH = 256L * ipt;
    long ncoord = 0, ntiles = 0;
    for (int q = 0; q < S; ++q) {
        const int nq = hpos[q + 1] - hpos[q];
        sa.tiles[q] = sa.rows[q] > 0 ? (int)(((long)sa.rows[q] + nq + CH - 1) / CH) : 0;
        sa.coord_base[q] = (int)ncoord;
        sa.tile_base[q] = (int)ntiles;
        ncoord += sa.tiles[q] + 1;
        ntiles += sa.tiles[q];
    }
    p->T = (int)ntiles;
    const size_t acc = p->xtype == SPMV_TYPE_DOUBLE ? 8 : 4;
    MUXI_HIP(hipMalloc(&p->d_pos, 2 * sizeof(int) * (size_t)ncoord));
    p->d_row = p->d_pos + ncoord;
    MUXI_HIP(hipMalloc(&p->d_carry, acc * (size_t)(ntiles + total_rows + 1)));  // carries, then P
    MUXI_HIP(hipMalloc(&p->d_slab_ckey, sizeof(int) * (size_t)(ntiles + 1)));
    for (int q = 0; q < S; ++q) {
        if (sa.tiles[q] == 0) continue;
        const int nq = hpos[q + 1] - hpos[q];
        st = launch_merge_partitions(sa.rows[q], nq, sa.tiles[q], p->d_pos + sa.coord_base[q], s);
        if (st == SPMV_SUCCESS)
            st = launch_merge_coords(p->d_slab_rp + sa.row_base[q] + q, sa.rows[q], nq, p->d_pos + sa.coord_base[q],
                                     sa.tiles[q] + 1, p->d_row + sa.coord_base[q], nullptr, s);
        if (st != SPMV_SUCCESS) return st;
    }
    int maxt = 0;
    for (int q = 0; q < S; ++q) maxt = sa.tiles[q] > maxt ? sa.tiles[q] : maxt;
    p->P = maxt;
    st = with_types(vt, p->xtype, false, [&](auto tag) -> spmv_status_t {
        using T = decltype(tag);
        return xslab_workgroups<typename T::VT, typename T::XT, typename T::AT>(p);
    });
    if (st != SPMV_SUCCESS) return st;
    (void)hipFreeAsync(cnt, s);
    (void)hipFreeAsync(flag, s);
    (void)hipFreeAsync(pos, s);
    (void)hipFreeAsync(lrow, s);
    (void)hipFreeAsync(npair, s);
    (void)hipFreeAsync(d_nb, s);
    MUXI_HIP(hipStreamSynchronize(s));
    p->scratch_bytes = sizeof(int) * ints + vs * (size_t)N + 2 * sizeof(int) * (size_t)ncoord +
                       acc * (size_t)(ntiles + total_rows) + sizeof(int) * ((size_t)ntiles + R + 1);
    return SPMV_SUCCESS;
}

template <typename VT, typename XT, typename AT>
spmv_status_t launch_xslab(const muxi_plan* p, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s,
                           hipEvent_t ev_end) {
    if (p->A.R <= 0) return SPMV_SUCCESS;
    const SlabArgs& sa = p->slab;
    AT* carry = (AT*)p->d_carry;
    AT* P = carry + p->T + 1;
    if (p->P > 0) {
        const dim3 grid((unsigned)((long)p->slab_W * sa.S));
        const XslabFn<VT, XT, AT> fn = xslab_fn<VT, XT, AT>(p->param0 & 0xff);
        hipLaunchKernelGGL(fn, grid, dim3(256), 0, s, sa, p->slab_W,
                           p->d_slab_rp, p->d_slab_ci, (const VT*)p->d_slab_v, x, p->d_pos, p->d_row, P, carry,
                           p->d_slab_ckey);
        MUXI_LAUNCH_CHECK();
        if (ev_end) MUXI_HIP(hipEventRecord(ev_end, s));
        hipLaunchKernelGGL(keyed_fixup_kernel<AT>, dim3(grid_for(p->T, 256)), dim3(256), 0, s, p->d_slab_ckey, p->T,
                           carry, P);
        MUXI_LAUNCH_CHECK();
    } else if (ev_end) {
        MUXI_HIP(hipEventRecord(ev_end, s));
    }
    hipLaunchKernelGGL((xslab_combine_kernel<XT, AT>), dim3(grid_for(p->A.R, 256)), dim3(256), 0, s, p->A.R,
                       p->d_slab_rpp, p->d_slab_ploc, P, y, alpha, beta);
    MUXI_LAUNCH_CHECK();
    return SPMV_SUCCESS;
}

template spmv_status_t launch_xslab<float, float, float>(const muxi_plan*, const float*, float*, float, float,
                                                         hipStream_t, hipEvent_t);
template spmv_status_t launch_xslab<bf16_t, float, float>(const muxi_plan*, const float*, float*, float, float,
                                                          hipStream_t, hipEvent_t);
template spmv_status_t launch_xslab<double, float, float>(const muxi_plan*, const float*, float*, float, float,
                                                          hipStream_t, hipEvent_t);
template spmv_status_t launch_xslab<double, double, double>(const muxi_plan*, const double*, double*, double,
                                                            double, hipStream_t, hipEvent_t);

}  // namespace muxi
