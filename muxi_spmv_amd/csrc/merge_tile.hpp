// One merge-path tile on a 256-lane workgroup: shared by the single-matrix
// merge kernel and the batched persistent kernel.
#pragma once

#include "kernels.hpp"

namespace muxi {

// One step of a wavefront's segmented scan by a DPP move: (f1,v1) then (f2,v2) -> (f1|f2, f2 ? v2 : v1+v2); a
// lane without a source takes (0, 0).
template <int CTRL, int ROWS, typename A>
__device__ __forceinline__ void seg_scan_step(A& v, int& f) {
    A u = v.template dpp<CTRL, ROWS>();
    const int fu = dpp_move<CTRL, ROWS>(f);
    if (!f) {
        u.add(v);
        v = u;
    }
    f |= fu;
}

// Workgroup-wide inclusive segmented scan over (flag, value) in lane order:
// (f1,v1) then (f2,v2) -> (f1|f2, f2 ? v2 : v1+v2). Returns the exclusive
// value (sum carried into this lane) and the workgroup total in *tot. Inside
// a wavefront: shifts of 1, 2, 4, 8 within each row of 16 lanes, then row 0's
// total into rows 1 and 3 and rows 0-1's total into rows 2 and 3 (DPP, no
// LDS); across wavefronts through LDS.
template <typename A, int NW>
__device__ __forceinline__ A block_seg_scan_excl(A v, int f, A* s_agg, int* s_flag, A* tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    seg_scan_step<0x111, 0xf>(v, f);
    seg_scan_step<0x112, 0xf>(v, f);
    seg_scan_step<0x114, 0xf>(v, f);
    seg_scan_step<0x118, 0xf>(v, f);
    seg_scan_step<0x142, 0xa>(v, f);
    seg_scan_step<0x143, 0xc>(v, f);
    if (lane == 63) {
        s_agg[w] = v;
        s_flag[w] = f;
    }
    __syncthreads();
    A pre;
    pre.zero();
    for (int u = 0; u < w; ++u) {
        if (s_flag[u])
            pre = s_agg[u];
        else
            pre.add(s_agg[u]);
    }
    A inc = v;
    if (!f) {
        A t = pre;
        t.add(v);
        inc = t;
    }
    A ex = inc.template dpp<0x138, 0xf>();  // wave_shr:1
    if (lane == 0) ex = pre;
    A t;
    t.zero();
    for (int u = 0; u < NW; ++u) {
        if (s_flag[u])
            t = s_agg[u];
        else
            t.add(s_agg[u]);
    }
    *tot = t;
    return ex;
}

template <typename AT, bool COMP, int IPT>
struct MergeSmem {
    int end[256 * IPT];
    Prod<AT, COMP> prod[256 * IPT];
    Acc<AT, COMP> agg[4];
    int flag[4];
};

// Processes path items [d0, d1) with coordinates (r0, d0-r0) .. (r1, d1-r1):
// writes y for rows r0..r1-1 and returns the partial sum of row r1 held in
// this tile (the same value in every lane). Ends with LDS still in use: the
// caller must __syncthreads() before reusing sm.
template <typename VT, typename XT, typename AT, bool COMP, int IPT>
__device__ __forceinline__ Acc<AT, COMP> merge_tile(MergeSmem<AT, COMP, IPT>& sm, const int* __restrict__ rp,
                                                    const int* __restrict__ ci, const VT* __restrict__ vals,
                                                    const XT* __restrict__ x, XT* __restrict__ y, int d0, int r0,
                                                    int d1, int r1, AT alpha, AT beta) {
    constexpr int B = 256;
    using A = Acc<AT, COMP>;
    using P = Prod<AT, COMP>;
    const int tid = threadIdx.x;
    const int n0 = d0 - r0, n1 = d1 - r1;
    const int nr = r1 - r0, nn = n1 - n0, tot = nr + nn;
    // Load phase. Indices past the tile end are clamped to its last item, so
    // no load sits behind a per-item branch: all IPT column, value and row-end
    // loads, then all IPT x gathers, are in flight together. (A guarded load
    // per item makes the compiler wait for each gather before the next.) The
    // nn > 0 / nr > 0 tests are uniform over the workgroup.
    int c[IPT], e[IPT];
    VT v[IPT];
    if (nn > 0) {
#pragma unroll
        for (int k = 0; k < IPT; ++k) c[k] = ntl(ci + n0 + min(tid + k * B, nn - 1));
#pragma unroll
        for (int k = 0; k < IPT; ++k) v[k] = ntl(vals + n0 + min(tid + k * B, nn - 1));
    }
    if (nr > 0) {
#pragma unroll
        for (int k = 0; k < IPT; ++k) e[k] = ntl(rp + r0 + 1 + min(tid + k * B, nr - 1));
    }
    if (nn > 0) {
        XT xv[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) xv[k] = x[c[k]];
#pragma unroll
        for (int k = 0; k < IPT; ++k)
            if (tid + k * B < nn) sm.prod[tid + k * B] = P::make(v[k], xv[k]);
    }
    if (nr > 0) {
#pragma unroll
        for (int k = 0; k < IPT; ++k)
            if (tid + k * B < nr) sm.end[tid + k * B] = e[k] - n0;
    }
    __syncthreads();

    const int di = min(tid * IPT, tot);
    const int de = min(di + IPT, tot);
    int lo = max(0, di - nn), hi = min(di, nr);
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (sm.end[mid] <= di - mid - 1)
            lo = mid + 1;
        else
            hi = mid;
    }
    const int rs = lo, ns = di - lo;

    // pass 1: does this lane close a row, and what is left open at its end
    A acc;
    acc.zero();
    int fl = 0;
    {
        int r = rs, n = ns;
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            if (di + k < de) {
                if (r < nr && sm.end[r] <= n) {
                    acc.zero();
                    fl = 1;
                    ++r;
                } else {
                    sm.prod[n].into(acc);
                    ++n;
                }
            }
        }
    }
    A total;
    A ex = block_seg_scan_excl<A, B / 64>(acc, fl, sm.agg, sm.flag, &total);

    // pass 2: emit rows, the first one including the sum carried in
    {
        A a = ex;
        int r = rs, n = ns;
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            if (di + k < de) {
                if (r < nr && sm.end[r] <= n) {
                    store_y(y, (long)r0 + r, a.value(), alpha, beta);
                    a.zero();
                    ++r;
                } else {
                    sm.prod[n].into(a);
                    ++n;
                }
            }
        }
    }
    return total;
}

}  // namespace muxi
