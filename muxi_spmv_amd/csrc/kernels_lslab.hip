// LDS column-slab SpMV ("lslab"): x gathers from LDS instead of L2.
//
// With uniformly scattered columns every nonzero of a plain CSR kernel costs
// one random 4-byte gather, and L2 serves those at one request per gather:
// on MI355X about 220-270 G gathers/s, a quarter of what the HBM stream of
// the matrix itself could sustain. This plan cuts the columns into slabs of
// 32,768 (the slab of FP32 x, 128 KiB, fits one CU's LDS) and runs two
// launches:
//
// 1. Slab kernel. The nonzeros of each slab, in row order, form one stream
//    of 16-bit column offsets (bit 15 set on the last nonzero of the row
//    inside the slab) and values. A stream is cut at row ends into chunks
//    of about equal size, padded to whole tiles of 64*EPL nonzeros; one
//    1024-lane workgroup per chunk copies its slab of x into LDS once and
//    then walks its tiles, one wavefront per tile: a lane loads its EPL
//    offsets and values with one vector load each, gathers x from LDS,
//    sums its products cutting at its row ends, and a segmented wave scan
//    in DPP carries open sums between lanes. The sum of a (row, slab) pair
//    goes to P[pair], pairs numbered in stream order. A pair that starts in
//    an earlier tile of the chunk is finished after a barrier from the
//    tiles' head and tail sums kept in LDS (added in a fixed order).
// 2. Combine kernel. Rows are cut into blocks (at most 4,096 rows and a
//    fixed number of pairs). A workgroup copies its block's pair sums from
//    every slab (one contiguous run per slab) into LDS, then reads the
//    block's pairs in row-major order through a static map (LDS index,
//    row, last-of-row bit), adds each row's pairs in slab order with a
//    segmented wave scan, and writes y for its rows in one coalesced pass.
//
// The matrix is read once as 6 bytes a nonzero (FP32 values), the pairs
// (about 0.44 per nonzero on the 2M power-law matrix) cost 4 bytes written
// and 8 read, and no access goes to L2 at random. Every sum has a fixed
// order, so results are reproducible bit for bit.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"

namespace muxi {

// ---------------------------------------------------------------- pair sums
// A (row, slab) pair's sum. K bit 0: Dot2 (Ogita, Rump, Oishi 2005): TwoProduct by FMA, TwoSum, errors summed in
// c. K bit 1 (checked plans): the sum of |products| rides along in a, for the row's a-posteriori error bound. The
// Dot2 arithmetic is written out with contraction off: a product fused into the next sum's addition would leave
// TwoSum computing the error of a different sum.
template <int K>
struct LsPair {
    static constexpr bool kComp = (K & 1) != 0, kAbs = (K & 2) != 0;
    static constexpr int NF = 1 + (kComp ? 1 : 0) + (kAbs ? 1 : 0);
    float f[NF];  // f[0] the sum; f[1] its compensation (Dot2); f[NF - 1] the sum of |products| (checked)

    __device__ __forceinline__ static LsPair zero() {
        LsPair r;
#pragma unroll
        for (int i = 0; i < NF; ++i) r.f[i] = 0.f;
        return r;
    }
    __device__ __forceinline__ void add(const LsPair& o) {
#pragma clang fp contract(off)
        if constexpr (kComp) {
            const float t = f[0] + o.f[0];
            const float bp = t - f[0];
            const float e = (f[0] - (t - bp)) + (o.f[0] - bp);
            f[0] = t;
            f[1] = f[1] + (o.f[1] + e);
        } else {
            f[0] += o.f[0];
        }
        if constexpr (kAbs) f[NF - 1] += o.f[NF - 1];
    }
    template <typename VT>
    __device__ __forceinline__ static LsPair prod(VT v, float xv) {
#pragma clang fp contract(off)
        const float a = to_f32(v);
        LsPair r;
        r.f[0] = a * xv;
        if constexpr (kComp) r.f[1] = __builtin_fmaf(a, xv, -r.f[0]);  // the product's exact rounding error
        if constexpr (kAbs) r.f[NF - 1] = __builtin_fabsf(r.f[0]);
        return r;
    }
    __device__ __forceinline__ float value() const {
#pragma clang fp contract(off)
        if constexpr (kComp) return f[0] + f[1];
        return f[0];
    }
    __device__ __forceinline__ float abs_sum() const { return kAbs ? f[NF - 1] : 0.f; }
};

// DPP cross-lane moves (GFX9 encodings: row_shr:n = 0x110 | n within a row of
// 16 lanes, row_bcast:15 = 0x142, row_bcast:31 = 0x143, wave_shr:1 = 0x138);
// lanes without a source read 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ int ls_dpp(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS, int K>
__device__ __forceinline__ LsPair<K> ls_dpp(const LsPair<K>& v) {
    LsPair<K> r;
#pragma unroll
    for (int i = 0; i < LsPair<K>::NF; ++i) r.f[i] = __int_as_float(ls_dpp<CTRL, ROWS>(__float_as_int(v.f[i])));
    return r;
}

// One step of an inclusive segmented scan: a lane whose flag is set starts a
// segment and takes nothing from below.
template <int CTRL, int ROWS, int K>
__device__ __forceinline__ void ls_seg_step(LsPair<K>& v, int& f) {
    LsPair<K> u = ls_dpp<CTRL, ROWS>(v);
    const int fu = ls_dpp<CTRL, ROWS>(f);
    if (!f) {
        u.add(v);
        v = u;
    }
    f |= fu;
}
// Shifts of 1, 2, 4, 8 inside each row of 16 lanes, then row 0's total into
// rows 1 and 3 and rows 0-1's total into rows 2 and 3.
template <int K>
__device__ __forceinline__ void ls_seg_scan(LsPair<K>& v, int& f) {
    ls_seg_step<0x111, 0xf>(v, f);
    ls_seg_step<0x112, 0xf>(v, f);
    ls_seg_step<0x114, 0xf>(v, f);
    ls_seg_step<0x118, 0xf>(v, f);
    ls_seg_step<0x142, 0xa>(v, f);
    ls_seg_step<0x143, 0xc>(v, f);
}
template <int K>
__device__ __forceinline__ LsPair<K> ls_readlane63(const LsPair<K>& v) {
    LsPair<K> r;
#pragma unroll
    for (int i = 0; i < LsPair<K>::NF; ++i)
        r.f[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.f[i]), 63));
    return r;
}

// Exclusive prefix over the wave of a per-lane count in [0, 15], and the total.
__device__ __forceinline__ int ls_count_prefix(int n, int& total) {
    int pre = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const unsigned long long m = __ballot((n >> b) & 1);
        pre += __popcll(m & __lanemask_lt()) << b;
        tot += __popcll(m) << b;
    }
    total = tot;
    return pre;
}

// N consecutive elements with one non-temporal vector load (aligned: tiles start on 64*EPL elements).
template <typename T, int N>
struct LsLoad {
    __device__ __forceinline__ static void run(const T* p, T* out) {
        typedef T V __attribute__((ext_vector_type(N)));
        const V v = __builtin_nontemporal_load((const V*)p);
#pragma unroll
        for (int j = 0; j < N; ++j) out[j] = v[j];
    }
};
template <int N>
struct LsLoad<bf16_t, N> {
    __device__ __forceinline__ static void run(const bf16_t* p, bf16_t* out) {
        typedef uint16_t V __attribute__((ext_vector_type(N)));
        const V v = __builtin_nontemporal_load((const V*)p);
#pragma unroll
        for (int j = 0; j < N; ++j) out[j].bits = v[j];
    }
};

// ---------------------------------------------------------------- the slab kernel
// chunks[4b ..] = (tile0, ntiles, first column of the slab, run table offset); ntiles = 0 is an empty launch
// slot. tiles[t] = (index of the pair holding tile t's first element, bit 31 set when that pair started in an
// earlier tile; the chunk run that pair falls in). Pairs are numbered in stream order (slab, then row); a
// chunk's pairs split into runs by combine block, runs[off + r] = (first pair of run r, its offset) with nine
// sentinels (first pair INT_MAX) after the last run, and pair j of run r goes to P[j + offset]: every combine
// block then finds its pair sums in one contiguous range of P. A checked plan's count of recomputed rows
// (chk_count, else null) is cleared here, before the combine kernel counts them.
template <typename VT, int K, int EPL, int W>
__global__ __launch_bounds__(kLsThreads) void lslab_slab_kernel(const int4* __restrict__ chunks,
                                                                const uint16_t* __restrict__ lc,
                                                                const VT* __restrict__ lv,
                                                                const int2* __restrict__ tiles,
                                                                const int2* __restrict__ runs,
                                                                const float* __restrict__ x, int C,
                                                                LsPair<K>* __restrict__ P,
                                                                int* __restrict__ chk_count) {
    using PT = LsPair<K>;
    constexpr int TILE = 64 * EPL;
    constexpr int NWAVE = kLsThreads / 64;
    static_assert(lslab_fits(W, EPL, (int)sizeof(PT)), "slab kernel LDS");
    __shared__ float xs[W];
    __shared__ PT stage[kLsThreads * EPL];  // per wavefront: one tile's pair sums in pair order
    __shared__ PT head[kLsMaxTiles];
    __shared__ PT tail[kLsMaxTiles];
    __shared__ unsigned char tflag[kLsMaxTiles];  // bit 0: the tile has a row end; bit 1: starts mid-pair
    __shared__ int2 tds[kLsMaxTiles];             // the chunk's tile descriptors
    if (LsPair<K>::kAbs && blockIdx.x == 0 && threadIdx.x == 0) *chk_count = 0;
    const int4 ch = chunks[blockIdx.x];
    const int ntiles = ch.y;
    if (ntiles == 0) return;  // whole workgroup, before any barrier
    const int x0 = ch.z;
    const int width = min(W, C - x0);
    const int tid = threadIdx.x;

    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // a lane's EPL offsets and values of tile lt (one vector load each)
    auto load_tile = [&](int lt, uint16_t* c, VT* v) {
        const long e0 = ((long)ch.x + lt) * TILE + lane * EPL;
        LsLoad<uint16_t, EPL>::run(lc + e0, c);
        LsLoad<VT, EPL>::run(lv + e0, v);
    };
    uint16_t cn[EPL];
    VT vn[EPL];
    // the slab of x into LDS: a whole slab with 16-byte loads, all in flight together (then the first tile's
    // loads join them), the last, partial slab one element at a time
    const float* xg = x + x0;
    for (int i = tid; i < ntiles; i += kLsThreads) tds[i] = tiles[(long)ch.x + i];  // with the slab of x
    if (width == W && (((uintptr_t)xg) & 15) == 0) {
        constexpr int NV = W / (4 * kLsThreads);
        float4 v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = *(const float4*)(xg + (tid + k * kLsThreads) * 4);
        if (wave < ntiles) load_tile(wave, cn, vn);
#pragma unroll
        for (int k = 0; k < NV; ++k) *(float4*)(xs + (tid + k * kLsThreads) * 4) = v[k];
    } else {
        if (wave < ntiles) load_tile(wave, cn, vn);
        for (int i = tid; i < width; i += kLsThreads) xs[i] = xg[i];
    }
    __syncthreads();

    PT* st = stage + wave * TILE;
    // one tile: products from LDS x, the lane's segmented sums, a DPP scan across lanes, the tile's pair sums
    // through the LDS stage to P
    auto process = [&](int lt, const uint16_t* c, const VT* v) {
        const int2 td = tds[lt];  // lt is uniform: one LDS broadcast read (a scalar load here waited on memory)
        const bool mid = td.x < 0;
        const int pair0 = td.x & 0x7fffffff;
        PT p[EPL];
        unsigned f = 0;
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            p[j] = PT::prod(v[j], xs[c[j] & 0x7fff]);
            f |= (unsigned)(c[j] >> 15) << j;
        }
        // the sum left open after the lane's last row end (all of it if none)
        PT seg = PT::zero();
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            seg.add(p[j]);
            if ((f >> j) & 1u) seg = PT::zero();
        }
        int total;
        const int before = ls_count_prefix(__popc(f), total);
        PT sc = seg;
        int fl = f != 0u;
        ls_seg_scan<K>(sc, fl);
        PT run = ls_dpp<0x138, 0xf>(sc);  // open sum entering the lane (0 for lane 0)
        // the lane's k-th row end closes the tile's pair before + k (pair pair0 + before + k)
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            run.add(p[j]);
            if ((f >> j) & 1u) {
                st[before + __popc(f & ((1u << j) - 1u))] = run;
                run = PT::zero();
            }
        }
        // the tile's pairs to P with whole-wavefront stores of consecutive pairs; the tile's first pair, when it
        // started in an earlier tile, is finished after the barrier. Runs are taken eight at a time, read with
        // scalar loads (the run index is uniform), and a lane picks its pair's offset by comparisons. (One fixed
        // store per lane and round, with idle lanes writing a spare slot so that the wait counts for the next
        // tile's loads could leave the stores in flight, measured 3.5 us slower on the 2M power-law matrix.)
        int r = td.y, j = pair0 + (mid ? 1 : 0);
        const int jend = pair0 + total;
        while (j < jend) {
            const int2* rt = runs + ch.w + r;
            int bnd[8], off[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                bnd[i] = rt[i + 1].x;  // first pair after run r + i (each chunk's table ends in 9 sentinels)
                off[i] = rt[i].y;
            }
            const int je = min(jend, bnd[7]);
            for (int q = j + lane; q < je; q += 64) {
                int d = off[0];
#pragma unroll
                for (int i = 1; i < 8; ++i) d = q >= bnd[i - 1] ? off[i] : d;
                P[q + d] = st[q - pair0];
            }
            j = je;
            r += 8;
        }
        if (mid && lane == 0 && total > 0) head[lt] = st[0];
        const PT open = ls_readlane63<K>(sc);  // the sum open at the tile end (0 after a final row end)
        if (lane == 0) {
            tail[lt] = open;
            tflag[lt] = (unsigned char)((total > 0 ? 1 : 0) | (mid ? 2 : 0));
        }
    };
    // two register sets in turn, so that no in-flight load result is moved (a copy of the prefetched registers
    // made the compiler wait for the prefetch at the loop's back edge): the next tile's loads fly while a tile is
    // processed; past the chunk's last tile the loads repeat its last tile
    uint16_t cb[EPL];
    VT vb[EPL];
    for (int lt = wave; lt < ntiles; lt += 2 * NWAVE) {
        load_tile(min(lt + NWAVE, ntiles - 1), cb, vb);
        process(lt, cn, vn);
        if (lt + NWAVE >= ntiles) break;
        load_tile(min(lt + 2 * NWAVE, ntiles - 1), cn, vn);
        process(lt + NWAVE, cb, vb);
    }
    __syncthreads();
    // pairs that started in an earlier tile of the chunk: head + tails of the tiles before, newest first (a
    // chunk starts on a pair boundary, so the walk ends at its first tile at the latest)
    for (int lt = tid; lt < ntiles; lt += kLsThreads) {
        if (tflag[lt] != 3) continue;
        PT s = head[lt];
        for (int q = lt - 1; q >= 0; --q) {
            s.add(tail[q]);
            if ((tflag[q] & 1) || q == 0) break;
        }
        const int2 td = tds[lt];
        P[(td.x & 0x7fffffff) + runs[ch.w + td.y].y] = s;
    }
}

// ---------------------------------------------------------------- the combine kernel
// blocks[8b ..] = (first row, rows, first pair, pairs | entries << 16, wave split 1, 2, 3, first map dword);
// wave split w = (first entry | first row << 16) of wavefront w, at a row boundary (split 0 = (0, 0), split 4 =
// (entries, rows)). The block's pair sums are P[first .. first + pairs) (slab by slab, rows ascending inside a
// slab) and its 16-bit map entries list them in row-major order: LDS index (bits 0-14), bit 15 on the row's last
// entry; an empty row has one entry, the LDS slot kept at zero. A row's offset in the block is the number of
// row ends before its entries. Wavefront w sums the rows of its entries and writes their y.

// A wavefront sums its rows of one block from LDS (its entry range starts and ends on row boundaries): rounds
// of 64 * E entries start on a multiple of E, a lane takes E consecutive entries (one 16-byte LDS read), and
// entries outside [a, b) read the zero slot pb[ZERO] and end nothing. Each row's last entry writes the row's sum
// to yl (and, in a checked plan, its sum of |products| to al), the row being the wavefront's first plus the row
// ends before.
template <int K, int ZERO>
__device__ __forceinline__ void ls_rows(const int* bi, const LsPair<K>* __restrict__ pb,
                                        const uint16_t* __restrict__ lmap, float* __restrict__ yl,
                                        float* __restrict__ al, int wave, int lane) {
    using PT = LsPair<K>;
    constexpr int E = 8;
    const int ne = bi[3] >> 16;
    const int a = wave == 0 ? 0 : bi[3 + wave] & 0xffff, b = wave == 3 ? ne : bi[4 + wave] & 0xffff;
    int row = wave == 0 ? 0 : bi[3 + wave] >> 16;  // the row of the next row end
    PT carry = PT::zero();
    for (int i0 = a & ~(E - 1); i0 < b; i0 += 64 * E) {
        const int ib = i0 + lane * E;
        unsigned m[E];
        {
            const uint4 u = *(const uint4*)(lmap + ib);
            m[0] = u.x & 0xffffu; m[1] = u.x >> 16; m[2] = u.y & 0xffffu; m[3] = u.y >> 16;
            m[4] = u.z & 0xffffu; m[5] = u.z >> 16; m[6] = u.w & 0xffffu; m[7] = u.w >> 16;
        }
        PT p[E];
        unsigned f = 0;
        // an entry outside [a, b) reads the zero slot: every lane issues its E reads at once (a guarded read
        // became a branch that waited for the reads before it)
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const int i = ib + j;
            const bool in = i >= a && i < b;
            p[j] = pb[in ? (m[j] & 0x7fffu) : (unsigned)ZERO];
            f |= (in ? (m[j] >> 15) : 0u) << j;
        }
        PT seg = PT::zero();
#pragma unroll
        for (int j = 0; j < E; ++j) {
            seg.add(p[j]);
            if ((f >> j) & 1u) seg = PT::zero();
        }
        // the row open at the end of the previous round continues in lane 0 (carry is 0 in the first round, and
        // 0 after a round whose last entry ended a row)
        PT sc = seg;
        if (lane == 0 && f == 0u) {
            PT c = carry;
            c.add(sc);
            sc = c;
        }
        int fl = f != 0u;
        ls_seg_scan<K>(sc, fl);
        PT run = ls_dpp<0x138, 0xf>(sc);
        if (lane == 0) run = carry;
        int ends;
        int rr = row + ls_count_prefix(__popc(f), ends);  // the row of this lane's first row end
#pragma unroll
        for (int j = 0; j < E; ++j) {
            run.add(p[j]);
            if ((f >> j) & 1u) {
                yl[rr] = run.value();
                if constexpr (PT::kAbs) al[rr] = run.abs_sum();
                ++rr;
                run = PT::zero();
            }
        }
        row += ends;
        carry = ls_readlane63<K>(sc);
    }
}

__device__ __forceinline__ void ls_block_info(const int* __restrict__ blocks, int b, int* bi) {
#pragma unroll
    for (int k = 0; k < 8; ++k) bi[k] = blocks[8 * (long)b + k];  // uniform: scalar loads
}

// Copies n 16-byte units from global src to LDS dst with the wavefront's direct global->LDS loads, 16 bytes a
// lane (no register round trip; nothing waits here). The workgroup's four wavefronts take every fourth 64-unit
// piece.
__device__ __forceinline__ void ls_copy_lds16(const uint4* __restrict__ src, uint4* __restrict__ dst, int n,
                                              int wave, int lane) {
    for (int e0 = wave * 64; e0 < n; e0 += 256)
        if (e0 + lane < n)
            __builtin_amdgcn_global_load_lds((const void*)(src + e0 + lane),
                                             (__attribute__((address_space(3))) void*)(dst + e0), 16, 0, 0);
}

// ---- checked plans: a row's a-posteriori error bound, and its FP64 recomputation
// With u = 2^-24, n the row's nonzeros, S its sum of |products| (summed in FP32 beside y) and
// g = (n + 2) u / (1 - (n + 2) u):
//   plain FP32 sums in any order: |y - exact| <= g S (Higham, Accuracy and Stability, (3.5)), S itself within a
//     factor 1 / (1 - g) of the exact sum of |products|: bound = g S / (1 - g);
//   Dot2: every product and TwoSum error is below u S (n + 1 of them), their FP32 sum adds g times their total and
//     the final s + c one rounding: bound = 2 u |y| + 2 g^2 S / (1 - g).
// A row fails when bound > rtol |y| (a row whose y is 0 with S > 0 fails).
template <int K>
__device__ __forceinline__ bool ls_row_fails(int n, float yv, float S, float rtol) {
    if (n <= 0) return false;
    const double u = 0x1p-24, m = (double)(n + 2) * u;
    if (m >= 0.5) return true;
    const double g = m / (1.0 - m);
    const double a = __builtin_fabs((double)yv);
    const double bound = (K & 1) ? 2.0 * u * a + 2.0 * g * g * (double)S / (1.0 - g) : g * (double)S / (1.0 - g);
    return bound > (double)rtol * a;
}

__device__ __forceinline__ double ls_to_f64(float v) { return (double)v; }
__device__ __forceinline__ double ls_to_f64(double v) { return v; }
__device__ __forceinline__ double ls_to_f64(bf16_t v) { return (double)to_f32(v); }

// The rows a checked exec found failing, recomputed in FP64: workgroup g takes list entries g, g + grid, ... (the
// list is complete when this kernel starts: the combine kernel before it wrote every entry). One row per
// workgroup of 1024 lanes: lane t takes nonzeros k0 + t, k0 + t + 1024, ..., four at a time, in Dot2 in double
// (TwoProduct by FMA, TwoSum); the lanes' (s, c) pairs are combined by a butterfly of TwoSums in each wavefront
// and then wavefront by wavefront in index order, so a row's value does not depend on the list order.
__device__ __forceinline__ void ls_two_sum_add(double& s, double& c, double p, double ep) {
#pragma clang fp contract(off)
    const double t = s + p;
    const double bp = t - s;
    c = c + (((s - (t - bp)) + (p - bp)) + ep);
    s = t;
}

template <typename VT>
__global__ __launch_bounds__(1024) void lslab_fp64_rows_kernel(const int* __restrict__ chk,
                                                               const int* __restrict__ rp,
                                                               const int* __restrict__ ci,
                                                               const VT* __restrict__ v,
                                                               const float* __restrict__ x, float* __restrict__ y,
                                                               float alpha, float beta) {
#pragma clang fp contract(off)
    __shared__ double ws[16], wc[16];
    const int count = chk[0];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int j = blockIdx.x; j < count; j += gridDim.x) {
        const long row = chk[1 + j];
        const int k0 = rp[row], k1 = rp[row + 1];
        double s = 0.0, c = 0.0;
        int k = k0 + tid;
        for (; k + 3 * 1024 < k1; k += 4 * 1024) {
            double a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = ls_to_f64(v[k + q * 1024]);
                b[q] = (double)x[ci[k + q * 1024]];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double pr = a[q] * b[q];
                ls_two_sum_add(s, c, pr, __builtin_fma(a[q], b[q], -pr));
            }
        }
        for (; k < k1; k += 1024) {
            const double a = ls_to_f64(v[k]), b = (double)x[ci[k]];
            const double pr = a * b;
            ls_two_sum_add(s, c, pr, __builtin_fma(a, b, -pr));
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const double os = __shfl_xor(s, m), oc = __shfl_xor(c, m);
            ls_two_sum_add(s, c, os, oc);
        }
        if (lane == 0) {
            ws[wave] = s;
            wc[wave] = c;
        }
        __syncthreads();
        if (tid == 0) {
            double ts = ws[0], tc = wc[0];
            for (int w = 1; w < 16; ++w) ls_two_sum_add(ts, tc, ws[w], wc[w]);
            double r = (double)alpha * (ts + tc);
            if (beta != 0.f) r += (double)beta * (double)y[row];
            y[row] = (float)r;
        }
        __syncthreads();  // ws / wc are reused by the next row
    }
}

// One block per workgroup: the block's map entries and pair sums (two contiguous ranges) go to LDS by direct
// loads, all in flight before anything waits; the rows are then summed from LDS alone and each wavefront writes
// y for its own rows. Small blocks let many workgroups share a CU and hide each other's load latency.
// Checked plans (K bit 1): each wavefront tests its rows' error bounds; a row that fails is appended to
// chk[1 ..] (chk[0] counts them) and its y left for lslab_fp64_rows_kernel, which runs next (y keeps its old
// value for beta meanwhile).
template <typename VT, int K, int PAIR_BYTES, int ROWS>
__global__ __launch_bounds__(256) void lslab_combine_kernel(const int* __restrict__ blocks,
                                                            const uint16_t* __restrict__ map,
                                                            const LsPair<K>* __restrict__ P, float* __restrict__ y,
                                                            float alpha, float beta, const int* __restrict__ rp,
                                                            float rtol, int* __restrict__ chk) {
    using PT = LsPair<K>;
    constexpr bool CHK = PT::kAbs;
    constexpr int CAP = PAIR_BYTES / (int)sizeof(PT);
    constexpr int MC = CAP + ROWS + 8;  // map entries: pairs and empty rows (+ the tail of a lane's last read)
    __shared__ __attribute__((aligned(16))) PT pb[CAP + 4];  // CAP + 3: a zero, out of the copies' reach
    __shared__ __attribute__((aligned(16))) uint16_t lm[MC];
    __shared__ float yl[ROWS];
    __shared__ float al[CHK ? ROWS : 1];
    __shared__ unsigned char fl[CHK ? ROWS : 1];    // 1: the row failed its bound (y written by the FP64 pass)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int bi[8];
    ls_block_info(blocks, blockIdx.x, bi);
    const int base = bi[2], np = bi[3] & 0xffff, ne = bi[3] >> 16;
    ls_copy_lds16((const uint4*)map + bi[7], (uint4*)lm, (ne + 7) / 8, wave, lane);
    ls_copy_lds16((const uint4*)((const char*)P + (long)base * (long)sizeof(PT)), (uint4*)pb,
                  (np * (int)sizeof(PT) + 15) / 16, wave, lane);
    if (tid == 0) pb[CAP + 3] = PT::zero();  // the slot of empty rows
    const int r0 = wave == 0 ? 0 : bi[3 + wave] >> 16, r1 = wave == 3 ? bi[1] : bi[4 + wave] >> 16;
    for (int i = r0 + lane; i < r1; i += 64) {  // this wavefront's rows
        yl[i] = 0.f;
        if constexpr (CHK) {
            al[i] = 0.f;
            fl[i] = 0;
        }
    }
    __builtin_amdgcn_s_waitcnt(0);                      // this wavefront's direct loads have landed
    __syncthreads();
    ls_rows<K, CAP + 3>(bi, pb, lm, yl, al, wave, lane);
    if constexpr (CHK) {
        for (int i0 = r0; i0 < r1; i0 += 64) {
            const int i = i0 + lane;
            bool bad = false;
            if (i < r1) {
                const long row = (long)bi[0] + i;
                bad = ls_row_fails<K>(rp[row + 1] - rp[row], yl[i], al[i], rtol);
            }
            if (bad) {
                fl[i] = 1;
                const int slot = atomicAdd(chk, 1);  // at most one slot per row: the list holds numRows ids
                chk[1 + slot] = (int)((long)bi[0] + i);
            }
        }
    }
    for (int i = r0 + lane; i < r1; i += 64)
        if (!(CHK && fl[i])) store_y(y, (long)bi[0] + i, yl[i], alpha, beta);
}

// values into stream order: out[i] = v[src[i]], 0 where src[i] < 0 (tile padding)
template <typename VT>
__global__ void lslab_gather_values_kernel(const int* __restrict__ src, long n, const VT* __restrict__ v,
                                           VT* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = src[i];
    VT z;
    __builtin_memset(&z, 0, sizeof(VT));
    out[i] = k >= 0 ? v[k] : z;
}

// ---------------------------------------------------------------- plan construction (host)
spmv_status_t lslab_build(muxi_plan* p, int param0, spmv_type_t vt, hipStream_t s) {
    const int R = p->A.R, C = p->A.C, N = p->A.N;
    const int epl = param0 & 0xff;
    if (!(epl == 4 || epl == 8)) {
        set_error("lslab: nonzeros per lane must be 4 or 8");
        return SPMV_ERROR_INTERNAL;
    }
    if (p->check && p->comp && epl != 4) {
        set_error("lslab: a checked Dot2 plan takes 4 nonzeros per lane (its 12-byte pair sums fill the LDS)");
        return SPMV_ERROR_INTERNAL;
    }
    const size_t ps = sizeof(float) * (1 + (p->comp ? 1 : 0) + (p->check ? 1 : 0));  // bytes per pair sum
    const int W = lslab_slab_cols(param0, (int)ps);  // columns per slab
    const int pair_bytes = kLsPairBytes, rows_max = kLsRowsMax;
    const int S = C > 0 ? (int)(((long)C + W - 1) / W) : 0;
    if (S > kLsMaxSlabs) {
        set_error("lslab: at most " + std::to_string(kLsMaxSlabs) + " column slabs (" +
                  std::to_string((long)kLsMaxSlabs * W) + " columns)");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const int TILE = 64 * epl;
    std::vector<int> rp((size_t)R + 1), ci((size_t)N);
    MUXI_HIP(hipMemcpyAsync(rp.data(), p->A.rp, sizeof(int) * ((size_t)R + 1), hipMemcpyDeviceToHost, s));
    if (N > 0) MUXI_HIP(hipMemcpyAsync(ci.data(), p->A.ci, sizeof(int) * (size_t)N, hipMemcpyDeviceToHost, s));
    MUXI_HIP(hipStreamSynchronize(s));
    if (rp[0] != 0 || rp[R] != N) {
        set_error("lslab: rowPtr[0] must be 0 and rowPtr[numRows] = nnz");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    // ---- the slab streams: slab, then row, then position in the row
    std::vector<long> off((size_t)S + 1, 0);
    for (int r = 0; r < R; ++r) {
        if (rp[r + 1] < rp[r]) {
            set_error("lslab: rowPtr decreases");
            return SPMV_ERROR_INVALID_MATRIX;
        }
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int c = ci[k];
            if (c < 0 || c >= C) {
                set_error("lslab: column index out of range");
                return SPMV_ERROR_INVALID_MATRIX;
            }
            ++off[(size_t)(c / W) + 1];
        }
    }
    for (int q = 0; q < S; ++q) off[q + 1] += off[q];
    std::vector<int> src((size_t)N), srow((size_t)N);
    std::vector<uint16_t> scol((size_t)N);
    {
        std::vector<long> fill(off.begin(), off.end() - 1);
        for (int r = 0; r < R; ++r)
            for (int k = rp[r]; k < rp[r + 1]; ++k) {
                const int q = ci[k] / W;
                const long i = fill[q]++;
                src[i] = k;
                srow[i] = r;
                scol[i] = (uint16_t)(ci[k] - q * W);
            }
    }
    // row ends inside each slab; pair j (a (row, slab) with nonzeros) is closed by the j-th end in stream order
    std::vector<unsigned char> end((size_t)N);
    std::vector<int> ends_before((size_t)N + 1);
    long npairs = 0;
    for (int q = 0; q < S; ++q)
        for (long i = off[q]; i < off[q + 1]; ++i) {
            ends_before[i] = (int)npairs;
            end[i] = (i + 1 == off[q + 1] || srow[i + 1] != srow[i]) ? 1 : 0;
            npairs += end[i];
        }
    ends_before[N] = (int)npairs;
    if (npairs >= INT_MAX / 2) {
        set_error("lslab: too many (row, slab) pairs");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    std::vector<int> pair_row((size_t)npairs), pair_slab((size_t)npairs);
    for (int q = 0; q < S; ++q)
        for (long i = off[q]; i < off[q + 1]; ++i)
            if (end[i]) {
                pair_row[ends_before[i]] = srow[i];
                pair_slab[ends_before[i]] = q;
            }
    std::vector<int> pslab((size_t)S + 1, 0);  // first pair of each slab
    for (int q = 0; q < S; ++q) pslab[q + 1] = ends_before[off[q + 1]];

    // ---- chunks (slab-kernel workgroups): each slab's stream cut at row ends into pieces. All pieces run at once
    // when their count is the number of CUs (one 1024-lane workgroup fills a CU), so the pieces are apportioned
    // to the slabs by their nonzeros (largest remainders, at least one a slab) for exactly `cus` of them -- more,
    // in whole multiples of `cus`, when a piece would pass kLsMaxTiles tiles -- and cut at the first row end at or
    // past each even share. Workgroup b runs on XCD b mod 8: the pieces go to the XCDs in slab order, an equal
    // count each, so that every XCD gets the same work and reads few slabs of x into its L2.
    int dev = 0, cus = 256;
    MUXI_HIP(hipGetDevice(&dev));
    MUXI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cus = std::max(8, cus / 8 * 8);
    // a (row, slab) pair without repeated columns holds at most W nonzeros, so a piece cut at the first row end
    // past a share of at most max_piece stays within kLsMaxTiles tiles
    const long max_piece = (long)kLsMaxTiles * TILE - W;
    long nchunks = (long)cus * std::max(1L, ((long)N + (long)cus * max_piece - 1) / ((long)cus * max_piece));
    long nslabs_used = 0;
    for (int q = 0; q < S; ++q) nslabs_used += off[q + 1] > off[q] ? 1 : 0;
    nchunks = std::max(nchunks, nslabs_used);
    std::vector<long> kq((size_t)S, 0);
    {
        std::vector<std::pair<double, int>> rem;
        long used = 0;
        for (int q = 0; q < S; ++q) {
            const long nq = off[q + 1] - off[q];
            if (nq == 0) continue;
            const double xq = (double)nchunks * (double)nq / (double)std::max(N, 1);
            kq[q] = std::max(1L, (long)xq);
            used += kq[q];
            rem.push_back({xq - (double)(long)xq, q});
        }
        std::sort(rem.begin(), rem.end(), [](const std::pair<double, int>& u, const std::pair<double, int>& v) {
            return u.first > v.first || (u.first == v.first && u.second < v.second);
        });
        for (size_t i = 0; used < nchunks && i < rem.size(); ++i, ++used) ++kq[rem[i].second];
        // past the count (the one-a-slab floor): take pieces back from the slabs with the largest shares
        while (used > nchunks) {
            int best = -1;
            for (int q = 0; q < S; ++q)
                if (kq[q] > 1 && (best < 0 || (off[q + 1] - off[q]) * kq[best] < (off[best + 1] - off[best]) * kq[q]))
                    best = q;
            if (best < 0) break;
            --kq[best];
            --used;
        }
    }
    struct Piece {
        int slab;
        long a, n;  // stream positions [a, a + n)
    };
    std::vector<Piece> pieces;
    for (int q = 0; q < S; ++q) {
        const long e = off[q + 1], nq = e - off[q];
        if (nq == 0) continue;
        long a = off[q];
        for (long c = 1; a < e; ++c) {
            // the c-th cut: the first row end at or past c / kq of the slab (a share an earlier cut already passed
            // is skipped), within max_piece of the last cut
            long want = std::min(off[q] + (nq * c + kq[q] - 1) / kq[q], e);
            if (c < kq[q] && want <= a) continue;
            want = std::min(std::max(want, a + 1), a + max_piece);
            long i = want - 1;
            while (i + 1 < e && !end[i]) ++i;
            if (c >= kq[q] && e - a <= max_piece) i = e - 1;
            if (i + 1 - a > (long)kLsMaxTiles * TILE) {
                set_error("lslab: a row holds more than " + std::to_string(W) +
                          " nonzeros in one column slab (repeated columns)");
                return SPMV_ERROR_INVALID_MATRIX;
            }
            pieces.push_back({q, a, i + 1 - a});
            a = i + 1;
        }
    }
    // XCD of each piece: equal counts in slab order; a slab's XCD is that of its first piece
    const int np_ = (int)pieces.size();
    std::vector<int> pxcd((size_t)np_);
    std::vector<int> slab_xcd((size_t)S, 0);
    for (int i = np_ - 1; i >= 0; --i) {
        pxcd[(size_t)i] = (int)((long)i * 8 / std::max(np_, 1));
        slab_xcd[(size_t)pieces[(size_t)i].slab] = pxcd[(size_t)i];
    }

    // ---- combine blocks: rows cut so that a block holds at most kLsRowsMax rows and its pair sums (and map
    // entries) fit the LDS; a block's pairs take P[rstart[first row] ..) in (slab, row) order
    const int cap = pair_bytes / (int)ps;
    std::vector<int> row_pairs((size_t)R + 1, 0);
    for (long j = 0; j < npairs; ++j) ++row_pairs[pair_row[j]];
    std::vector<int> rstart((size_t)R + 1, 0);
    for (int r = 0; r < R; ++r) rstart[r + 1] = rstart[r] + row_pairs[r];
    std::vector<int> rlist((size_t)std::max(npairs, 1L));
    {
        std::vector<int> fill(rstart.begin(), rstart.end() - 1);
        for (long j = 0; j < npairs; ++j) rlist[fill[pair_row[j]]++] = (int)j;  // slab order within a row
    }
    std::vector<int> brow;  // first row of each block, then R
    std::vector<int> row_block((size_t)std::max(R, 1));
    for (int r = 0; r < R;) {
        brow.push_back(r);
        int np = 0, nr = 0;
        while (r < R && nr < rows_max && (nr == 0 || np + row_pairs[r] <= cap)) {  // a row has <= S <= cap
            np += row_pairs[r];
            row_block[r] = (int)brow.size() - 1;
            ++nr;
            ++r;
        }
    }
    const int B = (int)brow.size();
    brow.push_back(R);
    // F[b][s]: the first pair of slab s with row >= brow[b]. Block b's pairs take P[rstart[brow[b]] ..) slab by
    // slab, so pair j of slab s goes to j + delta[b][s], delta = rstart[brow[b]] + (block pairs in slabs < s) -
    // F[b][s]; its map entry holds its LDS index (the same offset inside the block)
    std::vector<int> F((size_t)(B + 1) * std::max(S, 1), 0);
    for (int q = 0; q < S; ++q) {
        int j = pslab[q];
        for (int b = 0; b <= B; ++b) {
            while (j < pslab[q + 1] && pair_row[j] < brow[b]) ++j;
            F[(size_t)b * S + q] = j;
        }
    }
    // Inside a block the slabs' runs are ordered by the XCD their chunks run on (then by slab), so each XCD writes
    // one contiguous part of every block and few cache lines of P are written from two L2s.
    std::vector<int> xorder((size_t)S);
    for (int q = 0; q < S; ++q) xorder[q] = q;
    std::stable_sort(xorder.begin(), xorder.end(), [&](int u, int v) { return slab_xcd[u] < slab_xcd[v]; });
    std::vector<int> delta((size_t)std::max(B, 1) * std::max(S, 1), 0);
    // a block's pair sums start on a 16-byte boundary of P (the combine copies them to LDS 16 bytes a lane)
    const long palign = ps == 8 ? 2 : 4;
    long pnext = 0;
    std::vector<uint16_t> hmap;  // per block, from a dword boundary: its entries in row-major order
    hmap.reserve((size_t)npairs + (size_t)R + 2 * (size_t)B + 8);
    std::vector<int> hblk((size_t)8 * std::max(B, 1), 0);
    {
        std::vector<int> L((size_t)S + 1);  // L[q]: offset of slab q's run in the block
        for (int b = 0; b < B; ++b) {
            const int r0 = brow[b], r1 = brow[b + 1];
            const int np = rstart[r1] - rstart[r0];
            const int m0 = (int)pnext;  // the block's first pair sum in P
            pnext = (pnext + np + palign - 1) / palign * palign;
            int acc = 0;
            for (int i = 0; i < S; ++i) {
                const int q = xorder[i];
                L[q] = acc;
                acc += F[(size_t)(b + 1) * S + q] - F[(size_t)b * S + q];
                delta[(size_t)b * S + q] = m0 + L[q] - F[(size_t)b * S + q];
            }
            const size_t mb = hmap.size();  // the block's first entry (a multiple of 8: a 16-byte boundary)
            std::vector<int> rent((size_t)(r1 - r0) + 1);  // entry index of each row's first entry in the block
            for (int r = r0; r < r1; ++r) {
                rent[(size_t)(r - r0)] = (int)(hmap.size() - mb);
                if (rstart[r + 1] == rstart[r]) {  // an empty row: one entry, the zero slot, ending the row
                    hmap.push_back((uint16_t)((cap + 3) | 0x8000));
                    continue;
                }
                for (int k = rstart[r]; k < rstart[r + 1]; ++k) {
                    const int j = rlist[k], q = pair_slab[j];
                    const unsigned idx = (unsigned)(L[q] + j - F[(size_t)b * S + q]);
                    hmap.push_back((uint16_t)(idx | (k + 1 == rstart[r + 1] ? 0x8000u : 0u)));
                }
            }
            const int ne = (int)(hmap.size() - mb);
            rent[(size_t)(r1 - r0)] = ne;
            while (hmap.size() & 7) hmap.push_back(0);  // the next block's map on a 16-byte boundary
            int* d = &hblk[(size_t)8 * b];
            d[0] = r0;
            d[1] = r1 - r0;
            d[2] = m0;
            d[3] = np | (ne << 16);
            // wave splits at row boundaries, near w * ne / 4: (first entry | first row << 16) of wavefront w
            int r = r0;
            for (int w = 1; w < 4; ++w) {
                const long want = (long)ne * w / 4;
                while (r < r1 && rent[(size_t)(r - r0)] < want) ++r;
                d[3 + w] = rent[(size_t)(r - r0)] | ((r - r0) << 16);
            }
            d[7] = (int)(mb / 8);  // the block's map, in 16-byte units
        }
    }

    // launch order: workgroup b runs on XCD b mod 8, so XCD x's i-th piece takes slot 8 i + x
    std::vector<std::vector<int>> per_xcd(8);
    for (int i = 0; i < np_; ++i) per_xcd[(size_t)pxcd[(size_t)i]].push_back(i);
    size_t rounds = 0;
    for (auto& v : per_xcd) rounds = std::max(rounds, v.size());
    std::vector<int> order;
    for (size_t r = 0; r < rounds; ++r)
        for (int xc = 0; xc < 8; ++xc) order.push_back(r < per_xcd[xc].size() ? per_xcd[xc][r] : -1);
    while (!order.empty() && order.back() < 0) order.pop_back();
    // padded tile layout and run tables in piece order
    std::vector<long> tile0(pieces.size());
    std::vector<int> run0(pieces.size());
    std::vector<int> hrun;  // (first pair, offset) per run, then (INT_MAX, 0) per piece
    long ntiles = 0;
    for (size_t i = 0; i < pieces.size(); ++i) {
        const Piece& pc = pieces[i];
        tile0[i] = ntiles;
        ntiles += (pc.n + TILE - 1) / TILE;
        run0[i] = (int)(hrun.size() / 2);
        const int jA = ends_before[pc.a], jB = ends_before[pc.a + pc.n];
        int prev = -1;
        for (int j = jA; j < jB; ++j) {
            const int b = row_block[pair_row[j]];
            if (b != prev) {
                hrun.push_back(j);
                hrun.push_back(delta[(size_t)b * S + pc.slab]);
                prev = b;
            }
        }
        for (int k = 0; k < 9; ++k) {  // a tile reads up to 8 entries past its run: they stay in its chunk's table
            hrun.push_back(INT_MAX);
            hrun.push_back(0);
        }
    }
    if (hrun.empty()) hrun.assign(2, 0);
    const long npad = ntiles * TILE;
    if (ntiles >= INT_MAX / 2 || npad >= (long)INT_MAX) {
        set_error("lslab: padded nonzeros exceed INT32_MAX");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    std::vector<uint16_t> hlc((size_t)std::max(npad, 1L), 0);
    std::vector<int> hsrc((size_t)std::max(npad, 1L), -1), htile((size_t)2 * std::max(ntiles, 1L), 0);
    for (size_t i = 0; i < pieces.size(); ++i) {
        const Piece& pc = pieces[i];
        const long base = tile0[i] * TILE;
        for (long e = 0; e < pc.n; ++e) {
            const long si = pc.a + e;
            hlc[base + e] = (uint16_t)(scol[si] | (end[si] << 15));
            hsrc[base + e] = src[si];
        }
        const long nt = (pc.n + TILE - 1) / TILE;
        int r = run0[i];
        for (long lt = 0; lt < nt; ++lt) {
            const long si = pc.a + lt * TILE;
            const bool mid = lt > 0 && !end[si - 1];
            const int pj = ends_before[si];
            while (hrun[2 * (r + 1)] <= pj) ++r;  // the sentinel stops the walk
            htile[2 * (tile0[i] + lt)] = pj | (mid ? (int)0x80000000u : 0);
            htile[2 * (tile0[i] + lt) + 1] = r - run0[i];
        }
    }
    std::vector<int> hch(4 * std::max<size_t>(order.size(), 1), 0);
    for (size_t b = 0; b < order.size(); ++b) {
        if (order[b] < 0) continue;
        const size_t i = (size_t)order[b];
        const Piece& pc = pieces[i];
        hch[4 * b + 0] = (int)tile0[i];
        hch[4 * b + 1] = (int)((pc.n + TILE - 1) / TILE);
        hch[4 * b + 2] = pc.slab * W;
        hch[4 * b + 3] = run0[i];
    }
    // ---- device arrays
    p->ls_S = S;
    p->ls_chunks = (int)order.size();
    p->ls_blocks = B;
    p->ls_pairs = npairs;
    const size_t vs = vsize(vt);
    if (p->check) MUXI_HIP(hipMalloc(&p->d_ls_chk, sizeof(int) * ((size_t)R + 1)));
    MUXI_HIP(hipMalloc(&p->d_ls_lc, sizeof(uint16_t) * hlc.size()));
    MUXI_HIP(hipMalloc(&p->d_ls_lv, vs * hlc.size()));
    MUXI_HIP(hipMalloc(&p->d_ls_tp, sizeof(int) * htile.size()));
    MUXI_HIP(hipMalloc(&p->d_ls_ch, sizeof(int) * hch.size()));
    MUXI_HIP(hipMalloc(&p->d_ls_blk, sizeof(int) * hblk.size()));
    MUXI_HIP(hipMalloc(&p->d_ls_F, sizeof(int) * hrun.size()));
    if (hmap.empty()) hmap.assign(8, 0);
    p->ls_pspan = pnext;
    MUXI_HIP(hipMalloc(&p->d_ls_map, sizeof(uint16_t) * hmap.size()));
    MUXI_HIP(hipMalloc(&p->d_ls_P, ps * (size_t)pnext + 16));  // (+16: the last 16-byte copy may run over)
    MUXI_HIP(hipMemcpyAsync(p->d_ls_lc, hlc.data(), sizeof(uint16_t) * hlc.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_ls_tp, htile.data(), sizeof(int) * htile.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_ls_ch, hch.data(), sizeof(int) * hch.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_ls_blk, hblk.data(), sizeof(int) * hblk.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_ls_F, hrun.data(), sizeof(int) * hrun.size(), hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(p->d_ls_map, hmap.data(), sizeof(uint16_t) * hmap.size(), hipMemcpyHostToDevice, s));
    int* d_src = nullptr;
    MUXI_HIP(hipMalloc(&d_src, sizeof(int) * hsrc.size()));
    hipError_t e = hipMemcpyAsync(d_src, hsrc.data(), sizeof(int) * hsrc.size(), hipMemcpyHostToDevice, s);
    spmv_status_t st = SPMV_SUCCESS;
    if (e != hipSuccess) {
        set_error(std::string("lslab: ") + hipGetErrorString(e));
        st = SPMV_ERROR_DEVICE;
    } else {
        const long n = (long)hsrc.size();
        const unsigned g = grid_for(n, 256);
        if (vt == SPMV_TYPE_FLOAT)
            hipLaunchKernelGGL(lslab_gather_values_kernel<float>, dim3(g), dim3(256), 0, s, d_src, n,
                               (const float*)p->A.v, (float*)p->d_ls_lv);
        else if (vt == SPMV_TYPE_BF16)
            hipLaunchKernelGGL(lslab_gather_values_kernel<bf16_t>, dim3(g), dim3(256), 0, s, d_src, n,
                               (const bf16_t*)p->A.v, (bf16_t*)p->d_ls_lv);
        else
            hipLaunchKernelGGL(lslab_gather_values_kernel<double>, dim3(g), dim3(256), 0, s, d_src, n,
                               (const double*)p->A.v, (double*)p->d_ls_lv);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host arrays above go out of scope
        if (e != hipSuccess) {
            set_error(std::string("lslab: ") + hipGetErrorString(e));
            st = SPMV_ERROR_DEVICE;
        }
    }
    (void)hipFree(d_src);
    p->scratch_bytes = (sizeof(uint16_t) + vs) * hlc.size() +
                       sizeof(int) * (htile.size() + hch.size() + hblk.size() + hrun.size()) +
                       sizeof(uint16_t) * hmap.size() + ps * (size_t)pnext + 16 +
                       (p->check ? sizeof(int) * ((size_t)R + 1) : 0);
    return st;
}

template <typename VT, int K, int EPL, int W>
static void launch_slab(const muxi_plan* p, const float* x, hipStream_t s) {
    hipLaunchKernelGGL((lslab_slab_kernel<VT, K, EPL, W>), dim3(p->ls_chunks), dim3(kLsThreads), 0, s,
                       (const int4*)p->d_ls_ch, p->d_ls_lc, (const VT*)p->d_ls_lv, (const int2*)p->d_ls_tp,
                       (const int2*)p->d_ls_F, x, p->A.C, (LsPair<K>*)p->d_ls_P, p->d_ls_chk);
}

template <typename VT, int K, int EPL>
static void launch_slab_w(const muxi_plan* p, const float* x, hipStream_t s) {
    constexpr int PB = (int)sizeof(LsPair<K>);
    switch (lslab_slab_cols(p->param0, PB)) {
        case kLsSlabCols:
            if constexpr (lslab_fits(kLsSlabCols, EPL, PB)) launch_slab<VT, K, EPL, kLsSlabCols>(p, x, s);
            break;
        case 28672:
            if constexpr (lslab_fits(28672, EPL, PB)) launch_slab<VT, K, EPL, 28672>(p, x, s);
            break;
        case 24576:
            if constexpr (lslab_fits(24576, EPL, PB)) launch_slab<VT, K, EPL, 24576>(p, x, s);
            break;
        default:
            launch_slab<VT, K, EPL, kLsSlabCols / 2>(p, x, s);
    }
}

__global__ void lslab_clear_count_kernel(int* c) { *c = 0; }

template <typename VT, int K>
static spmv_status_t launch_lslab_k(const muxi_plan* p, const float* x, float* y, float alpha, float beta,
                                    hipStream_t s, hipEvent_t ev_end) {
    if (p->ls_chunks > 0) {
        // the slab width the plan was built with (lslab_slab_cols)
        if ((p->param0 & 0xff) == 4)
            launch_slab_w<VT, K, 4>(p, x, s);
        else if constexpr (K != 3)  // (a checked Dot2 plan is built with 4)
            launch_slab_w<VT, K, 8>(p, x, s);
        MUXI_LAUNCH_CHECK();
    } else if (K & 2) {
        hipLaunchKernelGGL(lslab_clear_count_kernel, dim3(1), dim3(1), 0, s, p->d_ls_chk);  // no slab kernel ran
        MUXI_LAUNCH_CHECK();
    }
    if (ev_end) MUXI_HIP(hipEventRecord(ev_end, s));
    if (p->ls_blocks > 0) {
        hipLaunchKernelGGL((lslab_combine_kernel<VT, K, kLsPairBytes, kLsRowsMax>), dim3(p->ls_blocks), dim3(256),
                           0, s, p->d_ls_blk, p->d_ls_map, (const LsPair<K>*)p->d_ls_P, y, alpha, beta, p->A.rp,
                           p->rtol, p->d_ls_chk);
        MUXI_LAUNCH_CHECK();
    }
    if constexpr ((K & 2) != 0) {
        // a fixed grid: the count is on the device; workgroups past it exit at once
        const unsigned g = (unsigned)std::max(1, std::min(p->A.R, 512));
        hipLaunchKernelGGL(lslab_fp64_rows_kernel<VT>, dim3(g), dim3(1024), 0, s, p->d_ls_chk, p->A.rp, p->A.ci,
                           (const VT*)p->A.v, x, y, alpha, beta);
        MUXI_LAUNCH_CHECK();
    }
    return SPMV_SUCCESS;
}

template <typename VT, bool COMP>
spmv_status_t launch_lslab(const muxi_plan* p, const float* x, float* y, float alpha, float beta, hipStream_t s,
                           hipEvent_t ev_end) {
    if (p->A.R <= 0) {
        if (ev_end) MUXI_HIP(hipEventRecord(ev_end, s));
        return SPMV_SUCCESS;
    }
    constexpr int K0 = COMP ? 1 : 0;
    return p->check ? launch_lslab_k<VT, K0 | 2>(p, x, y, alpha, beta, s, ev_end)
                    : launch_lslab_k<VT, K0>(p, x, y, alpha, beta, s, ev_end);
}

template spmv_status_t launch_lslab<float, false>(const muxi_plan*, const float*, float*, float, float, hipStream_t,
                                                  hipEvent_t);
template spmv_status_t launch_lslab<float, true>(const muxi_plan*, const float*, float*, float, float, hipStream_t,
                                                 hipEvent_t);
template spmv_status_t launch_lslab<bf16_t, false>(const muxi_plan*, const float*, float*, float, float, hipStream_t,
                                                   hipEvent_t);
template spmv_status_t launch_lslab<bf16_t, true>(const muxi_plan*, const float*, float*, float, float, hipStream_t,
                                                  hipEvent_t);
template spmv_status_t launch_lslab<double, false>(const muxi_plan*, const float*, float*, float, float, hipStream_t,
                                                   hipEvent_t);
template spmv_status_t launch_lslab<double, true>(const muxi_plan*, const float*, float*, float, float, hipStream_t,
                                                  hipEvent_t);

}  // namespace muxi
