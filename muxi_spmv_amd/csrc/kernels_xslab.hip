// Column-slab SpMV ("xslab"): keeps each XCD's x gathers in its own L2.
//
// MI355X has 8 XCDs, each with a private 4 MiB L2, and hands workgroup i to
// XCD i mod 8. When x is larger than one L2 (the 2M-column power-law matrix
// has an 8 MB x) and the columns are spread, every XCD pulls all of x through
// its L2: on that matrix the plain merge kernel moves ~1.1 GB through L2 for
// 139 MB of matrix and vectors (TCC_MISS x 128 B). The plan splits the
// columns into S slabs (column ranges holding equal shares of the nonzeros)
// and stores, per slab, a compressed CSR of the nonzeros that fall in it:
// slab-local row pointers over the rows that have any (no slab row is empty), columns
// and values, padded to whole tiles, plus one row-end bit per nonzero and a
// descriptor per tile (the slab row of its first nonzero). Workgroup b works
// on slab b mod S, so an XCD gathers from one slice of x that stays in its
// L2; the matrix stream uses non-temporal loads so it does not evict that
// slice.
//
// A (row, slab) pair with nonzeros is a "pair"; pairs are numbered row-major,
// so a row's pairs are contiguous. The slab kernel stores a row's partial
// sum to the row's slot in its slab (slab-local order; plain stores, no
// atomics); the partial of the row open at a tile end goes to a carry keyed
// by its slot, and a keyed fix-up adds carries in tile order. A combine
// kernel then forms y[r] = alpha * (sum of the row's pairs, in slab order) +
// beta * y[r]. Every sum has a fixed order: results are reproducible.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"

namespace muxi {

// N consecutive elements with one non-temporal vector load (the address is
// aligned to N elements: slab regions start on tile boundaries).
template <typename T, int N>
struct VecLoad {
    __device__ __forceinline__ static void run(const T* p, T* out) {
        typedef T V __attribute__((ext_vector_type(N)));
        const V v = __builtin_nontemporal_load((const V*)p);
#pragma unroll
        for (int j = 0; j < N; ++j) out[j] = v[j];
    }
};
template <int N>
struct VecLoad<bf16_t, N> {
    __device__ __forceinline__ static void run(const bf16_t* p, bf16_t* out) {
        typedef uint16_t V __attribute__((ext_vector_type(N)));
        const V v = __builtin_nontemporal_load((const V*)p);
#pragma unroll
        for (int j = 0; j < N; ++j) out[j].bits = v[j];
    }
};

// Cross-lane moves with DPP (GFX9 encodings: row_shr:n = 0x110 | n within a
// row of 16 lanes, row_bcast:15 = 0x142, row_bcast:31 = 0x143, wave_shr:1 =
// 0x138); lanes without a source read 0. These run in the ALUs, no LDS trip.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_v(float v) {
    return __int_as_float(dpp_i<CTRL, ROWS>(__float_as_int(v)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_v(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL, ROWS>((int)(b & 0xffffffffll)), hi = dpp_i<CTRL, ROWS>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One step of the segmented scan: (f1, v1) then (f2, v2) -> (f1 | f2, f2 ? v2 : v1 + v2).
template <int CTRL, int ROWS, typename AT>
__device__ __forceinline__ void seg_step(AT& v, int& f) {
    const AT vu = dpp_v<CTRL, ROWS>(v);
    const int fu = dpp_i<CTRL, ROWS>(f);
    if (!f) v = vu + v;
    f |= fu;
}

// Inclusive segmented scan over the 64 lanes: shifts of 1, 2, 4, 8 inside
// each row of 16, then row 0's total into rows 1 and 3 (row_bcast:15) and
// rows 0-1's total into rows 2 and 3 (row_bcast:31).
template <typename AT>
__device__ __forceinline__ void wave_seg_scan(AT& v, int& f) {
    seg_step<0x111, 0xf>(v, f);
    seg_step<0x112, 0xf>(v, f);
    seg_step<0x114, 0xf>(v, f);
    seg_step<0x118, 0xf>(v, f);
    seg_step<0x142, 0xa>(v, f);
    seg_step<0x143, 0xc>(v, f);
}

// Column-slab SpMV with row-end flags. A slab's compressed rows are never
// empty, so its nonzeros form one segmented sequence: bit e of the slab's
// flag words is set when nonzero e ends its row, and a tile descriptor holds
// the slab row of the tile's first nonzero. One wavefront per tile of 64*EPL
// nonzeros, lane l owning EPL consecutive ones (one vector load of columns
// and one of values, then all EPL gathers in flight together). A lane sums
// its products, cutting at its flags; one segmented wave scan (DPP) carries
// the open sums between lanes; the lane's k-th row end is slab row
// first + (flags before the lane) + k, whose partial goes to
// P[row_base[s] + that row]. No LDS, no barrier: wavefronts run
// independently, and a kernel this light keeps eight of them per SIMD (on
// MI355X this beat a persistent, software-pipelined walk over the tiles,
// whose second register set halves that occupancy). Workgroup b works on
// slab b mod S (one XCD's x slice in its L2); its four wavefronts take four
// consecutive tiles. The row open at a tile end goes to the tile's carry,
// keyed by its P slot.
template <typename VT, typename XT, typename AT, int EPL>
__global__ __launch_bounds__(256) void xslab_kernel(SlabArgs sa, const int* __restrict__ sci,
                                                    const VT* __restrict__ svals,
                                                    const uint32_t* __restrict__ sflags,
                                                    const int* __restrict__ tile_row, const XT* __restrict__ x,
                                                    AT* __restrict__ P, AT* __restrict__ carry,
                                                    int* __restrict__ carry_key, AT* __restrict__ PC) {
    constexpr int SIG = 64 * EPL, NW = SIG / 32;  // nonzeros and flag words per tile
    const int s = blockIdx.x % sa.S;
    const int t = __builtin_amdgcn_readfirstlane((int)(blockIdx.x / sa.S) * 4 + (int)(threadIdx.x >> 6));
    if (t >= sa.tiles[s]) return;  // whole wavefront
    const int lane = threadIdx.x & 63;
    const long e0 = sa.nnz_base[s] + (long)t * SIG + lane * EPL;  // padded layout: the whole tile is allocated
    int c[EPL];
    VT v[EPL];
    VecLoad<int, EPL>::run(sci + e0, c);
    VecLoad<VT, EPL>::run(svals + e0, v);
    const uint32_t* fw = sflags + (sa.nnz_base[s] + (long)t * SIG) / 32;
    uint32_t w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = fw[i];  // the tile's flags: uniform, scalar loads
    const int desc = tile_row[sa.tile_base[s] + t];
    const int row0 = desc & 0x7fffffff;  // bit 31: the open row's only carry goes to its PC slot
    XT xv[EPL];
#pragma unroll
    for (int j = 0; j < EPL; ++j) xv[j] = x[c[j]];

    // this lane's flag bits and the number of row ends before it in the tile
    constexpr int LPW = 32 / EPL;  // lanes per flag word
    const int wi = lane / LPW, sh = (lane % LPW) * EPL;
    uint32_t wsel = 0;
    int rk = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int pc = __popc(w[i]);
        rk += i < wi ? pc : 0;
        tot += pc;
        wsel = i == wi ? w[i] : wsel;
    }
    rk += __popc(wsel & ((1u << sh) - 1u));
    const uint32_t f = (wsel >> sh) & ((1u << EPL) - 1u);
    const int nreal = sa.nnz[s] - t * SIG - lane * EPL;  // real nonzeros of this lane (padding past the end)
    AT p[EPL];
#pragma unroll
    for (int j = 0; j < EPL; ++j) p[j] = j < nreal ? Conv<AT>::f(v[j]) * Conv<AT>::f(xv[j]) : AT(0);
    // the sum left open after the lane's last row end (all of it if none)
    AT seg = AT(0);
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
        seg += p[j];
        if ((f >> j) & 1u) seg = AT(0);
    }
    // segmented inclusive scan over lanes: a lane with a row end restarts the sum
    AT sc = seg;
    int fl = f != 0u;
    wave_seg_scan(sc, fl);
    AT* Ps = P + sa.row_base[s] + row0 + rk;
    seg = dpp_v<0x138, 0xf>(sc);  // open sum entering the lane: the lane before's inclusive sum (0 for lane 0)
    int k = 0;
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
        seg += p[j];
        if ((f >> j) & 1u) {
            Ps[k++] = seg;
            seg = AT(0);
        }
    }
    if (lane == 63) {
        // the open sum at the tile end belongs to slab row row0 + tot, unless the tile's last nonzero ended a row
        // or every row is closed (the slab's last tile)
        const int orow = row0 + tot;
        const bool open = orow < sa.rows[s] && !((f >> (EPL - 1)) & 1u);
        carry[sa.tile_base[s] + t] = open ? sc : AT(0);
        carry_key[sa.tile_base[s] + t] = open ? sa.row_base[s] + orow : -1;
        if (desc < 0 && open) PC[sa.row_base[s] + orow] = sc;
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

// Row masks of the fix-up-free combine: bits 0-7 the slabs holding a pair of
// the row, bit 8 + s set when the row's pair in slab s has a carry in PC,
// bit 31 set for a heavy row. Slot entries of heavy rows: the P slot, or
// -(j + 2) for a pair crossing several tile ends, -1 after the last pair.
constexpr unsigned XSLAB_HEAVY = 0x80000000u;

// Combine for plans of at most eight slabs, with no fix-up launch before it.
// The P slot of row r's pair in slab s is wbase[w * S + s] (the slot of the
// first pair in slab s at or after row 64w, w = r / 64) plus the number of
// rows of r's wavefront before r that hold a pair in slab s, which a ballot
// of the mask bit gives: the combine reads one 4-byte mask per row, and the
// pairs of consecutive rows in one slab are consecutive slots. A pair whose
// row crosses exactly one tile end has that carry in PC[slot] (the slab kernel
// writes it there; PC is zero for every other slot), so it is P[slot] +
// PC[slot]; only pairs with the mask bit load PC. A row with a pair that
// crosses several tile ends is "heavy": its K slot entries are in hent,
// the crossing pair is -(j + 2) with ct[3j..3j+2] = (slot, t0, t1), and one
// of the first workgroups' wavefronts sums carry[t0 .. t1) 64 at a time.
// Pairs are added in slab order, as in every other combine.
template <typename XT, typename AT, int K, int J>
__global__ __launch_bounds__(256) void xslab_combine_mask_kernel(int R, int S, const unsigned* __restrict__ rmask,
                                                                 const int* __restrict__ wbase,
                                                                 const AT* __restrict__ P, const AT* __restrict__ PC,
                                                                 const int* __restrict__ ct,
                                                                 const AT* __restrict__ carry, int H,
                                                                 const int* __restrict__ hrow,
                                                                 const int* __restrict__ hent, XT* __restrict__ y,
                                                                 AT alpha, AT beta) {
    const int HB = (H + 3) / 4;
    if ((int)blockIdx.x < HB) {  // heavy rows: one wavefront each
        const int h = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
        if (h >= H) return;  // whole wavefront
        const int r = hrow[h];
        AT s = AT(0);
        for (int k = 0; k < K; ++k) {
            const int q = hent[(long)K * h + k];
            if (q == -1) break;  // no more pairs
            if (q >= 0) {
                s += P[q] + PC[q];  // PC is 0 where no carry lands
            } else {
                const int j = -q - 2, t0 = ct[3 * j + 1], t1 = ct[3 * j + 2];
                AT c = AT(0);
                for (int t = t0 + lane; t < t1; t += 64) c += carry[t];
#pragma unroll
                for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
                s += c + P[ct[3 * j]];
            }
        }
        if (lane == 0) store_y(y, r, s, alpha, beta);
        return;
    }
    // Whole wavefronts stay to the end (the ballots); rows past R have no bits. A wavefront takes J groups of
    // 64 consecutive rows, so that J masks, then J * K partial loads, are in flight together.
    const int lane = threadIdx.x & 63;
    const long g0 = ((long)(blockIdx.x - HB) * (blockDim.x >> 6) + (threadIdx.x >> 6)) * J;  // first row group
    const unsigned long long lt = __lanemask_lt();
    unsigned m[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const long r = (g0 + j) * 64 + lane;
        m[j] = r < R ? __builtin_nontemporal_load(rmask + r) : 0u;
    }
    int slot[J][K];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const long w = (g0 + j) * 64 < R ? g0 + j : 0;  // a group wholly past R reads group 0's bases
#pragma unroll
        for (int q = 0; q < K; ++q) slot[j][q] = q < S ? wbase[w * S + q] : 0;
    }
    bool use[J][K];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const bool mine = !(m[j] & XSLAB_HEAVY);
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const bool has = q < S && ((m[j] >> q) & 1u);
            slot[j][q] += __popcll(__ballot(has) & lt);
            use[j][q] = has && mine;
        }
    }
    AT v[J][K], c[J][K];
#pragma unroll
    for (int j = 0; j < J; ++j) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            v[j][q] = use[j][q] ? ntl(P + slot[j][q]) : AT(0);
            c[j][q] = use[j][q] && ((m[j] >> (8 + q)) & 1u) ? PC[slot[j][q]] : AT(0);
        }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const long r = (g0 + j) * 64 + lane;
        AT s = AT(0);
#pragma unroll
        for (int q = 0; q < K; ++q)
            if (use[j][q]) s += v[j][q] + c[j][q];
        if (r < R && !(m[j] & XSLAB_HEAVY)) store_y(y, r, s, alpha, beta);
    }
}

// The pair open at the end of each tile of one slab (-1 = none): tile t ends
// inside slab row r = tile_row[t + 1] when that row starts before the next
// tile does; the slab's last tile ends its last row.
__global__ void xslab_carry_keys_kernel(const int* __restrict__ srp, const int* __restrict__ tile_row, int tiles,
                                        int SIG, int row_base, int* __restrict__ keys) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles) return;
    int k = -1;
    if (t + 1 < tiles) {
        const int r = tile_row[t + 1];
        if (srp[r] < (t + 1) * SIG) k = row_base + r;
    }
    keys[t] = k;
}

// Sets a[idx[i]] = 1.
__global__ void xslab_set_kernel(const int* __restrict__ idx, int n, int* __restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[idx[i]] = 1;
}

// Row r's pairs with a carry slot (has_pc[slot]) set bit 8 + s of its mask,
// s the slab of the pair (the k-th set bit of the mask for entry k).
__global__ void xslab_flag_pc_kernel(int R, int K, const int* __restrict__ ploc4, const int* __restrict__ has_pc,
                                     unsigned* __restrict__ rmask) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    unsigned m = rmask[r], rest = m & 0xffu, add = 0;
    for (int k = 0; k < K && rest; ++k) {
        const int sl = __builtin_ctz(rest);
        rest &= rest - 1;
        const int q = ploc4[K * r + k];
        if (q >= 0 && has_pc[q]) add |= 1u << (8 + sl);
    }
    if (add) rmask[r] = m | add;
}

// wbase[w * S + s]: the P slot of the first pair in slab s at or after row
// 64w, from the slab-major exclusive scan lrow of the pair flags.
__global__ void xslab_wbase_kernel(int R, int S, const int* __restrict__ lrow, int* __restrict__ wbase) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nw = ((long)R + 63) / 64;
    if (i >= nw * S) return;
    const long w = i / S, q = i % S;
    wbase[i] = lrow[q * R + 64 * w];
}

// Marks the descriptors of tiles whose carry is their open row's only one.
__global__ void xslab_mark_single_kernel(const int* __restrict__ tiles, int n, int* __restrict__ tile_row) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tile_row[tiles[i]] |= (int)0x80000000u;
}

__global__ void xslab_cross_id_kernel(const int* __restrict__ ct, long nct, int* __restrict__ cross_id) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nct) cross_id[ct[3 * j]] = (int)j;
}

// ploc4 entries of pairs crossing several tile ends become -(j + 2); their rows get hflag = 1.
__global__ void xslab_mark_heavy_kernel(int R, int K, int* __restrict__ ploc4, const int* __restrict__ cross_id,
                                        int* __restrict__ hflag) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    int heavy = 0;
    for (int k = 0; k < K; ++k) {
        const int q = ploc4[K * r + k];
        if (q < 0) continue;
        const int j = cross_id[q];
        if (j >= 0) {
            ploc4[K * r + k] = -(j + 2);
            heavy = 1;
        }
    }
    hflag[r] = heavy;
}

// Heavy row h = hrank[r]: hrow[h] = r, its K slot entries go to hent[Kh ..],
// and its mask gets XSLAB_HEAVY.
__global__ void xslab_heavy_rows_kernel(int R, int K, const int* __restrict__ hflag, const int* __restrict__ hrank,
                                        const int* __restrict__ ploc4, int* __restrict__ hrow, int* __restrict__ hent,
                                        unsigned* __restrict__ rmask) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R || !hflag[r]) return;
    const int h = hrank[r];
    hrow[h] = (int)r;
    for (int k = 0; k < K; ++k) hent[(long)K * h + k] = ploc4[K * r + k];
    rmask[r] |= XSLAB_HEAVY;
}

// ---- plan construction
// Slab q holds columns [cut[q], cut[q+1]). The cuts split the nonzeros into S
// near-equal parts (column quantiles), so every slab, and so every XCD, gets
// the same work whatever the column distribution (a row shard of a banded
// matrix touches only a narrow column window, for instance).
struct SlabCuts {
    int c[kMaxSlabs + 1];
};

__device__ __forceinline__ int slab_of(int col, const SlabCuts& cut, int S) {
    int q = 0;
    for (int k = 1; k < S; ++k) q += col >= cut.c[k];
    return q;
}

__global__ void xslab_colhist_kernel(const int* __restrict__ ci, long N, int C, int* __restrict__ hist) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const int c = ci[k];
        if ((unsigned)c < (unsigned)C) atomicAdd(hist + c, 1);
    }
}

// cut[q] = the first column c with (nonzeros in columns < c) >= ceil(q N / S); pre = exclusive prefix of the
// column histogram, C + 1 entries.
__global__ void xslab_cuts_kernel(const int* __restrict__ pre, int C, long N, int S, int* __restrict__ cut) {
    const int q = threadIdx.x;
    if (q > S) return;
    if (q == 0 || q == S) {
        cut[q] = q == 0 ? 0 : C;
        return;
    }
    const long t = ((long)q * N + S - 1) / S;
    int lo = 0, hi = C;
    while (lo < hi) {
        const int mid = (int)(((long)lo + hi) >> 1);
        if (pre[mid] < t) lo = mid + 1;
        else hi = mid;
    }
    cut[q] = lo;
}

// One wavefront per row (rows reach 10^5 nonzeros): the row is read in
// 64-wide chunks and lane q < S counts slab q's nonzeros from ballots.
__global__ __launch_bounds__(256) void xslab_count_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                          int R, SlabCuts cut, int S, int* __restrict__ cnt,
                                                          int* __restrict__ flag, int* __restrict__ npair) {
    const long r = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= R) return;  // whole wavefront
    const int b = rp[r], e = rp[r + 1];
    int c = 0;
    for (int i = b; i < e; i += 64) {
        const int s = i + lane < e ? slab_of(ci[i + lane], cut, S) : -1;
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
                                                         const VT* __restrict__ v, int R, SlabCuts cut, int S,
                                                         const int* __restrict__ cnt, const int* __restrict__ pos,
                                                         const int* __restrict__ lrow,
                                                         const int* __restrict__ nnz_base,
                                                         const int* __restrict__ pad_base,
                                                         const int* __restrict__ rpp, int* __restrict__ srp,
                                                         int* __restrict__ ploc, int* __restrict__ ploc4, int K,
                                                         unsigned* __restrict__ rmask, int* __restrict__ sci,
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
        if (rmask && lane == 0) rmask[r] = (unsigned)hm;  // S <= 8 here: bits 0-7
        if (has) {
            const int lr = lrow[k];
            const int rank = __popcll(hm & __lanemask_lt());
            ploc[rpp[r] + rank] = lr;  // the row's pairs in slab order
            if (ploc4) ploc4[r * K + rank] = lr;
            srp[lr + lane] = off - nnz_base[lane];
        }
    }
    const int b = rp[r], e = rp[r + 1];
    for (int i = b; i < e; i += 64) {
        const bool in = i + lane < e;
        const int col = in ? ci[i + lane] : 0;
        const int s = in ? slab_of(col, cut, S) : -1;
        int o = 0;
        for (int q = 0; q < S; ++q) {
            const unsigned long long m = __ballot(s == q);
            const int base = __shfl(off, q);
            if (s == q) o = base + __popcll(m & __lanemask_lt());
            if (lane == q) off += __popcll(m);
        }
        if (in) {
            const int op = o - nnz_base[s] + pad_base[s];  // into the slab's tile-aligned region
            sci[op] = col;
            sv[op] = v[i + lane];
        }
    }
}

// Tile padding of slab q: column lo (a column of the slab, so the gather stays in its x slice), value 0.
template <typename VT>
__global__ void xslab_pad_kernel(int* __restrict__ sci, VT* __restrict__ sv, long from, long to, int lo) {
    const long i = from + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= to) return;
    sci[i] = lo;
    VT z;
    __builtin_memset(&z, 0, sizeof(VT));
    sv[i] = z;
}

// Row-end flags of one slab: bit (pad_base + rowPtr[q+1] - 1) for each of its rows q.
__global__ void xslab_flags_kernel(const int* __restrict__ srp, int rows, long pad_base, uint32_t* __restrict__ flags) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= rows) return;
    const long e = pad_base + srp[q + 1] - 1;
    atomicOr(flags + e / 32, 1u << (e % 32));
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
using XslabFn = void (*)(SlabArgs, const int*, const VT*, const uint32_t*, const int*, const XT*, AT*, AT*, int*,
                         AT*);

template <typename VT, typename XT, typename AT>
static XslabFn<VT, XT, AT> xslab_fn(int epl) {
    return epl == 16 ? xslab_kernel<VT, XT, AT, 16> : epl == 8 ? xslab_kernel<VT, XT, AT, 8> : xslab_kernel<VT, XT, AT, 4>;
}

// Plan-build temporaries on one stream: freed (stream-ordered) when the build returns, on every path, or earlier
// by release().
struct StreamTemps {
    hipStream_t s;
    std::vector<void*> held;
    explicit StreamTemps(hipStream_t st) : s(st) {}
    StreamTemps(const StreamTemps&) = delete;
    StreamTemps& operator=(const StreamTemps&) = delete;
    ~StreamTemps() {
        for (void* q : held) (void)hipFreeAsync(q, s);
    }
    template <typename T>
    hipError_t alloc(T** out, size_t bytes) {
        *out = nullptr;
        void* q = nullptr;
        const hipError_t e = hipMallocAsync(&q, bytes, s);
        if (e == hipSuccess) {
            held.push_back(q);
            *out = (T*)q;
        }
        return e;
    }
    void release(void* q) {
        for (size_t i = 0; i < held.size(); ++i)
            if (held[i] == q) {
                (void)hipFreeAsync(q, s);
                held.erase(held.begin() + (long)i);
                return;
            }
    }
};

spmv_status_t xslab_build(muxi_plan* p, int S, int epl, spmv_type_t vt, hipStream_t s) {
    const int R = p->A.R, C = p->A.C;
    StreamTemps tmp(s);
    if (S < 1 || S > kMaxSlabs) {
        set_error("xslab: slabs must be in [1, 16]");
        return SPMV_ERROR_INTERNAL;
    }
    if ((long)S * R + 1 > INT_MAX) {
        set_error("xslab: rows * slabs exceeds INT32_MAX");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const int SIG = 64 * epl;
    const long SR = (long)S * R;
    SlabCuts cut{};
    {
        int *hist, *pre, *d_cut;
        MUXI_HIP(tmp.alloc(&hist, sizeof(int) * ((size_t)C + 1)));
        MUXI_HIP(tmp.alloc(&pre, sizeof(int) * ((size_t)C + 1)));
        MUXI_HIP(tmp.alloc(&d_cut, sizeof(int) * (kMaxSlabs + 1)));
        MUXI_HIP(hipMemsetAsync(hist, 0, sizeof(int) * ((size_t)C + 1), s));
        if (p->A.N > 0) {
            unsigned g = grid_for(p->A.N, 256);
            hipLaunchKernelGGL(xslab_colhist_kernel, dim3(g > 65536 ? 65536 : g), dim3(256), 0, s, p->A.ci,
                               (long)p->A.N, C, hist);
            MUXI_LAUNCH_CHECK();
        }
        spmv_status_t st = exscan(hist, pre, (long)C + 1, s);
        if (st != SPMV_SUCCESS) return st;
        hipLaunchKernelGGL(xslab_cuts_kernel, dim3(1), dim3(64), 0, s, pre, C, (long)p->A.N, S, d_cut);
        MUXI_LAUNCH_CHECK();
        MUXI_HIP(hipMemcpyAsync(cut.c, d_cut, sizeof(int) * (S + 1), hipMemcpyDeviceToHost, s));
        MUXI_HIP(hipStreamSynchronize(s));
        for (int* t : {hist, pre, d_cut}) tmp.release(t);
    }
    int *cnt, *flag, *pos, *lrow, *npair;
    MUXI_HIP(tmp.alloc(&cnt, sizeof(int) * (SR + 1)));
    MUXI_HIP(tmp.alloc(&flag, sizeof(int) * (SR + 1)));
    MUXI_HIP(tmp.alloc(&pos, sizeof(int) * (SR + 1)));
    MUXI_HIP(tmp.alloc(&lrow, sizeof(int) * (SR + 1)));
    MUXI_HIP(tmp.alloc(&npair, sizeof(int) * ((size_t)R + 1)));
    MUXI_HIP(hipMemsetAsync(cnt + SR, 0, sizeof(int), s));
    MUXI_HIP(hipMemsetAsync(flag + SR, 0, sizeof(int), s));
    MUXI_HIP(hipMemsetAsync(npair + R, 0, sizeof(int), s));
    hipLaunchKernelGGL(xslab_count_kernel, dim3(grid_for((long)R * 64, 256)), dim3(256), 0, s, p->A.rp, p->A.ci, R,
                       cut, S, cnt, flag, npair);
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
    // slab q's nonzeros start on a tile boundary of the padded arrays
    SlabArgs& sa = p->slab;
    sa.S = S;
    std::vector<int> hpad(S + 1);
    long tp = 0, ntiles = 0;
    int maxt = 0;
    for (int q = 0; q < S; ++q) {
        const int nq = hpos[q + 1] - hpos[q];
        sa.rows[q] = hrow[q + 1] - hrow[q];
        sa.row_base[q] = hrow[q];
        sa.nnz[q] = nq;
        sa.tiles[q] = (int)(((long)nq + SIG - 1) / SIG);
        sa.tile_base[q] = (int)ntiles;
        hpad[q] = (int)tp;
        sa.nnz_base[q] = (int)tp;
        tp += (long)sa.tiles[q] * SIG;
        ntiles += sa.tiles[q];
        maxt = sa.tiles[q] > maxt ? sa.tiles[q] : maxt;
    }
    hpad[S] = (int)tp;
    if (tp >= INT_MAX) {
        set_error("xslab: padded nonzeros exceed INT32_MAX");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    p->T = (int)ntiles;
    p->P = maxt;
    const long ncarry = ntiles;  // one carry per tile
    int* d_nb = nullptr;
    MUXI_HIP(tmp.alloc(&d_nb, sizeof(int) * 2 * S));
    MUXI_HIP(hipMemcpyAsync(d_nb, hpos.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
    MUXI_HIP(hipMemcpyAsync(d_nb + S, hpad.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
    const size_t vs = vsize(vt);
    const size_t nflag = (size_t)tp / 32;
    // rowPtr (pairs + S), pair slots (pairs), then columns, values and flags at 256-byte boundaries (the kernel
    // reads a lane's columns and values with one vector load each)
    const size_t head = ((size_t)total_rows * 2 + S + 63) / 64 * 64;
    const size_t ints = head + (size_t)tp;
    MUXI_HIP(hipMalloc(&p->d_slab, sizeof(int) * ints + vs * (size_t)(tp > 0 ? tp : 1) + 4 * (nflag + 1) + 512));
    p->d_slab_rp = p->d_slab;
    p->d_slab_ploc = p->d_slab_rp + total_rows + S;
    p->d_slab_ci = p->d_slab + head;
    uintptr_t va = (uintptr_t)(p->d_slab_ci + tp);
    va = (va + 255) & ~(uintptr_t)255;
    p->d_slab_v = (void*)va;
    uintptr_t fa = va + vs * (size_t)(tp > 0 ? tp : 1);
    fa = (fa + 255) & ~(uintptr_t)255;
    p->d_slab_flags = (uint32_t*)fa;
    MUXI_HIP(hipMemsetAsync(p->d_slab_flags, 0, 4 * (nflag + 1), s));
    std::vector<int> term(S);
    for (int q = 0; q < S; ++q) term[q] = hpos[q + 1] - hpos[q];
    for (int q = 0; q < S; ++q)
        MUXI_HIP(hipMemcpyAsync(p->d_slab_rp + sa.row_base[q] + q + sa.rows[q], &term[q], sizeof(int),
                                hipMemcpyHostToDevice, s));
    // At most eight slabs: the combine works from row masks and per-wavefront slot bases, with no fix-up
    // launch. ploc4 (a row's pair slots in one K-int entry, -1 = none) only serves the plan build.
    p->slab_k = S <= 4 ? 4 : (S <= 8 ? 8 : 0);
    int* ploc4 = nullptr;
    if (p->slab_k) {
        MUXI_HIP(tmp.alloc(&ploc4, sizeof(int) * p->slab_k * (size_t)R));
        MUXI_HIP(hipMemsetAsync(ploc4, 0xff, sizeof(int) * p->slab_k * (size_t)R, s));
        MUXI_HIP(hipMalloc(&p->d_slab_rmask, sizeof(unsigned) * (size_t)R));
        const long nwb = ((long)R + 63) / 64 * S;
        MUXI_HIP(hipMalloc(&p->d_slab_wbase, sizeof(int) * (size_t)nwb));
        hipLaunchKernelGGL(xslab_wbase_kernel, dim3(grid_for(nwb, 256)), dim3(256), 0, s, R, S, lrow,
                           p->d_slab_wbase);
        MUXI_LAUNCH_CHECK();
    }
    st = with_types(vt, p->xtype, p->comp, [&](auto tag) -> spmv_status_t {
        using T = decltype(tag);
        using VT = typename T::VT;
        hipLaunchKernelGGL(xslab_fill_kernel<VT>, dim3(grid_for((long)R * 64, 256)), dim3(256), 0, s, p->A.rp, p->A.ci,
                           (const VT*)p->A.v, R, cut, S, cnt, pos, lrow, d_nb, d_nb + S, p->d_slab_rpp, p->d_slab_rp,
                           p->d_slab_ploc, ploc4, p->slab_k, p->d_slab_rmask, p->d_slab_ci, (VT*)p->d_slab_v);
        MUXI_LAUNCH_CHECK();
        for (int q = 0; q < S; ++q) {
            const long from = (long)hpad[q] + sa.nnz[q], to = (long)hpad[q] + (long)sa.tiles[q] * SIG;
            const int lo = cut.c[q];  // first column of slab q (it has nonzeros, so lo < C)
            if (to > from) {
                hipLaunchKernelGGL(xslab_pad_kernel<VT>, dim3(grid_for(to - from, 256)), dim3(256), 0, s,
                                   p->d_slab_ci, (VT*)p->d_slab_v, from, to, lo);
                MUXI_LAUNCH_CHECK();
            }
        }
        return SPMV_SUCCESS;
    });
    if (st != SPMV_SUCCESS) return st;
    const size_t acc = p->xtype == SPMV_TYPE_DOUBLE ? 8 : 4;
    MUXI_HIP(hipMalloc(&p->d_tile_row, 2 * sizeof(int) * (size_t)(ntiles + 1)));  // descriptors, then offsets
    p->d_tile_off = p->d_tile_row + ntiles + 1;
    // carries, then P; one spare element so that P[0] exists even with no pairs (the combines' clamped loads)
    MUXI_HIP(hipMalloc(&p->d_carry, acc * (size_t)(ncarry + total_rows + 2)));
    MUXI_HIP(hipMalloc(&p->d_slab_ckey, sizeof(int) * (size_t)(ncarry + 1)));
    for (int q = 0; q < S; ++q) {
        if (sa.tiles[q] == 0) continue;
        const int* srp_q = p->d_slab_rp + sa.row_base[q] + q;
        hipLaunchKernelGGL(xslab_flags_kernel, dim3(grid_for(sa.rows[q], 256)), dim3(256), 0, s, srp_q, sa.rows[q],
                           (long)hpad[q], p->d_slab_flags);
        MUXI_LAUNCH_CHECK();
        st = launch_csr5_tiles(srp_q, sa.rows[q], sa.nnz[q], SIG, sa.tiles[q], p->d_tile_row + sa.tile_base[q],
                               p->d_tile_off + sa.tile_base[q], s);
        if (st != SPMV_SUCCESS) return st;
    }
    if (ploc4 && ncarry > 0) {
        // Carries without a fix-up launch: a tile whose open row crosses only this tile end writes the carry to
        // the row's PC slot (descriptor bit 31); a pair crossing several tile ends gets a row (slot, t0, t1) of
        // ct and its row is summed by a wavefront in the combine.
        MUXI_HIP(hipMalloc(&p->d_slab_pc, acc * ((size_t)total_rows + 1)));
        MUXI_HIP(hipMemsetAsync(p->d_slab_pc, 0, acc * ((size_t)total_rows + 1), s));
        int* keys = nullptr;
        MUXI_HIP(tmp.alloc(&keys, sizeof(int) * (size_t)(ncarry + 1)));
        for (int q = 0; q < S; ++q) {
            if (sa.tiles[q] == 0) continue;
            hipLaunchKernelGGL(xslab_carry_keys_kernel, dim3(grid_for(sa.tiles[q], 256)), dim3(256), 0, s,
                               p->d_slab_rp + sa.row_base[q] + q, p->d_tile_row + sa.tile_base[q], sa.tiles[q], SIG,
                               sa.row_base[q], keys + sa.tile_base[q]);
            MUXI_LAUNCH_CHECK();
        }
        std::vector<int> hk((size_t)ncarry);
        MUXI_HIP(hipMemcpyAsync(hk.data(), keys, sizeof(int) * ncarry, hipMemcpyDeviceToHost, s));
        MUXI_HIP(hipStreamSynchronize(s));
        tmp.release(keys);
        std::vector<int> single, ct;
        for (long t = 0; t < ncarry;) {
            if (hk[t] < 0) {
                ++t;
                continue;
            }
            long u = t;
            while (u + 1 < ncarry && hk[u + 1] == hk[t]) ++u;
            if (u == t) {
                single.push_back((int)t);
            } else {
                ct.push_back(hk[t]);
                ct.push_back((int)t);
                ct.push_back((int)(u + 1));
            }
            t = u + 1;
        }
        if (!single.empty()) {
            int* d_single = nullptr;
            MUXI_HIP(tmp.alloc(&d_single, sizeof(int) * single.size()));
            MUXI_HIP(hipMemcpyAsync(d_single, single.data(), sizeof(int) * single.size(), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(xslab_mark_single_kernel, dim3(grid_for((long)single.size(), 256)), dim3(256), 0, s,
                               d_single, (int)single.size(), p->d_tile_row);
            MUXI_LAUNCH_CHECK();
            MUXI_HIP(hipStreamSynchronize(s));
            tmp.release(d_single);
        }
        const long nct = (long)ct.size() / 3;
        p->nheavy = 0;
        if (nct > 0) {
            MUXI_HIP(hipMalloc(&p->d_slab_ct, sizeof(int) * ct.size()));
            MUXI_HIP(hipMemcpyAsync(p->d_slab_ct, ct.data(), sizeof(int) * ct.size(), hipMemcpyHostToDevice, s));
            int *cross_id = nullptr, *hflag = nullptr, *hrank = nullptr;
            MUXI_HIP(tmp.alloc(&cross_id, sizeof(int) * ((size_t)total_rows + 1)));
            MUXI_HIP(tmp.alloc(&hflag, sizeof(int) * ((size_t)R + 1)));
            MUXI_HIP(tmp.alloc(&hrank, sizeof(int) * ((size_t)R + 1)));
            MUXI_HIP(hipMemsetAsync(cross_id, 0xff, sizeof(int) * ((size_t)total_rows + 1), s));
            MUXI_HIP(hipMemsetAsync(hflag + R, 0, sizeof(int), s));
            hipLaunchKernelGGL(xslab_cross_id_kernel, dim3(grid_for(nct, 256)), dim3(256), 0, s, p->d_slab_ct, nct,
                               cross_id);
            MUXI_LAUNCH_CHECK();
            hipLaunchKernelGGL(xslab_mark_heavy_kernel, dim3(grid_for(R, 256)), dim3(256), 0, s, R, p->slab_k,
                               ploc4, cross_id, hflag);
            MUXI_LAUNCH_CHECK();
            st = exscan(hflag, hrank, (long)R + 1, s);
            if (st != SPMV_SUCCESS) return st;
            int H = 0;
            MUXI_HIP(hipMemcpyAsync(&H, hrank + R, sizeof(int), hipMemcpyDeviceToHost, s));
            MUXI_HIP(hipStreamSynchronize(s));
            if (H > 0) {
                MUXI_HIP(hipMalloc(&p->d_slab_heavy, sizeof(int) * (1 + p->slab_k) * (size_t)H));
                hipLaunchKernelGGL(xslab_heavy_rows_kernel, dim3(grid_for(R, 256)), dim3(256), 0, s, R, p->slab_k,
                                   hflag, hrank, ploc4, p->d_slab_heavy, p->d_slab_heavy + H, p->d_slab_rmask);
                MUXI_LAUNCH_CHECK();
            }
            p->nheavy = H;
            for (int* t : {cross_id, hflag, hrank}) tmp.release(t);
        }
        if (!single.empty()) {  // the masks say which pairs have a carry slot (the others skip PC)
            std::vector<int> skey(single.size());
            for (size_t i = 0; i < single.size(); ++i) skey[i] = hk[(size_t)single[i]];
            int *d_sk = nullptr, *has_pc = nullptr;
            MUXI_HIP(tmp.alloc(&d_sk, sizeof(int) * skey.size()));
            MUXI_HIP(tmp.alloc(&has_pc, sizeof(int) * ((size_t)total_rows + 1)));
            MUXI_HIP(hipMemsetAsync(has_pc, 0, sizeof(int) * ((size_t)total_rows + 1), s));
            MUXI_HIP(hipMemcpyAsync(d_sk, skey.data(), sizeof(int) * skey.size(), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(xslab_set_kernel, dim3(grid_for((long)skey.size(), 256)), dim3(256), 0, s, d_sk,
                               (int)skey.size(), has_pc);
            MUXI_LAUNCH_CHECK();
            hipLaunchKernelGGL(xslab_flag_pc_kernel, dim3(grid_for(R, 256)), dim3(256), 0, s, R, p->slab_k, ploc4,
                               has_pc, p->d_slab_rmask);
            MUXI_LAUNCH_CHECK();
            MUXI_HIP(hipStreamSynchronize(s));  // skey is host memory the copy reads
            for (int* t : {d_sk, has_pc}) tmp.release(t);
        }
    }
    MUXI_HIP(hipStreamSynchronize(s));
    p->scratch_bytes = sizeof(int) * ints + vs * (size_t)tp + 4 * nflag + 2 * sizeof(int) * (size_t)(ntiles + 1) +
                       acc * (size_t)(ncarry + total_rows) + sizeof(int) * ((size_t)ncarry + R + 1) +
                       (p->slab_k ? sizeof(unsigned) * (size_t)R + sizeof(int) * (((size_t)R + 63) / 64 * S) +
                                        sizeof(int) * (1 + p->slab_k) * (size_t)p->nheavy
                                  : 0);
    return SPMV_SUCCESS;
}

template <typename VT, typename XT, typename AT>
spmv_status_t launch_xslab(const muxi_plan* p, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s,
                           hipEvent_t ev_end) {
    if (p->A.R <= 0) return SPMV_SUCCESS;
    const SlabArgs& sa = p->slab;
    const int ncarry = p->T;
    AT* carry = (AT*)p->d_carry;
    AT* P = carry + ncarry + 1;
    if (p->P > 0) {
        const dim3 grid((unsigned)(((long)p->P + 3) / 4 * sa.S));  // four tiles per workgroup
        const XslabFn<VT, XT, AT> fn = xslab_fn<VT, XT, AT>(p->param0 & 0xff);
        hipLaunchKernelGGL(fn, grid, dim3(256), 0, s, sa, p->d_slab_ci, (const VT*)p->d_slab_v, p->d_slab_flags,
                           p->d_tile_row, x, P, carry, p->d_slab_ckey, (AT*)p->d_slab_pc);
        MUXI_LAUNCH_CHECK();
        if (ev_end) MUXI_HIP(hipEventRecord(ev_end, s));
        if (!p->slab_k) {
            hipLaunchKernelGGL(keyed_fixup_kernel<AT>, dim3(grid_for(ncarry, 256)), dim3(256), 0, s, p->d_slab_ckey,
                               ncarry, carry, P);
            MUXI_LAUNCH_CHECK();
        }
    } else if (ev_end) {
        MUXI_HIP(hipEventRecord(ev_end, s));
    }
    if (p->slab_k) {
        const int H = p->nheavy;
        // with no tiles there are no pairs: PC is never read then (every mask is 0)
        const AT* PC = p->d_slab_pc ? (const AT*)p->d_slab_pc : P;
        // J row groups of 64 per wavefront: 4 with four slot bases, 2 with eight
        if (p->slab_k == 4)
            hipLaunchKernelGGL((xslab_combine_mask_kernel<XT, AT, 4, 4>),
                               dim3((H + 3) / 4 + grid_for(p->A.R, 256 * 4)), dim3(256), 0, s, p->A.R, sa.S,
                               p->d_slab_rmask, p->d_slab_wbase, P, PC, p->d_slab_ct, carry, H, p->d_slab_heavy,
                               p->d_slab_heavy + H, y, alpha, beta);
        else
            hipLaunchKernelGGL((xslab_combine_mask_kernel<XT, AT, 8, 2>),
                               dim3((H + 3) / 4 + grid_for(p->A.R, 256 * 2)), dim3(256), 0, s, p->A.R, sa.S,
                               p->d_slab_rmask, p->d_slab_wbase, P, PC, p->d_slab_ct, carry, H, p->d_slab_heavy,
                               p->d_slab_heavy + H, y, alpha, beta);
    } else {
        hipLaunchKernelGGL((xslab_combine_kernel<XT, AT>), dim3(grid_for(p->A.R, 256)), dim3(256), 0, s, p->A.R,
                           p->d_slab_rpp, p->d_slab_ploc, P, y, alpha, beta);
    }
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
