// Plan object: one CSR matrix bound to one kernel with its tile parameters
// and every scratch buffer that kernel needs.
#pragma once

#include "kernels.hpp"

namespace muxi {
constexpr int kMaxSlabs = 16;
// LDS column slabs (kernels_lslab.hip)
constexpr int kLsSlabCols = 32768;   // columns per slab: 128 KiB of FP32 x in LDS, 15-bit offsets
constexpr int kLsThreads = 1024;     // slab-kernel workgroup
constexpr int kLsMaxTiles = 512;     // tiles per slab-kernel chunk (LDS head/tail sums)
constexpr int kLsLdsBytes = 163840;  // LDS of one CU (MI355X)
constexpr int kLsMaxSlabs = 512;     // slabs (16,777,216 columns of 32768)
constexpr long kLsSelectorMaxCols = 256L * kLsSlabCols;  // AUTO considers lslab up to 8,388,608 columns (its
                                                         // training domain; explicit plans go to kLsMaxSlabs)
constexpr int kLsPairBytes = 8192;   // combine: LDS for a row block's pair sums (and as much for its map)
constexpr int kLsRowsMax = 1024;     // combine: rows per block
constexpr int kLsMaxRuns = 1024;     // combine blocks one slab-kernel chunk may write to
constexpr int kLsHalfSlabs = 0x100;  // param0 bit: slabs of kLsSlabCols / 2 columns
// Columns per slab: the widest of 32768, 28672, 24576 and 16384 whose FP32 x fits one CU's LDS beside the slab
// kernel's staged pair sums (1024 lanes x EPL pair sums of pair_bytes), its per-tile head and tail sums and flags
// and the chunk's tile descriptors (8 bytes a tile); half slabs (16384) when asked. Full slabs take four nonzeros
// a lane and 4-byte pair sums.
constexpr bool lslab_fits(int W, int epl, int pair_bytes) {
    return W * 4 + kLsThreads * epl * pair_bytes + kLsMaxTiles * (2 * pair_bytes + 1 + 8) <= kLsLdsBytes;
}
inline int lslab_slab_cols(int param0, int pair_bytes) {
    const int epl = param0 & 0xff;
    if (param0 & kLsHalfSlabs) return kLsSlabCols / 2;
    for (int W : {kLsSlabCols, 28672, 24576})
        if (lslab_fits(W, epl, pair_bytes)) return W;
    return kLsSlabCols / 2;
}
struct SlabArgs {
    int S;
    int rows[kMaxSlabs];       // compressed rows of each slab (none is empty)
    int tiles[kMaxSlabs];      // tiles of 64*EPL nonzeros of each slab
    int row_base[kMaxSlabs];   // first row id of the slab; its rowPtr starts at row_base + s
    int nnz_base[kMaxSlabs];   // first element of the slab in the padded (tile-aligned) arrays
    int nnz[kMaxSlabs];        // nonzeros of the slab
    int tile_base[kMaxSlabs];  // first global tile (carry slot, tile descriptor) of the slab
};
}  // namespace muxi

struct muxi_plan {
    muxi::DevCsr A;
    spmv_type_t vtype, xtype;
    bool comp;
    muxi_kernel_t kernel;
    int param0;
    int device;
    // merge-path
    int P = 0;
    int* d_pos = nullptr;  // P+1 diagonals, then P+1 row coordinates
    int* d_row = nullptr;
    // CSR5
    int T = 0;
    int* d_tile_row = nullptr;  // T, then T offsets
    int* d_tile_off = nullptr;
    // merge / CSR5 carries
    void* d_carry = nullptr;
    // ELL
    int width = 0;
    void* d_ecol = nullptr;  // int columns, or short offsets col - (row + row_base) when ecol16
    bool ecol16 = false;
    int row_base = 0;        // a row block's first row in the full matrix (plan_create_rows), else 0
    int ell_win = -1;        // ELL: every column within ell_win of its row (x window in LDS), -1 = no window
    void* d_eval = nullptr;
    // adaptive row blocks: blk_row[nblocks+1], blk_chunk[nblocks], long_row[nlong], long_first[2*nlong]
    int nblocks = 0;
    int nlong = 0;
    int* d_blocks = nullptr;
    int* d_blk_chunk = nullptr;
    int* d_long_row = nullptr;
    int* d_long_first = nullptr;
    size_t scratch_bytes = 0;
    // SpMM: right-hand sides (0 for a SpMV plan)
    int nrhs = 0;
    // column slabs (xslab): one allocation holding rowPtr, pair slots, columns, values, row-end flags
    muxi::SlabArgs slab{};
    int* d_slab = nullptr;
    int* d_slab_rp = nullptr;
    int* d_slab_ploc = nullptr;  // P slot of each pair, pairs in row-major order
    unsigned* d_slab_rmask = nullptr; // with at most 8 slabs: per row, its slabs, carry-slot and heavy bits
    int* d_slab_wbase = nullptr;      // per wavefront of 64 rows and slab: the P slot of its first pair there
    int slab_k = 0;                   // K: 4 (up to 4 slabs), 8 (up to 8), 0 (more: row offsets and a fix-up)
    void* d_slab_pc = nullptr;   // with row masks: the single carry of each pair crossing one tile end (else 0)
    int* d_slab_ct = nullptr;    // pairs crossing several tile ends: (P slot, first carry tile, end tile)
    int* d_slab_heavy = nullptr; // rows holding such a pair (nheavy), then their K slot entries each
    int nheavy = 0;
    int* d_slab_ci = nullptr;
    void* d_slab_v = nullptr;
    uint32_t* d_slab_flags = nullptr;
    int* d_slab_rpp = nullptr;   // row-major pair offsets [R+1]
    int* d_slab_ckey = nullptr;  // carry keys (pair ids) per tile
    // LDS column slabs (lslab): slab streams (16-bit column offsets with row-end bits, values), tile
    // descriptors, chunk descriptors (launch order), combine blocks, run tables, row-major pair map, pair sums
    int ls_S = 0, ls_chunks = 0, ls_blocks = 0;
    long ls_pairs = 0, ls_pspan = 0;  // pair sums, and P's length in pair sums (blocks start 16-byte aligned)
    uint16_t* d_ls_lc = nullptr;
    void* d_ls_lv = nullptr;
    int* d_ls_tp = nullptr;   // per tile: (first pair | mid-pair bit, run in its chunk)
    int* d_ls_ch = nullptr;   // per launch slot: tile0, tiles, first column, run table offset
    int* d_ls_blk = nullptr;  // per combine block: first row, rows, first pair, pairs, wave splits
    int* d_ls_F = nullptr;    // run tables of the chunks: (first pair, P offset) per run, then (INT_MAX, 0)
    uint16_t* d_ls_map = nullptr;  // per combine block, from a dword boundary
    void* d_ls_P = nullptr;
    // checked plans (muxi_plan_create_checked): each row's a-posteriori error bound is tested against rtol * |y|
    // and the rows that fail are recomputed in FP64; d_ls_chk = [count of such rows in the last exec, their ids]
    bool check = false;
    float rtol = 0.f;
    int* d_ls_chk = nullptr;
};

namespace muxi {

template <typename V, typename X, typename A, bool C>
struct Tag {
    using VT = V;
    using XT = X;
    using AT = A;
    static constexpr bool COMP = C;
};

// Calls f(Tag<...>{}) for the element-type combination, or returns
// SPMV_ERROR_UNSUPPORTED_TYPE.
template <typename F>
spmv_status_t with_types(spmv_type_t vt, spmv_type_t xt, bool comp, F&& f) {
    if (xt == SPMV_TYPE_FLOAT) {
        if (vt == SPMV_TYPE_FLOAT)
            return comp ? f(Tag<float, float, float, true>{}) : f(Tag<float, float, float, false>{});
        if (vt == SPMV_TYPE_BF16)
            return comp ? f(Tag<bf16_t, float, float, true>{}) : f(Tag<bf16_t, float, float, false>{});
        if (vt == SPMV_TYPE_DOUBLE)
            return comp ? f(Tag<double, float, float, true>{}) : f(Tag<double, float, float, false>{});
    } else if (xt == SPMV_TYPE_DOUBLE && vt == SPMV_TYPE_DOUBLE && !comp) {
        return f(Tag<double, double, double, false>{});
    }
    set_error("unsupported (value type, x type, compensated) combination");
    return SPMV_ERROR_UNSUPPORTED_TYPE;
}

inline int merge_default_partitions(int L, int ipt) {
    long ch = 256L * ipt;
    long p = ((long)L + ch - 1) / ch;
    return (int)(p > 1 ? p : 1);
}

inline size_t vsize(spmv_type_t t) { return t == SPMV_TYPE_DOUBLE ? 8 : (t == SPMV_TYPE_BF16 ? 2 : 4); }

// Slabs for a column-slab plan when the caller gives none: the fewest (1, 2, 4
// or 8) that keep one slab of x within 2 MiB, half of an XCD's L2.
inline int xslab_default_slabs(long cols, int xbytes) {
    const long b = cols * xbytes;
    int S = 1;
    while (S < 8 && b > (long)S * (2L << 20)) S <<= 1;
    return S;
}

// A plan for a row block (rows [row_base, row_base + A->numRows) of a matrix of rows_full rows, numbered from 0,
// full-matrix columns): AUTO picks from the block's features against the full diagonal, and the ELL plan takes
// its 16-bit column offsets from the full-matrix row. muxi_plan_create is the case row_base = 0,
// rows_full = A->numRows.
spmv_status_t plan_create_rows(const muxi_csr_t* A, spmv_type_t vt, spmv_type_t xt, muxi_kernel_t kernel, int param0,
                               int compensated, void* stream, long row_base, long rows_full, muxi_plan** out);

// column-slab plan and launcher (kernels_xslab.hip); param0 = (slabs << 8) | nonzeros per lane (4 or 8)
spmv_status_t xslab_build(muxi_plan* p, int S, int ipt, spmv_type_t vt, hipStream_t s);
template <typename VT, typename XT, typename AT>
spmv_status_t launch_xslab(const muxi_plan* p, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s,
                           hipEvent_t ev_end);

// LDS column-slab plan and launcher (kernels_lslab.hip); param0 = nonzeros per lane (4 or 8) | kLsHalfSlabs;
// FP32 x and y
spmv_status_t lslab_build(muxi_plan* p, int param0, spmv_type_t vt, hipStream_t s);
template <typename VT, bool COMP>
spmv_status_t launch_lslab(const muxi_plan* p, const float* x, float* y, float alpha, float beta, hipStream_t s,
                           hipEvent_t ev_end);

}  // namespace muxi
