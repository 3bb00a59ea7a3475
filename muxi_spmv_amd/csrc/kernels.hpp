// Internal launcher declarations. Every launcher enqueues its kernels on the
// given stream and returns SPMV_SUCCESS or SPMV_ERROR_DEVICE; none allocates.
#pragma once

#include <vector>

#include "common.hpp"

namespace muxi {

struct DevCsr {
    int R, C, N;
    const int* rp;
    const int* ci;
    const void* v;
};

// Element-type combinations the library is built for: (value, x/y, accumulator, compensated).
#define MUXI_FOR_EACH_COMBO(M)            \
    M(float, float, float, false)         \
    M(double, double, double, false)      \
    M(float, float, float, true)          \
    M(bf16_t, float, float, false)        \
    M(bf16_t, float, float, true)         \
    M(double, float, float, false)        \
    M(double, float, float, true)

// ---- row-parallel CSR (kernels_csr.hip)
template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_scalar(const DevCsr& A, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s);
template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_vector(const DevCsr& A, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s);
// lanes: 1,2,4,8,16,32,64 lanes per row; rowPtr of the block's rows staged in LDS.
template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_vwave(const DevCsr& A, int lanes, const XT* x, XT* y, AT alpha, AT beta, hipStream_t s);

// ---- merge-path (kernels_merge.hip)
// Items (rows + nnz) one workgroup handles per partition, for ipt items per thread.
int merge_chunk_items(int ipt);
spmv_status_t launch_merge_partitions(int R, int N, int P, int* d_pos, hipStream_t s);
spmv_status_t launch_merge_coords(const int* rp, int R, int N, const int* d_pos, int count, int* d_row,
                                  int* d_nnz, hipStream_t s);
// Carry buffer: P entries of the accumulator type (Acc<AT,COMP>).
size_t merge_carry_bytes(int P, spmv_type_t xtype, bool comp);
template <typename VT, typename XT, typename AT, bool COMP>
// ev_main_end (optional) is recorded after the main kernel, before the carry fix-up.
spmv_status_t launch_merge(const DevCsr& A, int ipt, const int* d_pos, const int* d_row, int P, void* carry,
                           const XT* x, XT* y, AT alpha, AT beta, hipStream_t s, hipEvent_t ev_main_end = nullptr);

// ---- CSR5 tiles (kernels_csr5.hip)
spmv_status_t launch_csr5_tiles(const int* rp, int R, int N, int sigma, int T, int* d_tile_row, int* d_tile_off,
                                hipStream_t s);
bool csr5_sigma_supported(int sigma);
size_t csr5_carry_bytes(int T, spmv_type_t xtype, bool comp);
template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_csr5(const DevCsr& A, int sigma, const int* d_tile_row, int T, void* carry, const XT* x,
                          XT* y, AT alpha, AT beta, hipStream_t s, hipEvent_t ev_main_end = nullptr);

// ---- ELLPACK (kernels_ell.hip)
spmv_status_t launch_max_row_len(const int* rp, int R, int* d_out, hipStream_t s);
template <typename VT>
spmv_status_t launch_csr_to_ell(const DevCsr& A, int width, int* d_col, VT* d_val, hipStream_t s);
template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_ell(int R, int C, int width, const int* d_col, const VT* d_val, const XT* x, XT* y, AT alpha,
                         AT beta, hipStream_t s);
// the plan's layout: a row's slots grouped by four (kernels_ell.hip); with
// off16 the columns are 16-bit offsets from the row (d_col holds shorts),
// allowed when launch_ell_offsets_fit left *d_bad == 0
// (offsets from row + rb: rb is a row block's first row in its full matrix, 0 for a whole matrix;
// d_bad[0] = 1 when some offset needs more than 16 bits, d_bad[1] = the largest |offset|, capped at 32768)
spmv_status_t launch_ell_offsets_fit(const DevCsr& A, int width, int rb, int* d_bad, hipStream_t s);
template <typename VT>
spmv_status_t launch_csr_to_ellg(const DevCsr& A, int width, bool off16, int rb, void* d_col, VT* d_val,
                                 hipStream_t s);
// win >= 0 (with off16): every column lies within win of its row; the kernels stage that x window in LDS
template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_ellg(int R, int C, int width, bool off16, int rb, int win, const void* d_col, const VT* d_val,
                          const XT* x, XT* y, AT alpha, AT beta, hipStream_t s);

// ---- SpMM over merge-path partitions (kernels_spmm.hip); carry holds P*k elements.
int spmm_groups(int k);
int spmm_chunk_items(int k, int ipt);
template <typename VT, typename XT>
spmv_status_t launch_spmm(const DevCsr& A, int k, int ipt, const int* d_pos, const int* d_row, int P, void* carry,
                          const XT* X, XT* Y, XT alpha, XT beta, hipStream_t s, hipEvent_t ev = nullptr);

// ---- adaptive row blocks (kernels_adaptive.hip)
spmv_status_t adaptive_build(const int* d_rp, int R, int nb, hipStream_t s, std::vector<int>& blk_row,
                             std::vector<int>& blk_chunk, std::vector<int>& long_row, std::vector<int>& long_first);
template <typename VT, typename XT, typename AT, bool COMP>
spmv_status_t launch_adaptive(const DevCsr& A, int nb, int nblocks, const int* blk_row, const int* blk_chunk,
                              int nlong, const int* long_row, const int* long_first, void* partial, const XT* x, XT* y,
                              AT alpha, AT beta, hipStream_t s, hipEvent_t ev = nullptr);

// ---- misc (kernels_misc.hip)
template <typename VT, typename XT>
spmv_status_t launch_oracle_f64(const DevCsr& A, const XT* x, double* y, hipStream_t s);
spmv_status_t launch_convert(const void* src, spmv_type_t st, void* dst, spmv_type_t dt, size_t n, hipStream_t s);
template <typename XT>
spmv_status_t launch_scale_y(XT* y, long n, XT beta, hipStream_t s);
template <typename VT, typename XT>
spmv_status_t launch_transpose_scalar(const DevCsr& A, const XT* x, XT* y, XT alpha, hipStream_t s);

}  // namespace muxi
