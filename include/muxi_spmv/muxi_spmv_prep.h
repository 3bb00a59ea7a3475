/*
 * muxi_spmv_prep.h — preprocessing: row and column permutations.
 *
 * Host entry points take host arrays, like the reference's host functions;
 * the _dev entry points apply a permutation to device arrays on a stream.
 * Maps: forwardMap / colPermutation / rowMap are new -> old; the inverse maps
 * are old -> new.
 */
#ifndef MUXI_SPMV_PREP_H_
#define MUXI_SPMV_PREP_H_

#include "muxi_spmv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ColumnReorderStats (src/preprocessing/column_reorder.cuh:23-28). */
typedef struct {
    int numReorders;
    float avgDistanceBefore;
    float avgDistanceAfter;
    float improvementRatio;
} muxi_column_reorder_stats_t;

/* reorderColumns (column_reorder.cuh:42-141): columns ordered by their FP32
 * centre of mass (mean row index, accumulated in row order), ties by column.
 * Writes forwardMap[numCols] and newColIdx[nnz] (rowPtr and values are
 * unchanged). stats may be NULL. */
spmv_status_t muxi_reorder_columns(int numRows, int numCols, const int* rowPtr, const int* colIdx,
                                   int* forwardMap, int* newColIdx, muxi_column_reorder_stats_t* stats);

/* reorderXVector (column_reorder.cuh:150-161): xr[new] = x[forwardMap[new]]. */
spmv_status_t muxi_reorder_x_f32(const float* x, float* xr, const int* forwardMap, int numCols);

/* sortColumnsWithinRows (column_reorder.cuh:175-201): each row's (col, value)
 * pairs sorted ascending (col, then value), in place. */
spmv_status_t muxi_sort_columns_within_rows_f32(int numRows, const int* rowPtr, int* colIdx, float* values);

/* reorderColumnsRCM (tests/test_column_reordering.cu:156-197): columns sorted
 * by the first row that uses them (unused columns: numRows) with the
 * reference's unstable std::sort; colPermutation new -> old. */
spmv_status_t muxi_reorder_columns_minrow(int numRows, int numCols, const int* rowPtr, const int* colIdx,
                                          int* colPermutation, int* newColIdx);

/* reorderColumnsByFrequency (tests/test_column_reordering.cu:116-153):
 * columns by descending use count, unstable std::sort. */
spmv_status_t muxi_reorder_columns_by_frequency(int numRows, int numCols, const int* rowPtr, const int* colIdx,
                                                int* colPermutation, int* newColIdx);

/* reorderMatrixByRowLength (tests/test_row_reordering.cu:99-141): rows in
 * ascending length, stable. Outputs rowMap/invRowMap[numRows] and the
 * permuted CSR (newRowPtr[numRows+1], newColIdx[nnz], newValues[nnz]). */
spmv_status_t muxi_reorder_rows_by_length_f32(int numRows, const int* rowPtr, const int* colIdx,
                                              const float* values, int* rowMap, int* invRowMap, int* newRowPtr,
                                              int* newColIdx, float* newValues);

/* Reverse Cuthill-McKee on the symmetrised pattern of a square matrix (no
 * counterpart in the reference, whose "RCM" is the min-row sort above):
 * each connected component from a pseudo-peripheral start (George-Liu),
 * neighbours visited by ascending degree then index, the whole order
 * reversed. perm[new] = old. */
spmv_status_t muxi_rcm(int n, const int* rowPtr, const int* colIdx, int* perm);

/* Symmetric permutation B = P A P^T of a square matrix given perm (new ->
 * old): row i of B is row perm[i] of A with columns mapped old -> new and
 * re-sorted. Host arrays; values FP32 or FP64 by vtype. */
spmv_status_t muxi_permute_symmetric(int n, const int* rowPtr, const int* colIdx, const void* values,
                                     spmv_type_t vtype, const int* perm, int* newRowPtr, int* newColIdx,
                                     void* newValues);

/* Half-bandwidth max |i - j| over the nonzeros. */
int muxi_bandwidth(int numRows, const int* rowPtr, const int* colIdx);

/* Row shards for G GPUs. No counterpart in the reference (it has no multi-GPU
 * path); for callers that split one SpMV over several devices with x
 * replicated. The merge path (numRows row ends + nnz nonzeros) is cut at the G-1
 * diagonals d_g = floor(g * (numRows + nnz) / G), and each cut moves to the
 * start of the row that holds it (the merge-path coordinate row at d_g), so
 * shard g is rows [bounds[g], bounds[g+1]) with about (numRows + nnz) / G path
 * items, and no row is split. bounds[0] = 0, bounds[G] = numRows. Host arrays. */
spmv_status_t muxi_shard_rows(const int* rowPtr, int numRows, int G, int* bounds);

/* ---- device application ---- */
/* out[i] = map[in[i]] for i < n (column remap by an old -> new map). */
spmv_status_t muxi_remap_indices_dev(const int* d_in, const int* d_map, int* d_out, int n, void* stream);
/* dst[i] = src[idx[i]] (reorderXVector on the device); elem_bytes 2, 4 or 8. */
spmv_status_t muxi_gather_dev(const void* d_src, const int* d_idx, void* d_dst, int n, int elem_bytes, void* stream);
/* dst[idx[i]] = src[i] (restoreOrder, tests/test_combined_optimizations.cu:176-181). */
spmv_status_t muxi_scatter_dev(const void* d_src, const int* d_idx, void* d_dst, int n, int elem_bytes, void* stream);
/* Row gather: row i of the output is row rowMap[i] of the input; d_newRowPtr
 * must already hold the permuted row offsets. */
spmv_status_t muxi_permute_rows_dev(int numRows, const int* d_rowPtr, const int* d_colIdx, const void* d_values,
                                    int elem_bytes, const int* d_rowMap, const int* d_newRowPtr, int* d_newColIdx,
                                    void* d_newValues, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MUXI_SPMV_PREP_H_ */
