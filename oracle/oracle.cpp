// CPU oracle for the SpMV path — TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference's algorithms, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline as the checker and
// the CPU baseline. Nothing in the library links or calls it.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp), output
// oracle/liboracle.so. Every function cites the reference file:line it follows
// (paths relative to the reference repository root).
//
// Pinning (how each function is checked independently of retyping it):
//   spmv_csr_f32       diagonal known answer (tests/correctness/test_spmv.cu:63-67),
//                      Laplacian/banded closed forms, agreement with spmv_f64
//                      within the recursive-summation bound.
//   spmv_f64           closed forms; exact on small-integer data; Dot2 vs
//                      Python's exactly-rounded math.fsum per row.
//   spmv_csr_f64_plain exact integer product (numpy int64) and bit-equality
//                      with spmv_f64 on small-integer data; Laplacian and
//                      diagonal closed forms; recursive-summation bound vs fsum.
//   merge_*            brute force over all path diagonals on tiny matrices
//                      (the coordinate is the unique r with
//                      rowPtr[r]+r <= k <= rowPtr[r+1]+r) and the reference
//                      constants (P = 3328 at the cfg sizes).
//   csr5_tiles         numpy searchsorted(side="right")-1 on random rowPtr,
//                      incl. runs of empty rows.
//   csr_to_ell         dense reconstruction equals the CSR matrix.
//   spmv_ell_f64       exact integer product of the CSR matrix (padded widths)
//                      and of the matrix cut to the first w entries per row
//                      (truncated widths).
//   permutations       bijection checks, key order, golden fixtures; the
//                      unstable std::sort ties are libstdc++'s and are pinned
//                      only by the committed fixtures (see header of
//                      tests/golden/make_golden.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ---------------------------------------------------------------- SpMV loops
// tests/test_column_reorder.cu:87-93 (same loop: tests/deep_analysis.cu:189-195)
void oracle_spmv_csr_f32(int R, const int* rowPtr, const int* colIdx, const float* values, const float* x,
                         float* y) {
    for (int row = 0; row < R; row++) {
        float sum = 0;
        for (int idx = rowPtr[row]; idx < rowPtr[row + 1]; idx++) sum += values[idx] * x[colIdx[idx]];
        y[row] = sum;
    }
}

// Same loop, rows split over `threads` OpenMP threads (CPU baseline).
void oracle_spmv_csr_f32_mt(int R, const int* rowPtr, const int* colIdx, const float* values, const float* x,
                            float* y, int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 4096) num_threads(threads)
#endif
    for (int row = 0; row < R; row++) {
        float sum = 0;
        for (int idx = rowPtr[row]; idx < rowPtr[row + 1]; idx++) sum += values[idx] * x[colIdx[idx]];
        y[row] = sum;
    }
}

void oracle_spmv_csr_f64_plain(int R, const int* rowPtr, const int* colIdx, const double* values, const double* x,
                               double* y) {
    for (int row = 0; row < R; row++) {
        double sum = 0;
        for (int idx = rowPtr[row]; idx < rowPtr[row + 1]; idx++) sum += values[idx] * x[colIdx[idx]];
        y[row] = sum;
    }
}

// FP64 oracle: Dot2 (Ogita, Rump, Oishi 2005) in double, i.e. about twice
// double precision before the final rounding. Also returns sum |a_ij x_j|
// per row (absrow, may be NULL) for error bounds.
static inline void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    double bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}
static inline void two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    e = std::fma(a, b, -p);
}
void oracle_spmv_f64(int R, const int* rowPtr, const int* colIdx, const double* values, const double* x, double* y,
                     double* absrow) {
    for (int row = 0; row < R; row++) {
        double s = 0, c = 0, ab = 0;
        for (int idx = rowPtr[row]; idx < rowPtr[row + 1]; idx++) {
            double p, ep, t, es;
            two_prod(values[idx], x[colIdx[idx]], p, ep);
            two_sum(s, p, t, es);
            s = t;
            c += es + ep;
            ab += std::fabs(values[idx] * x[colIdx[idx]]);
        }
        y[row] = s + c;
        if (absrow) absrow[row] = ab;
    }
}

// ---------------------------------------------------------------- merge path
// spmv_csr5.cu:942-972 (ePP=16, mult=8, nSM=104 on a 64-wide wavefront),
// :1031-1038 (_v2: ePP=64, mult=4), ultra_optimized_test.cu:36-48 (ePP=8, mult=16)
int oracle_merge_partition_count(int numRows, int nnz, int ePP, int mult, int numSMs, int warpsPerBlock) {
    int mergePathLength = numRows + nnz;
    int targetPartitions = mergePathLength / ePP;
    int maxPartitions = numSMs * warpsPerBlock * mult;
    int p = std::min(targetPartitions, maxPartitions);
    return std::max(1, p);
}

// compute_merge_partitions_kernel, spmv_csr5.cu:912-925
void oracle_merge_partitions(int numRows, int nnz, int numPartitions, int* pos) {
    int mergePathLength = numRows + nnz;
    int pathPerPartition = (mergePathLength + numPartitions - 1) / numPartitions;
    for (int tid = 0; tid <= numPartitions; tid++) pos[tid] = std::min(tid * pathPerPartition, mergePathLength);
}

// merge_path_search, spmv_csr5.cu:497-533 (the reference's loop shape)
void oracle_merge_path_search(const int* rowPtr, int numRows, int nnz, int k, int* rowIdx, int* nnzIdx) {
    int lo = std::max(0, k - nnz);
    int hi = std::min(k, numRows);
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        int nnz_at_mid = k - mid;
        if (mid < numRows && nnz_at_mid > rowPtr[mid + 1]) {
            lo = mid + 1;
        } else if (mid > 0 && nnz_at_mid < rowPtr[mid]) {
            hi = mid;
        } else {
            lo = mid;
            hi = mid;
        }
    }
    int r = lo, n = k - lo;
    if (r > numRows) r = numRows;
    if (n > nnz) n = nnz;
    *rowIdx = r;
    *nnzIdx = n;
}

void oracle_merge_path_coords(const int* rowPtr, int numRows, int nnz, const int* pos, int count, int* rowIdx,
                              int* nnzIdx) {
    for (int t = 0; t < count; t++) oracle_merge_path_search(rowPtr, numRows, nnz, pos[t], &rowIdx[t], &nnzIdx[t]);
}

// ---------------------------------------------------------------- CSR5 tiles
// csr5_binary_search_row, spmv_csr5.cuh:44-61; tile kernel spmv_csr5.cu:36-58
static int csr5_binary_search_row(const int* rowPtr, int numRows, int nnzIdx) {
    int lo = 0, hi = numRows;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (rowPtr[mid] <= nnzIdx)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}
void oracle_csr5_tiles(const int* rowPtr, int numRows, int nnz, int sigma, int numTiles, int* tile_row_ptr,
                       int* tile_nnz_offset) {
    (void)nnz;
    for (int tid = 0; tid < numTiles; tid++) {
        int tileStart = tid * sigma;
        int row = csr5_binary_search_row(rowPtr, numRows, tileStart);
        tile_row_ptr[tid] = row;
        tile_nnz_offset[tid] = tileStart - rowPtr[row];
    }
}
// csr5_preprocess sigma default on a 64-wide wavefront: getCSR5Config, spmv_csr5.cu:13-32
int oracle_csr5_default_sigma(int warpSize) { return warpSize == 64 ? 256 : 128; }

// ---------------------------------------------------------------- ELLPACK
// csr_to_ellpack, spmv_ellpack.cu:11-70 (values as double; exact for f32/f64)
int oracle_ell_width(int numRows, const int* rowPtr) {
    int maxElements = 0;
    for (int i = 0; i < numRows; i++) maxElements = std::max(maxElements, rowPtr[i + 1] - rowPtr[i]);
    return maxElements;
}
void oracle_csr_to_ell(int numRows, const int* rowPtr, const int* colIdx, const double* values, int maxElements,
                       int* ecol, double* evals) {
    long total = (long)numRows * maxElements;
    for (long i = 0; i < total; i++) {
        ecol[i] = -1;
        evals[i] = 0.0;
    }
    for (int row = 0; row < numRows; row++) {
        int rowStart = rowPtr[row], rowLen = rowPtr[row + 1] - rowStart;
        for (int i = 0; i < rowLen && i < maxElements; i++) {
            long idx = row + (long)i * numRows;
            ecol[idx] = colIdx[rowStart + i];
            evals[idx] = values[rowStart + i];
        }
    }
}
// spmv_ellpack_kernel, spmv_ellpack.cuh:85-111 (FP64 arithmetic)
void oracle_spmv_ell_f64(int numRows, int numElements, const int* ecol, const double* evals, const double* x,
                         double* y) {
    for (int row = 0; row < numRows; row++) {
        double sum = 0;
        for (int i = 0; i < numElements; i++) {
            long idx = row + (long)i * numRows;
            int col = ecol[idx];
            if (col >= 0) sum += evals[idx] * x[col];
        }
        y[row] = sum;
    }
}

// ---------------------------------------------------------------- permutations
// reorderColumns, src/preprocessing/column_reorder.cuh:42-141: float centre of
// mass accumulated in row order, std::sort of pair<float,int>; returns
// forwardMap (new -> old) and the remapped colIdx (rowPtr/values unchanged).
void oracle_reorder_columns_com(int numRows, int numCols, const int* rowPtr, const int* colIdx, int* forwardMap,
                                int* newColIdx) {
    std::vector<float> com(numCols, 0.0f);
    std::vector<int> cnt(numCols, 0);
    for (int row = 0; row < numRows; row++)
        for (int idx = rowPtr[row]; idx < rowPtr[row + 1]; idx++) {
            int col = colIdx[idx];
            com[col] += row;
            cnt[col]++;
        }
    for (int col = 0; col < numCols; col++)
        if (cnt[col] > 0) com[col] /= cnt[col];
    std::vector<std::pair<float, int>> sorted(numCols);
    for (int col = 0; col < numCols; col++) sorted[col] = {com[col], col};
    std::sort(sorted.begin(), sorted.end());
    std::vector<int> reverse(numCols);
    for (int nc = 0; nc < numCols; nc++) {
        forwardMap[nc] = sorted[nc].second;
        reverse[sorted[nc].second] = nc;
    }
    int nnz = rowPtr[numRows];
    for (int idx = 0; idx < nnz; idx++) newColIdx[idx] = reverse[colIdx[idx]];
}

// reorderXVector, column_reorder.cuh:150-161
void oracle_reorder_x(int numCols, const float* x, const int* forwardMap, float* xr) {
    for (int nc = 0; nc < numCols; nc++) xr[nc] = x[forwardMap[nc]];
}

// sortColumnsWithinRows, column_reorder.cuh:175-201 (pair<int, float>, in place)
void oracle_sort_columns_within_rows(int numRows, const int* rowPtr, int* colIdx, float* values) {
    for (int row = 0; row < numRows; row++) {
        int start = rowPtr[row], end = rowPtr[row + 1];
        std::vector<std::pair<int, float>> el;
        for (int idx = start; idx < end; idx++) el.push_back({colIdx[idx], values[idx]});
        std::sort(el.begin(), el.end());
        for (size_t i = 0; i < el.size(); i++) {
            colIdx[start + i] = el[i].first;
            values[start + i] = el[i].second;
        }
    }
}

// reorderColumnsRCM, tests/test_column_reordering.cu:156-197 (= tests/test_rcm_all_matrices.cu:84-118):
// columns sorted by the first row that uses them (unused: numRows), std::sort
// with a comparator on .second only. colPermutation = new -> old.
void oracle_reorder_columns_minrow(int numRows, int numCols, const int* rowPtr, const int* colIdx,
                                   int* colPermutation, int* newColIdx) {
    std::vector<std::pair<int, int>> colMinRow(numCols);
    for (int i = 0; i < numCols; i++) colMinRow[i] = {i, numRows};
    for (int row = 0; row < numRows; row++)
        for (int idx = rowPtr[row]; idx < rowPtr[row + 1]; idx++) {
            int col = colIdx[idx];
            colMinRow[col].second = std::min(colMinRow[col].second, row);
        }
    std::sort(colMinRow.begin(), colMinRow.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return a.second < b.second;
    });
    std::vector<int> inv(numCols);
    for (int i = 0; i < numCols; i++) {
        colPermutation[i] = colMinRow[i].first;
        inv[colMinRow[i].first] = i;
    }
    int nnz = rowPtr[numRows];
    for (int i = 0; i < nnz; i++) newColIdx[i] = inv[colIdx[i]];
}

// reorderColumnsByFrequency, tests/test_column_reordering.cu:116-153 (descending count, std::sort)
void oracle_reorder_columns_freq(int numRows, int numCols, const int* rowPtr, const int* colIdx, int* colPermutation,
                                 int* newColIdx) {
    std::vector<std::pair<int, int>> colFreq(numCols);
    for (int i = 0; i < numCols; i++) colFreq[i] = {i, 0};
    int nnz = rowPtr[numRows];
    for (int i = 0; i < nnz; i++) colFreq[colIdx[i]].second++;
    std::sort(colFreq.begin(), colFreq.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return a.second > b.second;
    });
    std::vector<int> inv(numCols);
    for (int i = 0; i < numCols; i++) {
        colPermutation[i] = colFreq[i].first;
        inv[colFreq[i].first] = i;
    }
    for (int i = 0; i < nnz; i++) newColIdx[i] = inv[colIdx[i]];
}

// reorderMatrixByRowLength, tests/test_row_reordering.cu:99-141 (std::stable_sort ascending length)
void oracle_reorder_rows_by_length(int numRows, const int* rowPtr, const int* colIdx, const float* values, int* rowMap,
                                   int* invRowMap, int* newRowPtr, int* newColIdx, float* newValues) {
    std::vector<int> len(numRows), idx(numRows);
    for (int i = 0; i < numRows; i++) {
        len[i] = rowPtr[i + 1] - rowPtr[i];
        idx[i] = i;
    }
    std::stable_sort(idx.begin(), idx.end(), [&len](int a, int b) { return len[a] < len[b]; });
    for (int i = 0; i < numRows; i++) {
        rowMap[i] = idx[i];
        invRowMap[idx[i]] = i;
    }
    newRowPtr[0] = 0;
    int pos = 0;
    for (int i = 0; i < numRows; i++) {
        int o = rowMap[i];
        for (int j = rowPtr[o]; j < rowPtr[o + 1]; j++) {
            newColIdx[pos] = colIdx[j];
            newValues[pos] = values[j];
            pos++;
        }
        newRowPtr[i + 1] = pos;
    }
}

// restoreOrder, tests/test_combined_optimizations.cu:176-181
void oracle_restore_order(int numRows, const float* reordered_y, const int* rowMap, float* y) {
    for (int i = 0; i < numRows; i++) y[rowMap[i]] = reordered_y[i];
}

}  // extern "C"
