// Tuned schedules for the batched launch and SpMM: the tables the project autotuner
// (muxi_spmv_amd/sched_tune.py) fits from its own timing sweep, compiled in from the
// header build.py writes (_build/sched_tables.inc). Without the header the fixed
// defaults below apply.
#pragma once

#if __has_include("sched_tables.inc")
#include "sched_tables.inc"
#define MUXI_HAVE_SCHED_TABLES 1
#endif

namespace muxi {

// Bin of v against ascending edges e[0..n): the number of edges <= v.
inline int sched_bin(double v, const double* e, int n) {
    int b = 0;
    while (b < n && v >= e[b]) ++b;
    return b;
}

// Batched launch: merge items per lane and workgroups per CU, from the batch's mean nonzeros per row
// (all members together) and mean rows per member.
inline void batch_schedule(double mean_row_len, double mean_rows, int* ipt, int* wg_per_cu) {
#ifdef MUXI_HAVE_SCHED_TABLES
    const int a = sched_bin(mean_row_len, kBatchRowLenEdges, kBatchRowLenBins - 1);
    const int b = sched_bin(mean_rows, kBatchRowsEdges, kBatchRowsBins - 1);
    *ipt = kBatchIpt[a * kBatchRowsBins + b];
    *wg_per_cu = kBatchWgPerCu[a * kBatchRowsBins + b];
#else
    (void)mean_row_len;
    (void)mean_rows;
    *ipt = 16;
    *wg_per_cu = 8;
#endif
}

// SpMM: merge items per lane group for k right-hand sides (the largest tuned k not above k) and the
// matrix's merge-path length numRows + nnz.
inline int spmm_schedule(int k, double path_len) {
#ifdef MUXI_HAVE_SCHED_TABLES
    int ki = 0;
    while (ki + 1 < kSpmmKCount && kSpmmK[ki + 1] <= k) ++ki;
    const int a = sched_bin(path_len, kSpmmPathEdges, kSpmmPathBins - 1);
    return kSpmmIpt[ki * kSpmmPathBins + a];
#else
    (void)k;
    (void)path_len;
    return 8;
#endif
}

}  // namespace muxi
