/*
 * muxi_spmv_batch.h — many independent SpMVs in one persistent launch.
 *
 * No counterpart in the reference (its "batched" kernels,
 * src/spmv/csr/spmv_csr.cuh:545-869, batch rows of one matrix). A batch plan
 * binds `count` CSR matrices and their x / y vectors; every matrix is cut
 * into merge-path tiles of the same path length and one persistent kernel
 * (grid sized to the device, each XCD on its own contiguous eighth of the
 * tiles) works through all tiles of all matrices, then
 * one fix-up kernel adds the partial sums of rows that span tiles. Exec takes
 * no pointers, allocates nothing, and can be captured in a hipGraph
 * (muxi_batch_graph_*).
 *
 * Slab mode (FP32): the batch runs as one SpMV with the block-diagonal matrix
 * of its members, x and y laid end to end, through an LDS column-slab plan
 * (muxi_plan_create with MUXI_KERNEL_LSLAB) built once over that matrix; an
 * exec gathers the members' x, runs the slab plan and scatters y (four
 * kernels, capturable like the merge mode).
 */
#ifndef MUXI_SPMV_BATCH_H_
#define MUXI_SPMV_BATCH_H_

#include "muxi_spmv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct muxi_batch_plan muxi_batch_plan;

#define MUXI_BATCH_SLABS 0x10000 /* ipt flag: the slab mode */

/* mats[count] with device arrays; xs[count] / ys[count] device vectors of
 * numCols / numRows elements (x may be NULL for a member without nonzeros, y
 * for a member without rows). Each member's rowPtr runs from 0 to its nnz
 * (the slab mode checks it: SPMV_ERROR_INVALID_MATRIX). vtype FLOAT (xtype FLOAT) or DOUBLE (xtype
 * DOUBLE). ipt selects the mode and schedule:
 *   0: the slab mode when the batch qualifies -- FLOAT, at least 65,536
 *      columns and 8,388,608 nonzeros in all, at most 16,777,216 columns, rows
 *      + nonzeros within INT32_MAX -- else the tuned merge schedule;
 *   < 0: the merge mode with the schedule table the project's autotuner
 *      fitted (by the batch's mean nonzeros per row and mean rows per matrix;
 *      muxi_spmv_amd/sched_tune.py);
 *   MUXI_BATCH_SLABS set: the slab mode (SPMV_ERROR_UNSUPPORTED_FORMAT when
 *      the batch does not fit it);
 *   other > 0: merge items per lane (4, 8 or 16) | workgroups per CU << 8
 *      (1 to 32; 0 = 8).
 * The slab mode holds the block-diagonal matrix's row pointers, its slab plan
 * and one x and one y of the batch's total length; it copies the members'
 * column indices and values into that plan when it is created, so later
 * changes to them are not seen (the merge mode reads the members' arrays at
 * every exec). Both modes bind the xs / ys pointers at creation. */
spmv_status_t muxi_batch_plan_create(const muxi_csr_t* mats, int count, spmv_type_t vtype, spmv_type_t xtype,
                                     const void* const* xs, void* const* ys, int ipt, void* stream,
                                     muxi_batch_plan** plan);
spmv_status_t muxi_batch_plan_destroy(muxi_batch_plan* plan);
/* ys[m] = alpha * mats[m] * xs[m] + beta * ys[m] for every m. */
spmv_status_t muxi_batch_exec(const muxi_batch_plan* plan, double alpha, double beta, void* stream);
/* Like muxi_batch_exec, recording ev_before / ev_after around the whole batch. */
spmv_status_t muxi_batch_exec_ev(const muxi_batch_plan* plan, double alpha, double beta, void* stream,
                                 void* ev_before, void* ev_after);
/* The plain per-matrix path for comparison: one launch of the given row
 * kernel (MUXI_KERNEL_SCALAR, MUXI_KERNEL_VECTOR or MUXI_KERNEL_VWAVE with
 * param0 lanes; any other id is SPMV_ERROR_INTERNAL) per matrix, y = A*x. */
spmv_status_t muxi_batch_exec_per_matrix(const muxi_batch_plan* plan, muxi_kernel_t kernel, int param0,
                                         void* stream);
/* Tiles, grid size and scratch bytes of the plan (slab mode: 0 tiles, the
 * slab plan's workgroups, everything it holds). */
spmv_status_t muxi_batch_info(const muxi_batch_plan* plan, int* tiles, int* grid, size_t* scratch_bytes);
/* The schedule the plan runs: merge items per lane and workgroups per CU
 * (0 and 0 in the slab mode). */
spmv_status_t muxi_batch_schedule(const muxi_batch_plan* plan, int* ipt, int* wg_per_cu);

/* hipGraph of one muxi_batch_exec (per_matrix_kernel = 0) or of one launch
 * per matrix with that kernel id (SCALAR, VECTOR or VWAVE with param0 lanes;
 * any other id is SPMV_ERROR_INTERNAL), both computing
 * ys[m] = alpha * mats[m] * xs[m] + beta * ys[m], captured on `stream` (must
 * not be the null stream), then replayed by muxi_graph_launch. */
typedef struct muxi_graph muxi_graph;
spmv_status_t muxi_batch_graph_create(const muxi_batch_plan* plan, double alpha, double beta,
                                      int per_matrix_kernel, int param0, void* stream, muxi_graph** graph);
spmv_status_t muxi_graph_launch(const muxi_graph* graph, void* stream);
spmv_status_t muxi_graph_destroy(muxi_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* MUXI_SPMV_BATCH_H_ */
