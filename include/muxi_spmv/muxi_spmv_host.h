/*
 * muxi_spmv_host.h — SpMV with x and y in host memory: pinned buffers and
 * PCIe transfers overlapped with the kernels.
 *
 * Replaces the end-to-end experiments of the reference (pinned and pageable
 * host buffers, synchronous and asynchronous copies, several streams:
 * tests/test_pinned_memory.cu:191-330, row ranges on several streams:
 * tests/test_multi_stream.cu:121-290) with one reusable object.
 *
 * A host plan cuts A into `chunks` row ranges with equal merge-path work
 * (muxi_shard_rows) and builds a device plan for each (its own kernel choice
 * when kernel is MUXI_KERNEL_AUTO). muxi_host_plan_exec then runs, on three
 * streams of default priority:
 *   copy-in:  x in pieces: piece k = the columns below chunk k's largest column
 *             not sent before (so a banded matrix's chunk k starts as soon as
 *             its columns have arrived; a matrix with scattered columns sends
 *             all of x first), plus chunk k's y when beta != 0;
 *   compute:  chunk k's plan once its piece (and y chunk) is on the device;
 *   copy-out: chunk k's y as soon as its kernel is done, while chunk k+1 runs.
 * Host buffers from muxi_host_alloc are pinned (hipHostMalloc) and the copies
 * are then asynchronous DMA; pageable buffers also work, at the rate the
 * driver's staging allows. The device copies of A, x and y live in the plan.
 * exec returns when y is in host memory.
 */
#ifndef MUXI_SPMV_HOST_H_
#define MUXI_SPMV_HOST_H_

#include "muxi_spmv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct muxi_host_plan muxi_host_plan;

/* A's device arrays are used in place (they must outlive the plan); chunks in
 * [1, 64]; vtype / xtype / kernel / param0 as for muxi_plan_create (param0
 * applies to every chunk). */
spmv_status_t muxi_host_plan_create(const muxi_csr_t* A, spmv_type_t vtype, spmv_type_t xtype, int chunks,
                                    muxi_kernel_t kernel, int param0, muxi_host_plan** plan);
spmv_status_t muxi_host_plan_destroy(muxi_host_plan* plan);

/* h_y = alpha*A*h_x + beta*h_y with h_x (numCols) and h_y (numRows) in host
 * memory, elements of xtype. Not reentrant on one plan. */
spmv_status_t muxi_host_plan_exec(muxi_host_plan* plan, const void* h_x, void* h_y, double alpha, double beta);

/* Timings of the last exec, in ms, from events on the device: total (first
 * copy-in to last copy-out), the copy-in of x that precedes the first kernel,
 * the kernels (first start to last end) and the copy-out tail after the last
 * kernel. */
spmv_status_t muxi_host_plan_timing(const muxi_host_plan* plan, float* total_ms, float* first_copy_in_ms,
                                    float* kernels_ms, float* copy_out_tail_ms);

/* Chunk k's rows [row0, row1) and the device plan that runs it. */
spmv_status_t muxi_host_plan_chunk(const muxi_host_plan* plan, int k, int* row0, int* row1, int* kernel,
                                   int* param0);

/* Pinned host memory (hipHostMalloc) for h_x / h_y. */
spmv_status_t muxi_host_alloc(void** ptr, size_t bytes);
spmv_status_t muxi_host_free(void* ptr);

#ifdef __cplusplus
}
#endif

#endif /* MUXI_SPMV_HOST_H_ */
