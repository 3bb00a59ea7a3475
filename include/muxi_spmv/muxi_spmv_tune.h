/*
 * muxi_spmv_tune.h — device matrix generators, matrix features, the
 * autotune sweep's building blocks and the kernel-selector MLP.
 *
 * None of these has a counterpart in the reference (its generators are host
 * code with glibc rand(), src/generators/matrix_generator.cu, and its kernel
 * choice is a fixed avgNnz rule, src/spmv/csr/spmv_csr.cu:115). They exist so
 * that the selector can be trained on the library's own timings at full
 * matrix sizes.
 */
#ifndef MUXI_SPMV_TUNE_H_
#define MUXI_SPMV_TUNE_H_

#include "muxi_spmv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device generators ---------------------------------------------------
 * Families (p0..p2 meaning):
 *   MUXI_GEN_UNIFORM   row length ~ U{p0..p1}
 *   MUXI_GEN_POWERLAW  row length floor(p1 * u^(-1/(p0-1))) clipped to [p1, p2] (Pareto tail of exponent p0)
 *   MUXI_GEN_BANDED    p1 nonzeros per row inside [i-p0, i+p0]
 *   MUXI_GEN_BLOCKDIAG dense p0 x p0 diagonal blocks
 *   MUXI_GEN_POISSON   row length ~ Poisson(p0)
 * Columns of a row of length k over a range [lo, lo+span) are stratified:
 * the range is cut at floor(j*span/k) and one column drawn uniformly in each
 * stratum, so they are distinct and ascending. Values ~ U[-1,1] (or N(0,1)
 * when normal != 0). All randomness is SplitMix64 of (seed, index), so a
 * matrix is a pure function of its parameters.
 * The library allocates d_rowPtr/d_colIdx/d_values (values FLOAT or DOUBLE
 * by vtype); free them with muxi_gen_free. */
typedef enum {
    MUXI_GEN_UNIFORM = 0,
    MUXI_GEN_POWERLAW = 1,
    MUXI_GEN_BANDED = 2,
    MUXI_GEN_BLOCKDIAG = 3,
    MUXI_GEN_POISSON = 4
} muxi_gen_family_t;

spmv_status_t muxi_gen_matrix(muxi_gen_family_t family, int numRows, int numCols, double p0, double p1, double p2,
                              int normal, spmv_type_t vtype, uint64_t seed, void* stream, muxi_csr_t* out);
/* Random symmetric-in-distribution row and column permutation of a matrix
 * made by muxi_gen_matrix (rows gathered, column indices remapped; columns
 * are no longer sorted within rows). Replaces out's arrays. */
spmv_status_t muxi_gen_permute(muxi_csr_t* m, spmv_type_t vtype, uint64_t seed, void* stream);
spmv_status_t muxi_gen_free(muxi_csr_t* m);
/* x[i] ~ U[-1,1] from SplitMix64(seed, i); FLOAT or DOUBLE. */
spmv_status_t muxi_gen_vector(void* d_x, spmv_type_t type, long n, uint64_t seed, void* stream);

/* ---- features ------------------------------------------------------------
 * Eight matrix features, in this order: rows, nnz, mean row length,
 * row-length variance, max row length, bandwidth (mean over nonzeros of
 * |col - row*numCols/numRows|, divided by numCols), diagonal-block fill
 * (fraction of nonzeros whose row and column fall in the same 32x32
 * diagonal block), and the fraction of empty rows. Computed on the device
 * in FP64; writes out[8] on the host (synchronises the stream). */
spmv_status_t muxi_matrix_features(const muxi_csr_t* A, void* stream, double* out);

/* Writes `bytes` of a scratch buffer (evicts L2 and the Infinity Cache). */
spmv_status_t muxi_flush_caches(void* d_buf, size_t bytes, void* stream);

/* ---- selector MLP ----------------------------------------------------------
 * Fully connected in -> h -> h -> out, ReLU after the two hidden layers, FP32,
 * row-major weights W1[h][in], W2[h][h], W3[out][h]. One parameter buffer
 * holds W1,b1,W2,b2,W3,b3 in that order (muxi_mlp_param_count elements).
 * Forward for a batch X[B][in] -> Y[B][out]; the training step takes
 * Y_target[B][out] and a mask M[B][out] (1 = known target), minimises the
 * masked mean squared error and applies one Adam update (lr, beta1, beta2,
 * eps, step t >= 1). Workspace of muxi_mlp_workspace_bytes(B, ...) bytes. */
long muxi_mlp_param_count(int in, int hidden, int out);
size_t muxi_mlp_workspace_bytes(int batch, int in, int hidden, int out);
spmv_status_t muxi_mlp_forward(const float* d_params, int in, int hidden, int out, const float* d_X, int batch,
                               float* d_Y, void* d_work, void* stream);
/* Gradient of the masked MSE w.r.t. all parameters into d_grad (same layout
 * as d_params); the loss is written to d_loss[0]. */
spmv_status_t muxi_mlp_backward(const float* d_params, int in, int hidden, int out, const float* d_X, int batch,
                                const float* d_T, const float* d_M, float* d_grad, float* d_loss, void* d_work,
                                void* stream);
spmv_status_t muxi_mlp_adam(float* d_params, const float* d_grad, float* d_m, float* d_v, long n, float lr,
                            float beta1, float beta2, float eps, int t, void* stream);

/* ---- learned selection for MUXI_KERNEL_AUTO --------------------------------
 * A library built with trained weights (the package's selector_weights.json)
 * starts with that selector installed ("built-in"). muxi_selector_set
 * installs a trained selector (process-wide, replaces any previous one):
 * params as above with in = 10 or 11, mu/sd the input normalisation, and the
 * candidate each output scores (kernels[n_out], params0[n_out]). The model
 * input is log1p(rows), log1p(nnz), log1p(mean), log1p(sqrt(var)),
 * log1p(max), bandwidth, diagonal-block fill, empty fraction,
 * log1p(max/mean), log1p(nnz) - log1p(rows), and for in = 11 also 1 when x
 * and y are FP64 (0 otherwise). */
spmv_status_t muxi_selector_set(const float* params, int n_in, int hidden, int n_out, const float* mu,
                                const float* sd, const int* kernels, const int* params0);
/* Removes any selector: AUTO falls back to the fixed rule below. */
spmv_status_t muxi_selector_clear(void);
/* Returns to the built-in selector (or to the fixed rule if there is none). */
spmv_status_t muxi_selector_reset(void);
/* 0 = fixed rule, 1 = built-in selector, 2 = a selector set by the caller. */
int muxi_selector_state(void);
/* The (kernel, param0) an AUTO plan would use for A: the selector's argmin
 * over the candidates the sweep admits for A, or, with no selector
 * installed, merge-path (4 items per lane) or column slabs (param0 = 4: the
 * default slab count, 4 items per lane) when the x vector exceeds 4 MiB and
 * rows average under 64 nonzeros. */
spmv_status_t muxi_selector_pick(const muxi_csr_t* A, spmv_type_t xtype, void* stream, int* kernel, int* param0);

/* The schedule tables the autotuner fitted (muxi_spmv_amd/sched_tune.py), compiled into the library: what
 * muxi_batch_plan_create (ipt <= 0) picks for a batch with the given mean nonzeros per row and mean rows per
 * matrix, and what muxi_plan_create_spmm (ipt <= 0) picks for k right-hand sides and the given merge-path
 * length numRows + nnz. Host only. No counterpart in the reference, which hard-codes these (spmv_csr5.cu:953-967). */
spmv_status_t muxi_batch_schedule_for(double mean_row_len, double mean_rows, int* ipt, int* wg_per_cu);
int muxi_spmm_schedule_for(int k, double path_len);

#ifdef __cplusplus
}
#endif

#endif /* MUXI_SPMV_TUNE_H_ */
