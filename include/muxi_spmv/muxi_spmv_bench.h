/* muxi_spmv benchmark layer: the reference's performance measurement API
 * (src/benchmark/performance_benchmark.h:21-71, 79-111; performance_benchmark.cu:56-169, 272-335, 348-479)
 * behind the C ABI, timed with HIP events on MI355X. The C++ names (PerformanceResult, BenchmarkConfig,
 * measureCSRPerformance<T>, measureCSR5Performance<T>, writeResultsJSON, printPerformanceResult) are in
 * muxi_spmv.hpp, namespace muxi_spmv::benchmark. */
#ifndef MUXI_SPMV_BENCH_H_
#define MUXI_SPMV_BENCH_H_

#include "muxi_spmv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PerformanceResult (performance_benchmark.h:21-53), same fields in the same order. */
typedef struct {
    double timeMs;               /* median of the measured iterations */
    double gflops;               /* 2 * nnz / timeMs / 1e6 */
    double bandwidthGBs;         /* muxi_calculate_bandwidth at timeMs */
    double bandwidthUtilization; /* bandwidthGBs / peakBandwidthGBs * 100 */
    int iterations;
    bool correctnessPassed;
    double minTimeMs;
    double maxTimeMs;
    double avgTimeMs;
    double stdDevMs; /* population standard deviation */
    int numRows;
    int numCols;
    int nnz;
    double sparsity; /* nnz / (numRows * numCols) */
    int deviceId;
    int warpSize;
    double peakBandwidthGBs; /* 2 * memory clock (kHz) * bus width (bytes) / 1e6, from the device attributes */
    bool useCSR5;
    double conversionTimeMs; /* CSR -> CSR5 preprocessing, HIP events */
    double totalTimeMs;      /* timeMs + conversionTimeMs */
    int numTiles;
    int tileSize;
    double conversionOverheadPercent; /* conversionTimeMs / timeMs * 100 */
} muxi_perf_result_t;

/* BenchmarkConfig (performance_benchmark.h:58-71), same fields in the same order; `stream` is the hipStream_t
 * (NULL = null stream). compareCusparse is kept for the layout and must be false. */
typedef struct {
    int warmupIterations;
    int measureIterations;
    bool checkCorrectness;
    bool syncAfterEach;
    void* stream;
    spmv_type_t floatType; /* FLOAT or DOUBLE: which entry point (f32 / f64) is timed */
    bool compareCusparse;
} muxi_bench_config_t;

/* The reference's defaults: 3 warmup and 10 measured iterations, correctness check and a sync after each,
 * null stream, FLOAT. */
muxi_bench_config_t muxi_bench_default_config(void);

/* calculateGFlops (performance_benchmark.h:79-85): 2 * nnz / timeMs / 1e6. */
double muxi_calculate_gflops(int nnz, double timeMs);
/* calculateBandwidth (performance_benchmark.h:96-111): (nnz * (floatBytes + 4) + 4 * (numRows + 1) +
 * floatBytes * (numCols + numRows)) / timeMs / 1e6, counted in 64-bit (the reference's int product wraps
 * past 2^31 bytes). */
double muxi_calculate_bandwidth(int nnz, int numRows, int numCols, double timeMs, int floatBytes);

/* measureCSRPerformance<T> (performance_benchmark.cu:56-169): config->warmupIterations untimed calls of the
 * dispatching SpMV (muxi_spmv_csr_f32 / _f64 by config->floatType, alpha 1, beta 0, opts.sync =
 * syncAfterEach, on config->stream), then config->measureIterations calls each between two HIP events; timeMs
 * is the median, with min / max / mean / population standard deviation beside it. correctnessPassed: every
 * y is finite (the reference's check), true when checkCorrectness is off. handle gives deviceId and warpSize
 * (NULL: the current device, 64). */
spmv_status_t muxi_measure_csr_performance(const muxi_csr_t* A, const void* d_x, void* d_y,
                                           const muxi_bench_config_t* config, const muxi_spmv_handle_t* handle,
                                           muxi_perf_result_t* result);

/* measureCSR5Performance<T> (performance_benchmark.cu:348-479): times muxi_csr5_preprocess(csr, csr5, sigma)
 * with HIP events (conversionTimeMs), then the CSR5 SpMV (muxi_spmv_csr5_f32 / _f64) as above; fills
 * numTiles, tileSize, totalTimeMs and conversionOverheadPercent. csr5 must be empty (it receives the tiles). */
spmv_status_t muxi_measure_csr5_performance(const muxi_csr_t* csr, muxi_csr5_t* csr5, const void* d_x, void* d_y,
                                            const muxi_bench_config_t* config, const muxi_spmv_handle_t* handle,
                                            muxi_perf_result_t* result, int sigma);

/* compareWithCusparse (performance_benchmark.h:124-151, performance_benchmark.cu:173-270, compiled out in the
 * reference) on this device: the vendor library's CSR SpMV (rocSPARSE rocsparse_spmv by stages, loaded at run time
 * from librocsparse.so.1; alg 0 adaptive, 1 rowsplit, 2 LRB, 3 nnzsplit) on the same matrix and vectors, y cleared
 * first, its preprocess stage timed once (*analysisMs, may be NULL) and its compute stage timed as
 * muxi_measure_csr_performance times the dispatching SpMV (warmup, then the median of config->measureIterations,
 * HIP events on config->stream): *timeMs and *gflops (2 * nnz / timeMs / 1e6). A speed reference and a cross-check
 * of y; the library's own SpMV never calls it. SPMV_ERROR_UNSUPPORTED_FORMAT for an alg outside 0..3, when
 * rocSPARSE cannot be loaded, or when the loaded rocSPARSE rejects the algorithm; the argument errors of
 * muxi_measure_csr_performance otherwise (timeMs or gflops NULL: SPMV_ERROR_INVALID_VECTOR). */
spmv_status_t muxi_compare_with_rocsparse(const muxi_csr_t* A, const void* d_x, void* d_y,
                                          const muxi_bench_config_t* config, int alg, double* timeMs,
                                          double* gflops, double* analysisMs);

/* writeResultsJSON (performance_benchmark.cu:298-335): {"results": [{matrix, device, performance,
 * correctnessPassed}, ...]} with the reference's keys and number formats. */
spmv_status_t muxi_write_results_json(const muxi_perf_result_t* results, int count, const char* filename);

/* printPerformanceResult (performance_benchmark.cu:272-296), to stdout. */
void muxi_print_performance_result(const muxi_perf_result_t* result);

#ifdef __cplusplus
}
#endif

#endif /* MUXI_SPMV_BENCH_H_ */
