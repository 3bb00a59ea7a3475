// A C++ caller of the benchmark layer written as the reference's benchmark driver uses it
// (src/benchmark/performance_benchmark.h): a BenchmarkConfig with its defaults, measureCSRPerformance<float> and
// <double> through a handle, measureCSR5Performance<float> on a fresh CSR5Matrix, printPerformanceResult and
// writeResultsJSON, with hipStream_t in place of cudaStream_t. It checks the fields a reference caller reads:
// the median between min and max, GFLOP/s and bandwidth at the median by the reference's formulas, the CSR5
// conversion time and tile count, and that y equals a host CSR loop (integer-valued data, exact). compareWithRocsparse<T>
// runs rocSPARSE's four CSR algorithms on the same data (y NaN beforehand), each y exact as well.
// argv[1]: the JSON file to write. Exit code 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "muxi_spmv/muxi_spmv.hpp"

using namespace muxi_spmv;
using namespace muxi_spmv::benchmark;

static int fails = 0;
#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            std::printf("FAILED: %s (line %d)\n", #c, __LINE__);       \
            ++fails;                                                   \
        }                                                              \
    } while (0)

template <typename T>
static void fill(CSRMatrix<T>& A, int n) {
    long nnz = 0;
    for (int r = 0; r < n; ++r) nnz += 1 + (r * 5) % 11;
    A.allocateHost(n, n, (int)nnz);
    A.rowPtr[0] = 0;
    for (int r = 0; r < n; ++r) A.rowPtr[r + 1] = A.rowPtr[r] + 1 + (r * 5) % 11;
    for (int r = 0; r < n; ++r)
        for (int k = A.rowPtr[r]; k < A.rowPtr[r + 1]; ++k) {
            A.colIdx[k] = (int)(((long)(k - A.rowPtr[r]) * 40503L + 7L * r) % n);
            A.values[k] = (T)((k % 5) - 2);
        }
    A.allocateDevice();
    A.copyToDevice();
}

static void check_times(const PerformanceResult& r, int nnz, int rows, int cols, int fb) {
    CHECK(r.iterations == 10);
    CHECK(r.minTimeMs > 0.0 && r.minTimeMs <= r.timeMs && r.timeMs <= r.maxTimeMs);
    CHECK(r.minTimeMs <= r.avgTimeMs && r.avgTimeMs <= r.maxTimeMs && r.stdDevMs >= 0.0);
    CHECK(std::fabs(r.gflops - 2.0 * nnz / r.timeMs / 1e6) <= 1e-9 * r.gflops);
    const double bytes = (double)nnz * (fb + 4) + 4.0 * (rows + 1) + (double)fb * (cols + rows);
    CHECK(std::fabs(r.bandwidthGBs - bytes / r.timeMs / 1e6) <= 1e-9 * r.bandwidthGBs);
    CHECK(r.peakBandwidthGBs > 0.0);
    CHECK(std::fabs(r.bandwidthUtilization - r.bandwidthGBs / r.peakBandwidthGBs * 100.0) <= 1e-9 * r.bandwidthUtilization);
    CHECK(r.correctnessPassed);
    CHECK(r.numRows == rows && r.numCols == cols && r.nnz == nnz && r.warpSize == 64);
}

template <typename T>
static bool y_exact(const CSRMatrix<T>& A, const std::vector<T>& hx, const T* d_y) {
    std::vector<T> hy(A.numRows);
    (void)hipMemcpy(hy.data(), d_y, sizeof(T) * A.numRows, hipMemcpyDeviceToHost);
    for (int r = 0; r < A.numRows; ++r) {
        double s = 0.0;
        for (int k = A.rowPtr[r]; k < A.rowPtr[r + 1]; ++k) s += (double)A.values[k] * (double)hx[A.colIdx[k]];
        if ((double)hy[r] != s) return false;
    }
    return true;
}

template <typename T>
static PerformanceResult run_csr(int n, spmv_handle_t* h) {
    CSRMatrix<T> A;
    fill(A, n);
    std::vector<T> hx(n);
    for (int i = 0; i < n; ++i) hx[i] = (T)((i % 9) - 4);
    T *d_x = nullptr, *d_y = nullptr;
    (void)hipMalloc(&d_x, sizeof(T) * n);
    (void)hipMalloc(&d_y, sizeof(T) * n);
    (void)hipMemcpy(d_x, hx.data(), sizeof(T) * n, hipMemcpyHostToDevice);
    BenchmarkConfig config;  // 3 warmup, 10 measured, correctness check, sync after each
    PerformanceResult r{};
    CHECK(measureCSRPerformance<T>(A, d_x, d_y, config, h, r) == SPMV_SUCCESS);
    check_times(r, A.nnz, n, n, (int)sizeof(T));
    CHECK(!r.useCSR5);
    CHECK(y_exact(A, hx, d_y));
    double cs = 0.0, cg = 0.0;
    CHECK(compareWithCusparse<T>(A, d_x, d_y, config, cs, cg) == SPMV_ERROR_UNSUPPORTED_FORMAT);
    for (int alg = 0; alg < 4; ++alg) {  // the vendor library: adaptive, rowsplit, LRB, nnzsplit
        (void)hipMemset(d_y, 0xff, sizeof(T) * n);  // NaN: the comparison clears y itself
        const spmv_status_t rs = compareWithRocsparse<T>(A, d_x, d_y, config, cs, cg, alg);
        if (rs != SPMV_SUCCESS) std::printf("compareWithRocsparse(%d): %d %s\n", alg, (int)rs, ::muxi_spmv_last_error());
        CHECK(rs == SPMV_SUCCESS);
        CHECK(cs > 0.0 && std::fabs(cg - 2.0 * A.nnz / cs / 1e6) <= 1e-9 * cg);
        CHECK(y_exact(A, hx, d_y));
    }
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    return r;
}

int main(int argc, char** argv) {
    const char* json = argc > 1 ? argv[1] : "spmv_results_cpp.json";
    const int n = 60000;
    spmv_handle_t* h = nullptr;
    CHECK(spmv_create_handle(&h) == SPMV_SUCCESS);
    std::vector<PerformanceResult> results;
    results.push_back(run_csr<float>(n, h));
    results.push_back(run_csr<double>(n, h));

    // CSR5: the conversion timed with the SpMV
    CSRMatrix<float> A;
    fill(A, n);
    std::vector<float> hx(n);
    for (int i = 0; i < n; ++i) hx[i] = (float)((i % 9) - 4);
    float *d_x = nullptr, *d_y = nullptr;
    (void)hipMalloc(&d_x, sizeof(float) * n);
    (void)hipMalloc(&d_y, sizeof(float) * n);
    (void)hipMemcpy(d_x, hx.data(), sizeof(float) * n, hipMemcpyHostToDevice);
    hipStream_t stream;
    (void)hipStreamCreate(&stream);
    BenchmarkConfig config;
    config.stream = stream;
    CSR5Matrix<float> csr5;
    PerformanceResult r5{};
    CHECK(measureCSR5Performance<float>(A, csr5, d_x, d_y, config, h, r5, 256) == SPMV_SUCCESS);
    check_times(r5, A.nnz, n, n, 4);
    CHECK(r5.useCSR5 && r5.tileSize == 256 && r5.numTiles == csr5.numTiles);
    CHECK(r5.numTiles == (A.nnz + 255) / 256);
    CHECK(r5.conversionTimeMs > 0.0);
    CHECK(std::fabs(r5.totalTimeMs - (r5.timeMs + r5.conversionTimeMs)) <= 1e-12 * r5.totalTimeMs);
    CHECK(std::fabs(r5.conversionOverheadPercent - r5.conversionTimeMs / r5.timeMs * 100.0) <=
          1e-9 * r5.conversionOverheadPercent);
    CHECK(y_exact(A, hx, d_y));
    results.push_back(r5);
    printPerformanceResult(r5);
    CHECK(writeResultsJSON(results, json) == SPMV_SUCCESS);

    // a non-finite y fails the reference's correctness check
    (void)hipMemset(d_x, 0xff, sizeof(float) * n);  // every x a NaN
    PerformanceResult bad{};
    CHECK(measureCSRPerformance<float>(A, d_x, d_y, BenchmarkConfig(), h, bad) == SPMV_SUCCESS);
    CHECK(!bad.correctnessPassed);

    (void)hipFree(d_x);
    (void)hipFree(d_y);
    (void)hipStreamDestroy(stream);
    spmv_destroy_handle(h);
    std::printf("%s\n", fails == 0 ? "PASSED" : "FAILED");
    return fails == 0 ? 0 : 1;
}
