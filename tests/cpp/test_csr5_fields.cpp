// A C++ caller that uses CSR5Matrix<T> the way the reference's benchmark does
// (src/benchmark/performance_benchmark.cu:383-408: csr5_preprocess timed with
// stream events, then csr5.numTiles and csr5.sigma read as plain fields, then
// spmv_csr5 in a timed loop), with hipStream_t / hipEvent_t in place of the
// CUDA types. It also reads every tile's csr5.tile_row_ptr[t] and
// csr5.tile_nnz_offset[t] and checks them against the definition (the row
// holding nonzero t*sigma, and that nonzero's position inside its row), and
// y against a host CSR loop (exactly: integer-valued data), on a matrix with
// empty rows and a long row.
// Exit code 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "muxi_spmv/muxi_spmv.hpp"

using namespace muxi_spmv;

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 50000;
    CSRMatrix<float> A;
    // row r holds (r * 7) % 13 nonzeros (0 to 12: some rows empty), row n/2 holds 3000; columns spread by a hash
    std::vector<int> len(n);
    long nnz = 0;
    for (int r = 0; r < n; ++r) {
        len[r] = r == n / 2 ? 3000 : (r * 7) % 13;
        nnz += len[r];
    }
    A.allocateHost(n, n, (int)nnz);
    A.rowPtr[0] = 0;
    for (int r = 0; r < n; ++r) A.rowPtr[r + 1] = A.rowPtr[r] + len[r];
    for (int r = 0; r < n; ++r)
        for (int k = A.rowPtr[r]; k < A.rowPtr[r + 1]; ++k) {
            A.colIdx[k] = (int)(((long)(k - A.rowPtr[r]) * 2654435761L + r) % n);
            A.values[k] = (float)((k % 7) - 3) * 0.25f;
        }
    A.allocateDevice();
    A.copyToDevice();
    std::vector<float> hx(n), hy(n), ref(n);
    for (int i = 0; i < n; ++i) hx[i] = (float)(((long)i * 40503) % 9 - 4) * 0.5f;
    for (int r = 0; r < n; ++r) {
        double s = 0;
        for (int k = A.rowPtr[r]; k < A.rowPtr[r + 1]; ++k) s += (double)A.values[k] * hx[A.colIdx[k]];
        ref[r] = (float)s;
    }
    float *d_x, *d_y;
    (void)hipMalloc(&d_x, n * sizeof(float));
    (void)hipMalloc(&d_y, n * sizeof(float));
    (void)hipMemcpy(d_x, hx.data(), n * sizeof(float), hipMemcpyHostToDevice);

    hipStream_t stream;
    (void)hipStreamCreate(&stream);
    hipEvent_t start, stop;
    (void)hipEventCreate(&start);
    (void)hipEventCreate(&stop);

    const int sigma = 256;
    CSR5Matrix<float> csr5;
    (void)hipEventRecord(start, stream);
    spmv_status_t status = csr5_preprocess<float>(A, csr5, sigma, stream);
    (void)hipEventRecord(stop, stream);
    (void)hipEventSynchronize(stop);
    if (status != SPMV_SUCCESS) {
        std::printf("csr5_preprocess: %d (%s)\n", (int)status, muxi_spmv_last_error());
        return 2;
    }
    float convMs = 0.f;
    (void)hipEventElapsedTime(&convMs, start, stop);
    const int numTiles = csr5.numTiles;
    const int tileSize = csr5.sigma;
    bool ok = tileSize == sigma && numTiles == (int)((nnz + sigma - 1) / sigma);
    ok &= csr5.numRows == n && csr5.nnz == (int)nnz && csr5.d_rowPtr == A.d_rowPtr;
    int bad = 0;
    for (int t = 0; t < numTiles; ++t) {
        const int k = t * sigma;
        int lo = 0, hi = n;  // the row r with rowPtr[r] <= k < rowPtr[r + 1] (empty rows skipped)
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (A.rowPtr[mid] <= k) lo = mid; else hi = mid - 1;
        }
        if (csr5.tile_row_ptr[t] != lo || csr5.tile_nnz_offset[t] != k - A.rowPtr[lo]) {
            if (bad++ < 5)
                std::printf("tile %d: row %d offset %d, expected %d %d\n", t, csr5.tile_row_ptr[t],
                            csr5.tile_nnz_offset[t], lo, k - A.rowPtr[lo]);
        }
    }
    ok &= bad == 0;

    spmv_opts_t opts = spmv_default_opts();
    opts.stream = stream;
    opts.sync = 1;
    for (int i = 0; i < 3; ++i)
        if (spmv_csr5<float>(csr5, d_x, d_y, 1.0f, 0.0f, opts) != SPMV_SUCCESS) return 3;
    std::vector<float> times;
    for (int i = 0; i < 10; ++i) {
        (void)hipEventRecord(start, stream);
        if (spmv_csr5<float>(csr5, d_x, d_y, 1.0f, 0.0f, opts) != SPMV_SUCCESS) return 4;
        (void)hipEventRecord(stop, stream);
        (void)hipEventSynchronize(stop);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, start, stop);
        times.push_back(ms);
    }
    (void)hipMemcpy(hy.data(), d_y, n * sizeof(float), hipMemcpyDeviceToHost);
    // values are multiples of 1/4 and x of 1/2 with |sum| < 2^13: every partial sum is exact in FP32, so y must
    // equal the host product exactly, in any summation order
    for (int r = 0; r < n; ++r) {
        if (hy[r] != ref[r]) {
            if (bad++ < 5) std::printf("row %d: %f vs %f\n", r, hy[r], ref[r]);
            ok = false;
        }
    }
    std::printf("rows %d nnz %ld tiles %d sigma %d conversion %.3f ms spmv %.3f ms\n", n, nnz, numTiles, tileSize,
                convMs, times.back());
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    (void)hipStreamDestroy(stream);
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    std::printf("\nTest: %s\n", ok ? "PASSED" : "FAILED");
    return ok ? 0 : 1;
}
