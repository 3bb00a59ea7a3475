// Reference caller compiled against the C++ shim: the flow of
// tests/correctness/test_spmv.cu (diagonal matrix, x[i] = i+1, y == x within
// 1e-5) with cudaStream_t replaced by hipStream_t, plus CSR5 and ELLPACK on
// the same matrix. Exit code 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "muxi_spmv/muxi_spmv.hpp"

using namespace muxi_spmv;

static void generateDiagonalMatrix(int n, CSRMatrix<float>& m) {
    m.allocateHost(n, n, n);
    for (int i = 0; i <= n; i++) m.rowPtr[i] = i;
    for (int i = 0; i < n; i++) {
        m.colIdx[i] = i;
        m.values[i] = 1.0f;
    }
}

static bool verify(const float* y, const float* e, int n, float tol = 1e-5f) {
    for (int i = 0; i < n; i++)
        if (std::fabs(y[i] - e[i]) > tol) {
            std::printf("Mismatch at index %d: got %f, expected %f\n", i, y[i], e[i]);
            return false;
        }
    return true;
}

int main(int argc, char** argv) {
    int n = argc > 1 ? std::atoi(argv[1]) : 1024;
    CSRMatrix<float> A;
    generateDiagonalMatrix(n, A);
    float* hx = (float*)std::malloc(n * sizeof(float));
    float* hy = (float*)std::malloc(n * sizeof(float));
    for (int i = 0; i < n; i++) hx[i] = (float)(i + 1);
    A.allocateDevice();
    A.copyToDevice();
    float *dx, *dy;
    (void)hipMalloc(&dx, n * sizeof(float));
    (void)hipMalloc(&dy, n * sizeof(float));
    (void)hipMemcpy(dx, hx, n * sizeof(float), hipMemcpyHostToDevice);
    (void)hipMemset(dy, 0, n * sizeof(float));

    spmv_handle_t* h;
    if (spmv_create_handle(&h) != SPMV_SUCCESS) return 1;
    std::printf("warpSize %d, CUs %d\n", h->warpSize, h->smCount);
    spmv_opts_t opts = spmv_default_opts();
    opts.sync = 1;
    bool ok = true;

    if (spmv_csr<float>(A, dx, dy, 1.0f, 0.0f, opts) != SPMV_SUCCESS) return 2;
    (void)hipMemcpy(hy, dy, n * sizeof(float), hipMemcpyDeviceToHost);
    ok &= verify(hy, hx, n);

    CSR5Matrix<float> A5;
    if (csr5_preprocess<float>(A, A5, 0, 0) != SPMV_SUCCESS) return 3;
    (void)hipMemset(dy, 0, n * sizeof(float));
    if (spmv_csr5<float>(A5, dx, dy, 1.0f, 0.0f, opts) != SPMV_SUCCESS) return 4;
    (void)hipMemcpy(hy, dy, n * sizeof(float), hipMemcpyDeviceToHost);
    ok &= verify(hy, hx, n);
    ok &= A5.numTiles == (n + 255) / 256 && A5.tile_row_ptr[A5.numTiles - 1] == (A5.numTiles - 1) * 256;

    ELLPACKMatrix<float> E;
    if (csr_to_ellpack<float>(A, E) != SPMV_SUCCESS) return 5;
    (void)hipMemset(dy, 0, n * sizeof(float));
    if (spmv_ellpack<float>(E, dx, dy, 0) != SPMV_SUCCESS) return 6;
    (void)hipDeviceSynchronize();
    (void)hipMemcpy(hy, dy, n * sizeof(float), hipMemcpyDeviceToHost);
    ok &= verify(hy, hx, n) && E.numElements == 1;
    (void)hipFree(E.d_colIdx);
    (void)hipFree(E.d_values);

    (void)hipMemset(dy, 0, n * sizeof(float));
    if (spmv_exec<float>(h, SPMV_FORMAT_CSR, &A, dx, dy) != SPMV_SUCCESS) return 7;
    (void)hipDeviceSynchronize();
    (void)hipMemcpy(hy, dy, n * sizeof(float), hipMemcpyDeviceToHost);
    ok &= verify(hy, hx, n);

    spmv_destroy_handle(h);
    (void)hipFree(dx);
    (void)hipFree(dy);
    std::free(hx);
    std::free(hy);
    std::printf("\nTest: %s\n", ok ? "PASSED" : "FAILED");
    return ok ? 0 : 1;
}
