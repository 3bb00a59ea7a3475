// Reference-style caller of the COO / BSR / Matrix Market / transpose names in
// the C++ shim: writes a matrix with io::writeMTXFile, reads it back with
// io::readMTXFile and io::readMTXToCOO, converts CSR -> COO -> CSR and
// CSR -> BSR -> CSR on the device, and checks spmv_csr_transpose against a
// dense A^T x. Argument: a scratch directory. Exit code 0 on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "muxi_spmv/muxi_spmv.hpp"

using namespace muxi_spmv;

static int fail(const char* what) {
    std::printf("FAILED: %s (%s)\n", what, muxi_spmv_last_error());
    return 1;
}

template <typename T>
static std::vector<T> to_host(const T* d, int n) {
    std::vector<T> h(n > 0 ? n : 1);
    if (n > 0) (void)hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost);
    return h;
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int R = 57, C = 43;
    // A[i][j] = i + 0.25 * j on a fixed pattern (j = (7 i + 3 k) mod C, k < 1 + i mod 5, distinct and sorted)
    CSRMatrix<double> A;
    std::vector<int> rp(1, 0), ci;
    std::vector<double> va;
    for (int i = 0; i < R; ++i) {
        std::vector<int> cols;
        for (int k = 0; k < 1 + i % 5; ++k) cols.push_back((7 * i + 3 * k * k + k) % C);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        for (int j : cols) {
            ci.push_back(j);
            va.push_back(i + 0.25 * j);
        }
        rp.push_back((int)ci.size());
    }
    A.allocateHost(R, C, (int)ci.size());
    std::copy(rp.begin(), rp.end(), A.rowPtr);
    std::copy(ci.begin(), ci.end(), A.colIdx);
    std::copy(va.begin(), va.end(), A.values);
    const int N = A.nnz;

    const std::string f = dir + "/a.mtx";
    if (io::writeMTXFile<double>(f.c_str(), A) != SPMV_SUCCESS) return fail("writeMTXFile");
    io::MTXHeader hd;
    if (io::readMTXHeader(f.c_str(), hd) != SPMV_SUCCESS || hd.numRows != R || hd.numCols != C || hd.nnz != N)
        return fail("readMTXHeader");
    CSRMatrix<double> B;
    if (io::readMTXFile<double>(f.c_str(), B) != SPMV_SUCCESS || B.nnz != N) return fail("readMTXFile");
    for (int k = 0; k < N; ++k)
        if (B.colIdx[k] != A.colIdx[k] || B.values[k] != A.values[k]) return fail("MTX round trip");
    COOMatrix<double> Ch;
    if (io::readMTXToCOO<double>(f.c_str(), Ch) != SPMV_SUCCESS || Ch.nnz != N) return fail("readMTXToCOO");
    for (int i = 0; i < R; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (Ch.rowIdx[k] != i) return fail("readMTXToCOO rows");

    A.allocateDevice();
    A.copyToDevice();
    (void)hipDeviceSynchronize();

    COOMatrix<double> coo = csr_to_coo<double>(A);
    if (coo.numRows != R) return fail("csr_to_coo");
    CSRMatrix<double> A2 = coo_to_csr<double>(coo);
    if (A2.numRows != R) return fail("coo_to_csr");
    auto rp2 = to_host(A2.d_rowPtr, R + 1);
    auto ci2 = to_host(A2.d_colIdx, N);
    auto va2 = to_host(A2.d_values, N);
    for (int i = 0; i <= R; ++i)
        if (rp2[i] != rp[i]) return fail("coo_to_csr rowPtr");
    for (int k = 0; k < N; ++k)
        if (ci2[k] != ci[k] || va2[k] != va[k]) return fail("coo_to_csr entries");

    BSRMatrix<double, 4> bsr;
    if (csr_to_bsr<double, 4>(A, bsr) != SPMV_SUCCESS) return fail("csr_to_bsr");
    CSRMatrix<double> A3 = bsr_to_csr<double, 4>(bsr);
    if (A3.numRows != R) return fail("bsr_to_csr");
    // the BSR round trip keeps each block's stored zeros: compare as dense
    std::vector<double> D(R * C, 0.0), D3(R * C, 0.0);
    for (int i = 0; i < R; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) D[i * C + ci[k]] = va[k];
    auto rp3 = to_host(A3.d_rowPtr, R + 1);
    auto ci3 = to_host(A3.d_colIdx, A3.nnz);
    auto va3 = to_host(A3.d_values, A3.nnz);
    for (int i = 0; i < R; ++i)
        for (int k = rp3[i]; k < rp3[i + 1]; ++k) D3[i * C + ci3[k]] += va3[k];
    for (int t = 0; t < R * C; ++t)
        if (D[t] != D3[t]) return fail("BSR round trip");

    // y = 2 A^T x + 0.5 y0 against the dense product
    std::vector<double> x(R), y0(C), y(C);
    for (int i = 0; i < R; ++i) x[i] = 1.0 + 0.001 * i;
    for (int j = 0; j < C; ++j) y0[j] = -1.0 + 0.01 * j;
    double *dx, *dy;
    (void)hipMalloc(&dx, sizeof(double) * R);
    (void)hipMalloc(&dy, sizeof(double) * C);
    (void)hipMemcpy(dx, x.data(), sizeof(double) * R, hipMemcpyHostToDevice);
    (void)hipMemcpy(dy, y0.data(), sizeof(double) * C, hipMemcpyHostToDevice);
    spmv_opts_t opts = spmv_default_opts();
    opts.sync = 1;
    if (spmv_csr_transpose<double>(A, dx, dy, 2.0, 0.5, opts) != SPMV_SUCCESS) return fail("spmv_csr_transpose");
    (void)hipMemcpy(y.data(), dy, sizeof(double) * C, hipMemcpyDeviceToHost);
    for (int j = 0; j < C; ++j) {
        double e = 0.5 * y0[j];
        for (int i = 0; i < R; ++i) e += 2.0 * D[i * C + j] * x[i];
        if (std::fabs(y[j] - e) > 1e-12 * (1.0 + std::fabs(e))) return fail("A^T x");
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    std::printf("PASSED (%d x %d, %d nnz, %d blocks of 4 x 4)\n", R, C, N, bsr.numBlocks);
    return 0;
}
