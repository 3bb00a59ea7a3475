// rocSPARSE's CSR SpMV timed on a matrix given as raw arrays, outside any PyTorch process, so that the ROCm
// installation's own librocsparse is the one loaded (a PyTorch process holds its bundled, older copy under the
// same soname). A reference-style caller of the benchmark layer: compareWithRocsparse<float> for each of the
// library's four CSR algorithms, warm, every launch between two HIP events, the median reported; each y checked
// against a host FP64 CSR loop.
//
// usage: rocsparse_timing DIR ROWS COLS NNZ [ITERATIONS]
//   DIR holds rowPtr.i32 (ROWS + 1), colIdx.i32 (NNZ), values.f32 (NNZ), x.f32 (COLS), raw little-endian.
// Prints one JSON line: {"library": path, "iterations": n, "us": {alg: t}, "analysis_ms": {...},
// "rel_l2_vs_fp64": {...}}; exit code 0 when every algorithm ran and every y is within 1e-5 relative L2.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "muxi_spmv/muxi_spmv.hpp"

using namespace muxi_spmv;
using namespace muxi_spmv::benchmark;

template <typename T>
static bool read_raw(const std::string& path, T* dst, size_t n) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const size_t got = std::fread(dst, sizeof(T), n, f);
    std::fclose(f);
    return got == n;
}

static std::string loaded_rocsparse() {
    FILE* f = std::fopen("/proc/self/maps", "r");
    if (!f) return "?";
    char line[4096];
    std::string found = "?";
    while (std::fgets(line, sizeof line, f)) {
        const char* p = std::strstr(line, "librocsparse");
        if (!p) continue;
        const char* s = std::strchr(line, '/');
        if (!s) continue;
        found.assign(s);
        while (!found.empty() && (found.back() == '\n' || found.back() == ' ')) found.pop_back();
        break;
    }
    std::fclose(f);
    return found;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s DIR ROWS COLS NNZ [ITERATIONS]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int rows = std::atoi(argv[2]), cols = std::atoi(argv[3]), nnz = std::atoi(argv[4]);
    const int iters = argc > 5 ? std::atoi(argv[5]) : 200;
    if (rows < 0 || cols < 0 || nnz < 0 || iters <= 0) {
        std::fprintf(stderr, "bad sizes\n");
        return 2;
    }
    CSRMatrix<float> A;
    A.allocateHost(rows, cols, nnz);
    std::vector<float> hx(cols);
    if (!read_raw(dir + "/rowPtr.i32", A.rowPtr, (size_t)rows + 1) || !read_raw(dir + "/colIdx.i32", A.colIdx, nnz) ||
        !read_raw(dir + "/values.f32", A.values, nnz) || !read_raw(dir + "/x.f32", hx.data(), cols)) {
        std::fprintf(stderr, "cannot read the arrays under %s\n", dir.c_str());
        return 2;
    }
    // a well-formed CSR before any device work: rocSPARSE indexes x and y with these arrays unchecked
    bool valid = A.rowPtr[0] == 0 && A.rowPtr[rows] == nnz;
    for (int r = 0; valid && r < rows; ++r) valid = A.rowPtr[r] <= A.rowPtr[r + 1];
    for (int k = 0; valid && k < nnz; ++k) valid = A.colIdx[k] >= 0 && A.colIdx[k] < cols;
    if (!valid) {
        std::fprintf(stderr, "not a CSR matrix of %d x %d with %d nonzeros\n", rows, cols, nnz);
        return 2;
    }
    A.allocateDevice();
    A.copyToDevice();
    float *d_x = nullptr, *d_y = nullptr;
    if (hipMalloc(&d_x, sizeof(float) * (cols > 0 ? cols : 1)) != hipSuccess ||
        hipMalloc(&d_y, sizeof(float) * (rows > 0 ? rows : 1)) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 2;
    }
    (void)hipMemcpy(d_x, hx.data(), sizeof(float) * cols, hipMemcpyHostToDevice);

    std::vector<double> y64(rows);
    double n64 = 0.0;
    for (int r = 0; r < rows; ++r) {
        double s = 0.0;
        for (int k = A.rowPtr[r]; k < A.rowPtr[r + 1]; ++k) s += (double)A.values[k] * (double)hx[A.colIdx[k]];
        y64[r] = s;
        n64 += s * s;
    }
    n64 = std::sqrt(n64);

    BenchmarkConfig config;
    config.warmupIterations = 10;
    config.measureIterations = iters;
    config.syncAfterEach = false;  // launches back to back, as a solver's loop issues them
    const char* names[4] = {"adaptive", "rowsplit", "lrb", "nnzsplit"};
    std::string us = "{", an = "{", err = "{";
    auto add = [](std::string& obj, const char* name, const char* fmt, double v) {
        char buf[160];
        std::snprintf(buf, sizeof buf, fmt, v);
        obj += std::string(obj.size() > 1 ? ", \"" : "\"") + name + "\": " + buf;
    };
    bool ok = true;
    std::vector<float> hy(rows);
    for (int alg = 0; alg < 4; ++alg) {
        double t = 0.0, g = 0.0, a = 0.0;
        const spmv_status_t st = compareWithRocsparse<float>(A, d_x, d_y, config, t, g, alg, &a);
        if (st != SPMV_SUCCESS) {
            std::fprintf(stderr, "%s: %d %s\n", names[alg], (int)st, ::muxi_spmv_last_error());
            us += std::string(us.size() > 1 ? ", \"" : "\"") + names[alg] + "\": null";
            ok = false;
            continue;
        }
        (void)hipMemcpy(hy.data(), d_y, sizeof(float) * rows, hipMemcpyDeviceToHost);
        double d = 0.0;
        for (int r = 0; r < rows; ++r) d += ((double)hy[r] - y64[r]) * ((double)hy[r] - y64[r]);
        const double rel = n64 > 0.0 ? std::sqrt(d) / n64 : std::sqrt(d);
        ok = ok && rel <= 1e-5;
        add(us, names[alg], "%.3f", t * 1e3);
        add(an, names[alg], "%.3f", a);
        add(err, names[alg], "%.3e", rel);
    }
    std::printf("{\"library\": \"%s\", \"iterations\": %d, \"us\": %s}, \"analysis_ms\": %s}, "
                "\"rel_l2_vs_fp64\": %s}}\n",
                loaded_rocsparse().c_str(), iters, us.c_str(), an.c_str(), err.c_str());
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    return ok ? 0 : 1;
}
