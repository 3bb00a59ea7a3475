// Benchmark layer behind the C ABI: the reference's measureCSRPerformance / measureCSR5Performance timing
// protocol (src/benchmark/performance_benchmark.cu:56-169, 348-479), its metrics (performance_benchmark.h:79-111),
// results file (:298-335) and printout (:272-296), timed with HIP events on the configured stream.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

// rocSPARSE's types and prototypes only: the library is opened at run time by the comparison that uses it (a
// process that has loaded torch already holds a copy of it, under the same name, which dlopen then returns), so
// libmuxi_spmv_amd.so itself depends on nothing beyond the HIP runtime.
#include <rocsparse/rocsparse.h>

#include "kernels.hpp"
#include "muxi_spmv/muxi_spmv_bench.h"

namespace muxi {
namespace {

// Two HIP events, destroyed on every return path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// minTimeMs / maxTimeMs / avgTimeMs / stdDevMs as calculateStats (performance_benchmark.cu:37-54; population
// standard deviation), and timeMs = the median as the reference takes it (element size / 2 of the sorted times).
void fill_stats(std::vector<double> t, muxi_perf_result_t* r) {
    std::sort(t.begin(), t.end());
    double sum = 0.0;
    for (double v : t) sum += v;
    const double avg = sum / (double)t.size();
    double var = 0.0;
    for (double v : t) var += (v - avg) * (v - avg);
    r->minTimeMs = t.front();
    r->maxTimeMs = t.back();
    r->avgTimeMs = avg;
    r->stdDevMs = std::sqrt(var / (double)t.size());
    r->timeMs = t[t.size() / 2];
}

// The matrix, device and peak-bandwidth fields every measurement shares.
spmv_status_t fill_header(const muxi_csr_t* A, const muxi_bench_config_t* cfg, const muxi_spmv_handle_t* h,
                          muxi_perf_result_t* r) {
    *r = muxi_perf_result_t{};
    r->numRows = A->numRows;
    r->numCols = A->numCols;
    r->nnz = A->nnz;
    r->sparsity = (A->numRows > 0 && A->numCols > 0) ? (double)A->nnz / ((double)A->numRows * (double)A->numCols)
                                                     : 0.0;
    int dev = 0;
    MUXI_HIP(hipGetDevice(&dev));
    r->deviceId = h ? h->deviceId : dev;
    r->warpSize = h ? h->warpSize : 64;
    r->iterations = cfg->measureIterations;
    int clk = 0, bus = 0;
    MUXI_HIP(hipDeviceGetAttribute(&clk, hipDeviceAttributeMemoryClockRate, r->deviceId));
    MUXI_HIP(hipDeviceGetAttribute(&bus, hipDeviceAttributeMemoryBusWidth, r->deviceId));
    r->peakBandwidthGBs = 2.0 * (double)clk * (double)(bus / 8) / 1e6;
    return SPMV_SUCCESS;
}

spmv_status_t check_config(const muxi_bench_config_t* cfg) {
    if (!cfg || cfg->measureIterations <= 0 || cfg->warmupIterations < 0) {
        set_error("benchmark: config is NULL or measureIterations <= 0 or warmupIterations < 0");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    if (cfg->compareCusparse) {
        set_error("benchmark: compareCusparse has no counterpart on this device");
        return SPMV_ERROR_UNSUPPORTED_FORMAT;
    }
    if (!(cfg->floatType == SPMV_TYPE_FLOAT || cfg->floatType == SPMV_TYPE_DOUBLE)) {
        set_error("benchmark: floatType must be FLOAT or DOUBLE");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    return SPMV_SUCCESS;
}

// Warmup, then measured iterations of run() each between two events on stream s; the median and its spread go
// to r, and the GFLOP/s and bandwidth at the median.
template <typename F>
spmv_status_t time_iterations(const muxi_bench_config_t* cfg, hipStream_t s, const muxi_csr_t* A,
                              muxi_perf_result_t* r, F&& run) {
    for (int i = 0; i < cfg->warmupIterations; ++i) {
        const spmv_status_t st = run();
        if (st != SPMV_SUCCESS) return st;
    }
    MUXI_HIP(hipDeviceSynchronize());
    EventPair ev;
    MUXI_HIP(hipEventCreate(&ev.a));
    MUXI_HIP(hipEventCreate(&ev.b));
    std::vector<double> times;
    times.reserve((size_t)cfg->measureIterations);
    for (int i = 0; i < cfg->measureIterations; ++i) {
        MUXI_HIP(hipEventRecord(ev.a, s));
        const spmv_status_t st = run();
        if (st != SPMV_SUCCESS) return st;
        MUXI_HIP(hipEventRecord(ev.b, s));
        MUXI_HIP(hipEventSynchronize(ev.b));
        float ms = 0.f;
        MUXI_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
        times.push_back((double)ms);
    }
    fill_stats(times, r);
    const int fb = cfg->floatType == SPMV_TYPE_DOUBLE ? 8 : 4;
    r->gflops = muxi_calculate_gflops(A->nnz, r->timeMs);
    r->bandwidthGBs = muxi_calculate_bandwidth(A->nnz, A->numRows, A->numCols, r->timeMs, fb);
    r->bandwidthUtilization = r->peakBandwidthGBs > 0.0 ? r->bandwidthGBs / r->peakBandwidthGBs * 100.0 : 0.0;
    return SPMV_SUCCESS;
}

// The reference's check: every y is finite.
spmv_status_t check_finite(const muxi_bench_config_t* cfg, const void* d_y, int rows, muxi_perf_result_t* r) {
    r->correctnessPassed = true;
    if (!cfg->checkCorrectness || rows <= 0) return SPMV_SUCCESS;
    const size_t es = cfg->floatType == SPMV_TYPE_DOUBLE ? 8 : 4;
    std::vector<unsigned char> h((size_t)rows * es);
    MUXI_HIP(hipMemcpy(h.data(), d_y, h.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < rows; ++i) {
        const double v = es == 8 ? ((const double*)h.data())[i] : (double)((const float*)h.data())[i];
        if (!std::isfinite(v)) {
            r->correctnessPassed = false;
            break;
        }
    }
    return SPMV_SUCCESS;
}

// The rocSPARSE entry points the comparison calls, looked up once. The SpMV is the stage-based rocsparse_spmv
// (buffer size, preprocess, compute), which every rocSPARSE of ROCm 6 and 7 exports (ROCm 7 marks it deprecated
// in favour of rocsparse_v2_spmv).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
using RsSpmv = decltype(&rocsparse_spmv);
#pragma clang diagnostic pop
struct RsApi {
    decltype(&rocsparse_create_handle) create_handle = nullptr;
    decltype(&rocsparse_destroy_handle) destroy_handle = nullptr;
    decltype(&rocsparse_set_stream) set_stream = nullptr;
    decltype(&rocsparse_create_csr_descr) create_csr_descr = nullptr;
    decltype(&rocsparse_destroy_spmat_descr) destroy_spmat_descr = nullptr;
    decltype(&rocsparse_create_dnvec_descr) create_dnvec_descr = nullptr;
    decltype(&rocsparse_destroy_dnvec_descr) destroy_dnvec_descr = nullptr;
    RsSpmv spmv = nullptr;
    bool ok = false;
    std::string why;
    std::string path;  // the file the symbols came from
};

const RsApi& rs_api() {
    static RsApi api;
    static std::once_flag once;
    std::call_once(once, [] {
        // the copy already in the process (or on the loader's path), else the one in the ROCm installation
        void* h = dlopen("librocsparse.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            const char* root = std::getenv("ROCM_PATH");
            const std::string path = std::string(root && *root ? root : "/opt/rocm") + "/lib/librocsparse.so.1";
            h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (!h) {
                const char* e = dlerror();
                api.why = "cannot open librocsparse.so.1 (" + path + "): " + (e ? e : "");
                return;
            }
        }
#define MUXI_RS_SYM(field, name)                                  \
    api.field = reinterpret_cast<decltype(api.field)>(dlsym(h, name)); \
    if (!api.field) {                                               \
        api.why = std::string("librocsparse.so.1 lacks ") + name;   \
        return;                                                     \
    }
        MUXI_RS_SYM(create_handle, "rocsparse_create_handle")
        MUXI_RS_SYM(destroy_handle, "rocsparse_destroy_handle")
        MUXI_RS_SYM(set_stream, "rocsparse_set_stream")
        MUXI_RS_SYM(create_csr_descr, "rocsparse_create_csr_descr")
        MUXI_RS_SYM(destroy_spmat_descr, "rocsparse_destroy_spmat_descr")
        MUXI_RS_SYM(create_dnvec_descr, "rocsparse_create_dnvec_descr")
        MUXI_RS_SYM(destroy_dnvec_descr, "rocsparse_destroy_dnvec_descr")
        MUXI_RS_SYM(spmv, "rocsparse_spmv")
#undef MUXI_RS_SYM
        Dl_info info{};
        api.path = dladdr(reinterpret_cast<void*>(api.spmv), &info) && info.dli_fname ? info.dli_fname : "?";
        api.ok = true;
    });
    return api;
}

// rocSPARSE objects of one comparison, released on every return path.
struct Rocsparse {
    const RsApi& f;
    explicit Rocsparse(const RsApi& api) : f(api) {}
    rocsparse_handle h = nullptr;
    rocsparse_spmat_descr mat = nullptr;
    rocsparse_dnvec_descr x = nullptr, y = nullptr;
    void* buf = nullptr;
    ~Rocsparse() {
        (void)hipFree(buf);
        if (x) f.destroy_dnvec_descr(x);
        if (y) f.destroy_dnvec_descr(y);
        if (mat) f.destroy_spmat_descr(mat);
        if (h) f.destroy_handle(h);
    }
};

#define MUXI_RS(call)                                                                          \
    do {                                                                                       \
        const rocsparse_status rs_ = (call);                                                   \
        if (rs_ != rocsparse_status_success) {                                                 \
            set_error(std::string(#call) + ": rocsparse status " + std::to_string((int)rs_)); \
            return SPMV_ERROR_INTERNAL;                                                        \
        }                                                                                      \
    } while (0)


spmv_opts_t opts_for(const muxi_bench_config_t* cfg) {
    spmv_opts_t o = muxi_spmv_default_opts();
    o.sync = cfg->syncAfterEach ? 1 : 0;
    o.stream = cfg->stream;
    return o;
}

}  // namespace
}  // namespace muxi

using namespace muxi;

extern "C" muxi_bench_config_t muxi_bench_default_config(void) {
    muxi_bench_config_t c{};
    c.warmupIterations = 3;
    c.measureIterations = 10;
    c.checkCorrectness = true;
    c.syncAfterEach = true;
    c.stream = nullptr;
    c.floatType = SPMV_TYPE_FLOAT;
    c.compareCusparse = false;
    return c;
}

extern "C" double muxi_calculate_gflops(int nnz, double timeMs) { return (2.0 * (double)nnz) / timeMs / 1e6; }

extern "C" double muxi_calculate_bandwidth(int nnz, int numRows, int numCols, double timeMs, int floatBytes) {
    const long long bytes = (long long)nnz * floatBytes + (long long)nnz * 4 + ((long long)numRows + 1) * 4 +
                            (long long)numCols * floatBytes + (long long)numRows * floatBytes;
    return (double)bytes / timeMs / 1e6;
}

extern "C" spmv_status_t muxi_measure_csr_performance(const muxi_csr_t* A, const void* d_x, void* d_y,
                                                      const muxi_bench_config_t* config,
                                                      const muxi_spmv_handle_t* handle, muxi_perf_result_t* result) {
    return no_throw("muxi_measure_csr_performance", [&]() -> spmv_status_t {
        if (!A || !result) {
            set_error("benchmark: matrix or result is NULL");
            return SPMV_ERROR_INVALID_MATRIX;
        }
        if (!d_x || !d_y) {
            set_error("benchmark: x or y is NULL");
            return SPMV_ERROR_INVALID_VECTOR;
        }
        spmv_status_t st = check_config(config);
        if (st != SPMV_SUCCESS) return st;
        st = fill_header(A, config, handle, result);
        if (st != SPMV_SUCCESS) return st;
        const spmv_opts_t o = opts_for(config);
        hipStream_t s = (hipStream_t)config->stream;
        auto run = [&]() -> spmv_status_t {
            if (config->floatType == SPMV_TYPE_DOUBLE)
                return muxi_spmv_csr_f64(A, (const double*)d_x, (double*)d_y, 1.0, 0.0, &o);
            return muxi_spmv_csr_f32(A, (const float*)d_x, (float*)d_y, 1.f, 0.f, &o);
        };
        st = time_iterations(config, s, A, result, run);
        if (st != SPMV_SUCCESS) return st;
        return check_finite(config, d_y, A->numRows, result);
    });
}

extern "C" spmv_status_t muxi_measure_csr5_performance(const muxi_csr_t* csr, muxi_csr5_t* csr5, const void* d_x,
                                                       void* d_y, const muxi_bench_config_t* config,
                                                       const muxi_spmv_handle_t* handle, muxi_perf_result_t* result,
                                                       int sigma) {
    return no_throw("muxi_measure_csr5_performance", [&]() -> spmv_status_t {
        if (!csr || !csr5 || !result) {
            set_error("benchmark: matrix, CSR5 matrix or result is NULL");
            return SPMV_ERROR_INVALID_MATRIX;
        }
        if (!d_x || !d_y) {
            set_error("benchmark: x or y is NULL");
            return SPMV_ERROR_INVALID_VECTOR;
        }
        spmv_status_t st = check_config(config);
        if (st != SPMV_SUCCESS) return st;
        st = fill_header(csr, config, handle, result);
        if (st != SPMV_SUCCESS) return st;
        result->useCSR5 = true;
        result->tileSize = sigma;
        hipStream_t s = (hipStream_t)config->stream;
        {  // the conversion, between two events
            EventPair ev;
            MUXI_HIP(hipEventCreate(&ev.a));
            MUXI_HIP(hipEventCreate(&ev.b));
            MUXI_HIP(hipEventRecord(ev.a, s));
            st = muxi_csr5_preprocess(csr, csr5, sigma, s);
            MUXI_HIP(hipEventRecord(ev.b, s));
            MUXI_HIP(hipEventSynchronize(ev.b));
            if (st != SPMV_SUCCESS) return st;
            float ms = 0.f;
            MUXI_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
            result->conversionTimeMs = (double)ms;
        }
        result->numTiles = csr5->numTiles;
        result->tileSize = csr5->sigma;
        const spmv_opts_t o = opts_for(config);
        auto run = [&]() -> spmv_status_t {
            if (config->floatType == SPMV_TYPE_DOUBLE)
                return muxi_spmv_csr5_f64(csr5, (const double*)d_x, (double*)d_y, 1.0, 0.0, &o);
            return muxi_spmv_csr5_f32(csr5, (const float*)d_x, (float*)d_y, 1.f, 0.f, &o);
        };
        st = time_iterations(config, s, csr, result, run);
        if (st != SPMV_SUCCESS) return st;
        result->totalTimeMs = result->timeMs + result->conversionTimeMs;
        result->conversionOverheadPercent = result->conversionTimeMs / result->timeMs * 100.0;
        return check_finite(config, d_y, csr->numRows, result);
    });
}

extern "C" spmv_status_t muxi_compare_with_rocsparse(const muxi_csr_t* A, const void* d_x, void* d_y,
                                                     const muxi_bench_config_t* config, int alg, double* timeMs,
                                                     double* gflops, double* analysisMs) {
    return no_throw("muxi_compare_with_rocsparse", [&]() -> spmv_status_t {
        if (!A || A->numRows < 0 || A->numCols < 0 || A->nnz < 0 || (A->numRows > 0 && !A->d_rowPtr) ||
            (A->nnz > 0 && (!A->d_colIdx || !A->d_values))) {
            set_error("compare_with_rocsparse: invalid matrix");
            return SPMV_ERROR_INVALID_MATRIX;
        }
        if (!d_x || !d_y || !timeMs || !gflops) {
            set_error("compare_with_rocsparse: x, y, timeMs or gflops is NULL");
            return SPMV_ERROR_INVALID_VECTOR;
        }
        spmv_status_t st = check_config(config);
        if (st != SPMV_SUCCESS) return st;
        static const rocsparse_spmv_alg algs[] = {rocsparse_spmv_alg_csr_adaptive, rocsparse_spmv_alg_csr_rowsplit,
                                                  rocsparse_spmv_alg_csr_lrb, rocsparse_spmv_alg_csr_nnzsplit};
        if (alg < 0 || alg > 3) {
            set_error("compare_with_rocsparse: alg must be 0 (adaptive), 1 (rowsplit), 2 (LRB) or 3 (nnzsplit)");
            return SPMV_ERROR_UNSUPPORTED_FORMAT;
        }
        const RsApi& f = rs_api();
        if (!f.ok) {
            set_error("compare_with_rocsparse: " + f.why);
            return SPMV_ERROR_UNSUPPORTED_FORMAT;
        }
        const bool f64 = config->floatType == SPMV_TYPE_DOUBLE;
        const rocsparse_datatype dt = f64 ? rocsparse_datatype_f64_r : rocsparse_datatype_f32_r;
        hipStream_t s = (hipStream_t)config->stream;
        Rocsparse rs(f);
        MUXI_RS(f.create_handle(&rs.h));
        MUXI_RS(f.set_stream(rs.h, s));
        MUXI_RS(f.create_csr_descr(&rs.mat, A->numRows, A->numCols, A->nnz, A->d_rowPtr, A->d_colIdx, A->d_values,
                                   rocsparse_indextype_i32, rocsparse_indextype_i32, rocsparse_index_base_zero, dt));
        MUXI_RS(f.create_dnvec_descr(&rs.x, A->numCols, const_cast<void*>(d_x), dt));
        MUXI_RS(f.create_dnvec_descr(&rs.y, A->numRows, d_y, dt));
        const rocsparse_spmv_alg a = algs[alg];
        const float a32 = 1.f, b32 = 0.f;
        const double a64 = 1.0, b64 = 0.0;
        const void* alpha = f64 ? (const void*)&a64 : (const void*)&a32;
        const void* beta = f64 ? (const void*)&b64 : (const void*)&b32;
        auto stage = [&](rocsparse_spmv_stage sg, size_t* bytes, void* buf) {
            return f.spmv(rs.h, rocsparse_operation_none, alpha, rs.mat, rs.x, beta, rs.y, dt, a, sg, bytes, buf);
        };
        size_t bytes = 0;
        const rocsparse_status bs = stage(rocsparse_spmv_stage_buffer_size, &bytes, nullptr);
        if (bs == rocsparse_status_invalid_value && alg != 0) {
            // an algorithm newer than the rocSPARSE in the process (a framework may bring an older copy)
            set_error("compare_with_rocsparse: " + f.path + " rejects algorithm " + std::to_string(alg) +
                      " (rocsparse status 7, invalid value)");
            return SPMV_ERROR_UNSUPPORTED_FORMAT;
        }
        MUXI_RS(bs);
        MUXI_HIP(hipMalloc(&rs.buf, std::max<size_t>(bytes, 16)));
        // The library's atomic-accumulating algorithms (adaptive on long rows, LRB, nnzsplit) scale y by beta and
        // then add into it, so with beta 0 a NaN already in y survives: y starts at zero, as the first call's
        // result would leave it for every later one.
        if (A->numRows > 0) MUXI_HIP(hipMemsetAsync(d_y, 0, (size_t)A->numRows * (f64 ? 8 : 4), s));
        EventPair ev;
        MUXI_HIP(hipEventCreate(&ev.a));
        MUXI_HIP(hipEventCreate(&ev.b));
        MUXI_HIP(hipEventRecord(ev.a, s));
        MUXI_RS(stage(rocsparse_spmv_stage_preprocess, &bytes, rs.buf));
        MUXI_HIP(hipEventRecord(ev.b, s));
        MUXI_HIP(hipEventSynchronize(ev.b));
        float ams = 0.f;
        MUXI_HIP(hipEventElapsedTime(&ams, ev.a, ev.b));
        if (analysisMs) *analysisMs = (double)ams;
        muxi_perf_result_t r{};
        r.iterations = config->measureIterations;
        auto run = [&]() -> spmv_status_t {
            MUXI_RS(stage(rocsparse_spmv_stage_compute, &bytes, rs.buf));
            if (config->syncAfterEach) MUXI_HIP(hipStreamSynchronize(s));
            return SPMV_SUCCESS;
        };
        st = time_iterations(config, s, A, &r, run);
        if (st != SPMV_SUCCESS) return st;
        *timeMs = r.timeMs;
        *gflops = r.gflops;
        return SPMV_SUCCESS;
    });
}

extern "C" spmv_status_t muxi_write_results_json(const muxi_perf_result_t* results, int count, const char* filename) {
    if (!filename || count < 0 || (count > 0 && !results)) {
        set_error("writeResultsJSON: filename or results is NULL");
        return SPMV_ERROR_INVALID_VECTOR;
    }
    FILE* fp = std::fopen(filename, "w");
    if (!fp) {
        set_error(std::string("writeResultsJSON: cannot open ") + filename + " for writing");
        return SPMV_ERROR_INVALID_MATRIX;  // the reference's status for this case
    }
    std::fprintf(fp, "{\n  \"results\": [\n");
    for (int i = 0; i < count; ++i) {
        const muxi_perf_result_t& r = results[i];
        std::fprintf(fp, "    {\n");
        std::fprintf(fp, "      \"matrix\": {\"rows\": %d, \"cols\": %d, \"nnz\": %d, \"sparsity\": %.6f},\n",
                     r.numRows, r.numCols, r.nnz, r.sparsity);
        std::fprintf(fp, "      \"device\": {\"id\": %d, \"warpSize\": %d, \"peakBandwidthGBs\": %.2f},\n", r.deviceId,
                     r.warpSize, r.peakBandwidthGBs);
        std::fprintf(fp, "      \"performance\": {\n");
        std::fprintf(fp, "        \"timeMs\": %.4f,\n", r.timeMs);
        std::fprintf(fp, "        \"minTimeMs\": %.4f,\n", r.minTimeMs);
        std::fprintf(fp, "        \"maxTimeMs\": %.4f,\n", r.maxTimeMs);
        std::fprintf(fp, "        \"avgTimeMs\": %.4f,\n", r.avgTimeMs);
        std::fprintf(fp, "        \"stdDevMs\": %.4f,\n", r.stdDevMs);
        std::fprintf(fp, "        \"gflops\": %.2f,\n", r.gflops);
        std::fprintf(fp, "        \"bandwidthGBs\": %.2f,\n", r.bandwidthGBs);
        std::fprintf(fp, "        \"bandwidthUtilization\": %.1f\n", r.bandwidthUtilization);
        std::fprintf(fp, "      },\n");
        std::fprintf(fp, "      \"correctnessPassed\": %s\n", r.correctnessPassed ? "true" : "false");
        std::fprintf(fp, "    }%s\n", i + 1 < count ? "," : "");
    }
    std::fprintf(fp, "  ]\n}\n");
    const bool ok = std::fclose(fp) == 0;
    if (!ok) {
        set_error(std::string("writeResultsJSON: error writing ") + filename);
        return SPMV_ERROR_INVALID_MATRIX;
    }
    return SPMV_SUCCESS;
}

extern "C" void muxi_print_performance_result(const muxi_perf_result_t* r) {
    if (!r) return;
    std::printf("\n=== Performance Results ===\n");
    std::printf("Matrix: %d x %d, NNZ: %d, Sparsity: %.4f%%\n", r->numRows, r->numCols, r->nnz, r->sparsity * 100);
    std::printf("Device: %d, Warp Size: %d\n", r->deviceId, r->warpSize);
    if (r->useCSR5) {
        std::printf("Format: CSR5 (tiles: %d, sigma: %d)\n", r->numTiles, r->tileSize);
        std::printf("Conversion Time: %.4f ms (%.1f%% overhead)\n", r->conversionTimeMs, r->conversionOverheadPercent);
    }
    std::printf("\n");
    std::printf("Time (median): %.4f ms\n", r->timeMs);
    std::printf("Time (min/max): %.4f / %.4f ms\n", r->minTimeMs, r->maxTimeMs);
    std::printf("Time (avg/std): %.4f / %.4f ms\n", r->avgTimeMs, r->stdDevMs);
    if (r->useCSR5) std::printf("Total Time (incl. conversion): %.4f ms\n", r->totalTimeMs);
    std::printf("\n");
    std::printf("GFLOPS: %.2f\n", r->gflops);
    std::printf("Bandwidth: %.2f GB/s\n", r->bandwidthGBs);
    std::printf("Bandwidth Utilization: %.1f%% (%.2f GB/s peak)\n", r->bandwidthUtilization, r->peakBandwidthGBs);
    std::printf("\n");
    std::printf("Correctness: %s\n", r->correctnessPassed ? "PASSED" : "FAILED");
    std::fflush(stdout);
}
