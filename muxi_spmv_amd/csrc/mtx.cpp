// Matrix Market coordinate files <-> host CSR (see muxi_spmv_io.h for the
// rules). Reference: src/generators/mtx_io.cu (readMTXHeader :70-112,
// readMTXFile :114-207, writeMTXFile :209-254).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <exception>
#include <new>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"
#include "muxi_spmv/muxi_spmv_formats.h"
#include "muxi_spmv/muxi_spmv_io.h"

using namespace muxi;

namespace {

struct File {
    FILE* f = nullptr;
    explicit File(const char* name, const char* mode) : f(std::fopen(name, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

const char* skip_ws(const char* p) {
    while (*p == ' ' || *p == '\t') ++p;
    return p;
}

bool starts(const char* p, const char* word) { return std::strncmp(p, word, std::strlen(word)) == 0; }

// "%%MatrixMarket matrix coordinate <type> <symmetry>"
bool parse_banner(const char* line, muxi_mtx_header_t* h) {
    if (!starts(line, "%%MatrixMarket")) return false;
    const char* p = skip_ws(line + 14);
    if (!starts(p, "matrix")) return false;
    p = skip_ws(p + 6);
    if (!starts(p, "coordinate")) return false;
    p = skip_ws(p + 10);
    h->isSymmetric = h->isComplex = h->isPattern = h->isInteger = false;
    if (starts(p, "real")) {
        p += 4;
    } else if (starts(p, "complex")) {
        h->isComplex = true;
        p += 7;
    } else if (starts(p, "pattern")) {
        h->isPattern = true;
        p += 7;
    } else if (starts(p, "integer")) {
        h->isInteger = true;
        p += 7;
    }
    p = skip_ws(p);
    h->isSymmetric = starts(p, "symmetric") || starts(p, "Symmetric");
    return true;
}

// Reads the banner and the dimension line; leaves f at the first entry.
spmv_status_t read_head(FILE* f, const char* name, muxi_mtx_header_t* h) {
    char line[1024];
    if (!std::fgets(line, sizeof line, f) || !parse_banner(line, h)) {
        set_error(std::string("mtx: no Matrix Market coordinate banner in ") + name);
        return SPMV_ERROR_INVALID_MATRIX;
    }
    bool got = false;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '%') continue;
        const char* p = skip_ws(line);
        if (*p == '\n' || *p == '\r' || *p == 0) continue;  // blank line
        got = true;
        break;
    }
    if (!got || std::sscanf(line, "%d %d %d", &h->numRows, &h->numCols, &h->nnz) != 3 || h->numRows < 0 ||
        h->numCols < 0 || h->nnz < 0) {
        set_error(std::string("mtx: bad dimension line in ") + name);
        return SPMV_ERROR_INVALID_MATRIX;
    }
    return SPMV_SUCCESS;
}

template <typename T>
spmv_status_t read_csr(const char* name, int expand, muxi_csr_t* out) {
    if (!name || !out) return SPMV_ERROR_INVALID_MATRIX;
    File file(name, "r");
    if (!file.f) {
        set_error(std::string("mtx: cannot open ") + name);
        return SPMV_ERROR_INVALID_MATRIX;
    }
    muxi_mtx_header_t h;
    spmv_status_t st = read_head(file.f, name, &h);
    if (st != SPMV_SUCCESS) return st;
    if (h.isComplex) {
        set_error("mtx: complex matrices are not supported");
        return SPMV_ERROR_UNSUPPORTED_TYPE;
    }
    const bool mirror = h.isSymmetric && expand;
    std::vector<int> ri, cj;
    std::vector<T> vv;
    // reserve what the header declares, up to 2^26 entries (more grow as they are read)
    const size_t want = std::min<size_t>((size_t)h.nnz * (mirror ? 2 : 1), (size_t)1 << 26);
    ri.reserve(want);
    cj.reserve(want);
    vv.reserve(want);
    for (int e = 0; e < h.nnz; ++e) {
        int r, c;
        double v = 1.0;  // pattern entries are 1
        const int want = h.isPattern ? 2 : 3;
        const int got = h.isPattern ? std::fscanf(file.f, "%d %d", &r, &c) : std::fscanf(file.f, "%d %d %lf", &r, &c, &v);
        if (got != want) {
            set_error(std::string("mtx: entry ") + std::to_string(e + 1) + " unreadable in " + name);
            return SPMV_ERROR_INVALID_MATRIX;
        }
        --r;
        --c;
        if (r < 0 || r >= h.numRows || c < 0 || c >= h.numCols) {
            set_error(std::string("mtx: entry ") + std::to_string(e + 1) + " outside the matrix in " + name);
            return SPMV_ERROR_INVALID_MATRIX;
        }
        ri.push_back(r);
        cj.push_back(c);
        vv.push_back((T)v);
        if (mirror && r != c) {
            ri.push_back(c);
            cj.push_back(r);
            vv.push_back((T)v);
        }
    }
    const size_t n = vv.size();
    if (n > (size_t)INT32_MAX) {
        set_error("mtx: more than INT32_MAX entries");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return ri[a] != ri[b] ? ri[a] < ri[b] : cj[a] < cj[b]; });
    int* rp = (int*)std::malloc(sizeof(int) * ((size_t)h.numRows + 1));
    int* ci = (int*)std::malloc(sizeof(int) * (n ? n : 1));
    T* va = (T*)std::malloc(sizeof(T) * (n ? n : 1));
    if (!rp || !ci || !va) {
        std::free(rp);
        std::free(ci);
        std::free(va);
        return SPMV_ERROR_MEMORY_ALLOC;
    }
    std::fill(rp, rp + h.numRows + 1, 0);
    for (size_t k = 0; k < n; ++k) {
        rp[ri[ord[k]] + 1]++;
        ci[k] = cj[ord[k]];
        va[k] = vv[ord[k]];
    }
    for (int r = 0; r < h.numRows; ++r) rp[r + 1] += rp[r];
    std::memset(out, 0, sizeof(*out));
    out->numRows = h.numRows;
    out->numCols = h.numCols;
    out->nnz = (int)n;
    out->rowPtr = rp;
    out->colIdx = ci;
    out->values = va;
    out->ownsHostMemory = true;
    return SPMV_SUCCESS;
}

template <typename T>
spmv_status_t write_csr(const char* name, const muxi_csr_t* A, int symmetric) {
    if (!name || !A || A->numRows < 0 || (A->numRows > 0 && !A->rowPtr) || (A->nnz > 0 && (!A->colIdx || !A->values)))
        return SPMV_ERROR_INVALID_MATRIX;
    File file(name, "w");
    if (!file.f) {
        set_error(std::string("mtx: cannot open ") + name + " for writing");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const T* v = (const T*)A->values;
    long count = 0;
    for (int r = 0; r < A->numRows; ++r)
        for (int k = A->rowPtr[r]; k < A->rowPtr[r + 1]; ++k) count += !symmetric || A->colIdx[k] >= r;
    std::fprintf(file.f, "%%%%MatrixMarket matrix coordinate real %s\n", symmetric ? "symmetric" : "general");
    std::fprintf(file.f, "%% written by muxi_spmv_amd\n");
    std::fprintf(file.f, "%% Rows: %d, Columns: %d, NNZ: %d\n", A->numRows, A->numCols, A->nnz);
    std::fprintf(file.f, "%d %d %ld\n", A->numRows, A->numCols, count);
    for (int r = 0; r < A->numRows; ++r)
        for (int k = A->rowPtr[r]; k < A->rowPtr[r + 1]; ++k) {
            const int c = A->colIdx[k];
            if (symmetric && c < r) continue;  // the upper triangle only
            if (sizeof(T) == 8)
                std::fprintf(file.f, "%d %d %.17g\n", r + 1, c + 1, (double)v[k]);
            else
                std::fprintf(file.f, "%d %d %.9g\n", r + 1, c + 1, (double)v[k]);
        }
    if (std::ferror(file.f)) {
        set_error(std::string("mtx: write failed on ") + name);
        return SPMV_ERROR_INVALID_MATRIX;
    }
    return SPMV_SUCCESS;
}

// readMTXToCOO: the file's entries as stored (no symmetric mirroring), in
// (row, column) order, rows expanded from the CSR read.
template <typename T>
spmv_status_t read_coo(const char* name, muxi_coo_t* out) {
    if (!out) return SPMV_ERROR_INVALID_MATRIX;
    muxi_csr_t c;
    spmv_status_t st = read_csr<T>(name, 0, &c);
    if (st != SPMV_SUCCESS) return st;
    int* ri = (int*)std::malloc(sizeof(int) * (c.nnz ? (size_t)c.nnz : 1));
    if (!ri) {
        std::free(c.rowPtr);
        std::free(c.colIdx);
        std::free(c.values);
        return SPMV_ERROR_MEMORY_ALLOC;
    }
    for (int r = 0; r < c.numRows; ++r)
        for (int k = c.rowPtr[r]; k < c.rowPtr[r + 1]; ++k) ri[k] = r;
    std::free(c.rowPtr);
    std::memset(out, 0, sizeof(*out));
    out->numRows = c.numRows;
    out->numCols = c.numCols;
    out->nnz = c.nnz;
    out->rowIdx = ri;
    out->colIdx = c.colIdx;
    out->values = c.values;
    out->ownsHostMemory = true;
    return SPMV_SUCCESS;
}

// writeCOOToMTX: a "general" file with the entries in the given order.
template <typename T>
spmv_status_t write_coo(const char* name, const muxi_coo_t* A) {
    if (!name || !A || A->numRows < 0 || A->numCols < 0 || A->nnz < 0 ||
        (A->nnz > 0 && (!A->rowIdx || !A->colIdx || !A->values)))
        return SPMV_ERROR_INVALID_MATRIX;
    for (int k = 0; k < A->nnz; ++k)
        if (A->rowIdx[k] < 0 || A->rowIdx[k] >= A->numRows || A->colIdx[k] < 0 || A->colIdx[k] >= A->numCols) {
            set_error("mtx: COO entry outside the matrix");
            return SPMV_ERROR_INVALID_MATRIX;
        }
    File file(name, "w");
    if (!file.f) {
        set_error(std::string("mtx: cannot open ") + name + " for writing");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const T* v = (const T*)A->values;
    std::fprintf(file.f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(file.f, "%% written by muxi_spmv_amd\n");
    std::fprintf(file.f, "%d %d %d\n", A->numRows, A->numCols, A->nnz);
    for (int k = 0; k < A->nnz; ++k) {
        if (sizeof(T) == 8)
            std::fprintf(file.f, "%d %d %.17g\n", A->rowIdx[k] + 1, A->colIdx[k] + 1, (double)v[k]);
        else
            std::fprintf(file.f, "%d %d %.9g\n", A->rowIdx[k] + 1, A->colIdx[k] + 1, (double)v[k]);
    }
    if (std::ferror(file.f)) {
        set_error(std::string("mtx: write failed on ") + name);
        return SPMV_ERROR_INVALID_MATRIX;
    }
    return SPMV_SUCCESS;
}

}  // namespace

extern "C" spmv_status_t muxi_mtx_read_header(const char* filename, muxi_mtx_header_t* header) {
    if (!filename || !header) return SPMV_ERROR_INVALID_MATRIX;
    File file(filename, "r");
    if (!file.f) {
        set_error(std::string("mtx: cannot open ") + filename);
        return SPMV_ERROR_INVALID_MATRIX;
    }
    return read_head(file.f, filename, header);
}

// Host containers may throw (a file that declares billions of entries); no
// exception crosses the C ABI.
template <typename F>
static spmv_status_t guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        set_error("mtx: out of host memory");
        return SPMV_ERROR_MEMORY_ALLOC;
    } catch (const std::exception& e) {
        set_error(std::string("mtx: ") + e.what());
        return SPMV_ERROR_INTERNAL;
    }
}

extern "C" spmv_status_t muxi_mtx_read_csr_f32(const char* filename, int expandSymmetric, muxi_csr_t* out) {
    return guarded([&] { return read_csr<float>(filename, expandSymmetric, out); });
}
extern "C" spmv_status_t muxi_mtx_read_csr_f64(const char* filename, int expandSymmetric, muxi_csr_t* out) {
    return guarded([&] { return read_csr<double>(filename, expandSymmetric, out); });
}
extern "C" spmv_status_t muxi_mtx_write_csr_f32(const char* filename, const muxi_csr_t* A, int writeSymmetric) {
    return guarded([&] { return write_csr<float>(filename, A, writeSymmetric); });
}
extern "C" spmv_status_t muxi_mtx_write_csr_f64(const char* filename, const muxi_csr_t* A, int writeSymmetric) {
    return guarded([&] { return write_csr<double>(filename, A, writeSymmetric); });
}

extern "C" spmv_status_t muxi_csr_free_host(muxi_csr_t* A) {
    if (!A) return SPMV_ERROR_INVALID_MATRIX;
    std::free(A->rowPtr);
    std::free(A->colIdx);
    std::free(A->values);
    A->rowPtr = nullptr;
    A->colIdx = nullptr;
    A->values = nullptr;
    A->ownsHostMemory = false;
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_mtx_read_coo_f32(const char* filename, muxi_coo_t* out) {
    return guarded([&] { return read_coo<float>(filename, out); });
}
extern "C" spmv_status_t muxi_mtx_read_coo_f64(const char* filename, muxi_coo_t* out) {
    return guarded([&] { return read_coo<double>(filename, out); });
}
extern "C" spmv_status_t muxi_mtx_write_coo_f32(const char* filename, const muxi_coo_t* A) {
    return write_coo<float>(filename, A);
}
extern "C" spmv_status_t muxi_mtx_write_coo_f64(const char* filename, const muxi_coo_t* A) {
    return write_coo<double>(filename, A);
}

extern "C" spmv_status_t muxi_coo_free_host(muxi_coo_t* A) {
    if (!A) return SPMV_ERROR_INVALID_MATRIX;
    std::free(A->rowIdx);
    std::free(A->colIdx);
    std::free(A->values);
    A->rowIdx = nullptr;
    A->colIdx = nullptr;
    A->values = nullptr;
    A->ownsHostMemory = false;
    return SPMV_SUCCESS;
}
