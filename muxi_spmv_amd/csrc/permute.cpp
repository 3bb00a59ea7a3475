// Host permutations (the reference computes all of these on the host).
//
// Orders that are total (centre of mass with the column as tie-break, row
// length with a stable tie-break, (column, value) within a row) are computed
// here with counting and radix sorts. reorderColumnsRCM and
// reorderColumnsByFrequency order by a key with ties and use an unstable
// std::sort in the reference; their tie order is whatever libstdc++'s
// introsort produces, so they call std::sort with the same comparator on the
// same initial sequence to reproduce it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "common.hpp"
#include "muxi_spmv/muxi_spmv_prep.h"

using namespace muxi;

namespace {

// Order-preserving map of a finite float to uint32.
inline uint32_t float_key(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Stable LSD radix sort of idx by 32-bit keys (4 passes of 8 bits).
void radix_sort_stable(std::vector<int>& idx, const std::vector<uint32_t>& key) {
    const size_t n = idx.size();
    std::vector<int> tmp(n);
    for (int pass = 0; pass < 4; ++pass) {
        const int sh = pass * 8;
        size_t cnt[257] = {0};
        for (size_t i = 0; i < n; ++i) cnt[((key[idx[i]] >> sh) & 0xff) + 1]++;
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (size_t i = 0; i < n; ++i) tmp[cnt[(key[idx[i]] >> sh) & 0xff]++] = idx[i];
        idx.swap(tmp);
    }
}

bool valid_csr(int R, int C, const int* rp, const int* ci) {
    if (R < 0 || C < 0 || (R > 0 && !rp)) return false;
    if (R > 0 && rp[R] > 0 && !ci) return false;
    return true;
}

}  // namespace

extern "C" spmv_status_t muxi_reorder_columns(int R, int C, const int* rp, const int* ci, int* fwd, int* nci,
                                              muxi_column_reorder_stats_t* stats) {
    if (!valid_csr(R, C, rp, ci) || (C > 0 && !fwd)) {
        set_error("reorder_columns: invalid matrix");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const int N = R > 0 ? rp[R] : 0;
    std::vector<float> com(C, 0.0f);
    std::vector<int> cnt(C, 0);
    for (int r = 0; r < R; ++r)
        for (int i = rp[r]; i < rp[r + 1]; ++i) {
            com[ci[i]] += (float)r;  // FP32 sum in row order
            cnt[ci[i]]++;
        }
    std::vector<uint32_t> key(C);
    for (int c = 0; c < C; ++c) {
        if (cnt[c] > 0) com[c] /= (float)cnt[c];
        key[c] = float_key(com[c]);
    }
    std::vector<int> order(C);
    for (int c = 0; c < C; ++c) order[c] = c;
    radix_sort_stable(order, key);  // ties stay in column order == pair<float,int> order
    std::vector<int> rev(C);
    int moved = 0;
    for (int n = 0; n < C; ++n) {
        fwd[n] = order[n];
        rev[order[n]] = n;
        moved += order[n] != n;
    }
    if (nci)
        for (int i = 0; i < N; ++i) nci[i] = rev[ci[i]];
    if (stats) {
        // column_reorder.cuh:119-138: first 10000 rows, FP32 sums
        float before = 0, after = 0;
        int count = 0;
        for (int r = 0; r < R && r < 10000; ++r)
            for (int i = rp[r]; i < rp[r + 1] - 1; ++i) {
                before += std::abs(ci[i + 1] - ci[i]);
                after += std::abs(rev[ci[i + 1]] - rev[ci[i]]);
                count++;
            }
        stats->numReorders = moved;
        stats->avgDistanceBefore = count ? before / count : 0.f;
        stats->avgDistanceAfter = count ? after / count : 0.f;
        stats->improvementRatio = count ? stats->avgDistanceBefore / std::max(stats->avgDistanceAfter, 1.0f) : 0.f;
    }
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_reorder_x_f32(const float* x, float* xr, const int* fwd, int C) {
    if (C < 0 || (C > 0 && (!x || !xr || !fwd))) return SPMV_ERROR_INVALID_VECTOR;
    for (int n = 0; n < C; ++n) xr[n] = x[fwd[n]];
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_sort_columns_within_rows_f32(int R, const int* rp, int* ci, float* v) {
    if (!valid_csr(R, 1, rp, ci) || (R > 0 && rp[R] > 0 && !v)) return SPMV_ERROR_INVALID_MATRIX;
    std::vector<std::pair<int, float>> el;
    for (int r = 0; r < R; ++r) {
        const int b = rp[r], e = rp[r + 1];
        const int n = e - b;
        if (n < 2) continue;
        el.resize(n);
        for (int i = 0; i < n; ++i) el[i] = {ci[b + i], v[b + i]};
        // (column, value) lexicographic; short rows by insertion
        auto less = [](const std::pair<int, float>& a, const std::pair<int, float>& c) {
            return a.first < c.first || (a.first == c.first && a.second < c.second);
        };
        if (n <= 32) {
            for (int i = 1; i < n; ++i) {
                auto t = el[i];
                int j = i - 1;
                while (j >= 0 && less(t, el[j])) {
                    el[j + 1] = el[j];
                    --j;
                }
                el[j + 1] = t;
            }
        } else {
            std::stable_sort(el.begin(), el.end(), less);
        }
        for (int i = 0; i < n; ++i) {
            ci[b + i] = el[i].first;
            v[b + i] = el[i].second;
        }
    }
    return SPMV_SUCCESS;
}

static spmv_status_t column_perm_by_key(int R, int C, const int* rp, const int* ci, int* perm, int* nci,
                                        bool minrow) {
    if (!valid_csr(R, C, rp, ci) || (C > 0 && !perm)) {
        set_error("column permutation: invalid matrix");
        return SPMV_ERROR_INVALID_MATRIX;
    }
    const int N = R > 0 ? rp[R] : 0;
    std::vector<std::pair<int, int>> kv(C);
    for (int c = 0; c < C; ++c) kv[c] = {c, minrow ? R : 0};
    if (minrow) {
        for (int r = 0; r < R; ++r)
            for (int i = rp[r]; i < rp[r + 1]; ++i) kv[ci[i]].second = std::min(kv[ci[i]].second, r);
        std::sort(kv.begin(), kv.end(),
                  [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.second < b.second; });
    } else {
        for (int i = 0; i < N; ++i) kv[ci[i]].second++;
        std::sort(kv.begin(), kv.end(),
                  [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.second > b.second; });
    }
    std::vector<int> inv(C);
    for (int i = 0; i < C; ++i) {
        perm[i] = kv[i].first;
        inv[kv[i].first] = i;
    }
    if (nci)
        for (int i = 0; i < N; ++i) nci[i] = inv[ci[i]];
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_reorder_columns_minrow(int R, int C, const int* rp, const int* ci, int* perm,
                                                     int* nci) {
    return column_perm_by_key(R, C, rp, ci, perm, nci, true);
}

extern "C" spmv_status_t muxi_reorder_columns_by_frequency(int R, int C, const int* rp, const int* ci, int* perm,
                                                           int* nci) {
    return column_perm_by_key(R, C, rp, ci, perm, nci, false);
}

extern "C" spmv_status_t muxi_reorder_rows_by_length_f32(int R, const int* rp, const int* ci, const float* v,
                                                         int* rowMap, int* invRowMap, int* nrp, int* nci, float* nv) {
    if (!valid_csr(R, 1, rp, ci) || (R > 0 && (!rowMap || !invRowMap))) return SPMV_ERROR_INVALID_MATRIX;
    // stable counting sort by length
    int maxlen = 0;
    for (int r = 0; r < R; ++r) maxlen = std::max(maxlen, rp[r + 1] - rp[r]);
    std::vector<int> start(maxlen + 2, 0);
    for (int r = 0; r < R; ++r) start[rp[r + 1] - rp[r] + 1]++;
    for (int l = 0; l <= maxlen; ++l) start[l + 1] += start[l];
    for (int r = 0; r < R; ++r) {
        int pos = start[rp[r + 1] - rp[r]]++;
        rowMap[pos] = r;
        invRowMap[r] = pos;
    }
    if (nrp) {
        nrp[0] = 0;
        int pos = 0;
        for (int i = 0; i < R; ++i) {
            int o = rowMap[i];
            for (int j = rp[o]; j < rp[o + 1]; ++j) {
                if (nci) nci[pos] = ci[j];
                if (nv) nv[pos] = v[j];
                pos++;
            }
            nrp[i + 1] = pos;
        }
    }
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_shard_rows(const int* rp, int R, int G, int* bounds) {
    if (R < 0 || G < 1 || !bounds || (R > 0 && !rp)) return SPMV_ERROR_INVALID_MATRIX;
    const long N = R > 0 ? rp[R] : 0, L = (long)R + N;
    bounds[0] = 0;
    for (int g = 1; g < G; ++g) {
        const long d = (long)g * L / G;
        // merge-path coordinate row at diagonal d (the search the merge plan kernels use): the number of row
        // ends among the first d path items, the smallest r in [max(0, d-N), min(d, R)] with rp[r+1] >= d - r
        long lo = d - N > 0 ? d - N : 0, hi = d < R ? d : R;
        while (lo < hi) {
            const long mid = (lo + hi) >> 1;
            if ((long)rp[mid + 1] <= d - mid - 1)
                lo = mid + 1;
            else
                hi = mid;
        }
        bounds[g] = (int)lo;
    }
    bounds[G] = R;
    return SPMV_SUCCESS;
}

extern "C" int muxi_bandwidth(int R, const int* rp, const int* ci) {
    int bw = 0;
    for (int r = 0; r < R; ++r)
        for (int i = rp[r]; i < rp[r + 1]; ++i) bw = std::max(bw, std::abs(ci[i] - r));
    return bw;
}

extern "C" spmv_status_t muxi_rcm(int n, const int* rp, const int* ci, int* perm) {
    if (n < 0 || (n > 0 && (!rp || !perm))) return SPMV_ERROR_INVALID_MATRIX;
    if (n == 0) return SPMV_SUCCESS;
    // symmetrised adjacency without self loops
    std::vector<int> deg(n, 0);
    for (int r = 0; r < n; ++r)
        for (int i = rp[r]; i < rp[r + 1]; ++i) {
            int c = ci[i];
            if (c == r || c < 0 || c >= n) continue;
            deg[r]++;
            deg[c]++;
        }
    std::vector<long> off(n + 1, 0);
    for (int r = 0; r < n; ++r) off[r + 1] = off[r] + deg[r];
    std::vector<int> adj(off[n]);
    std::vector<long> fill(off.begin(), off.end() - 1);
    for (int r = 0; r < n; ++r)
        for (int i = rp[r]; i < rp[r + 1]; ++i) {
            int c = ci[i];
            if (c == r || c < 0 || c >= n) continue;
            adj[fill[r]++] = c;
            adj[fill[c]++] = r;
        }
    // dedupe and sort each neighbour list by (degree, index)
    for (int r = 0; r < n; ++r) {
        auto b = adj.begin() + off[r], e = adj.begin() + off[r + 1];
        std::sort(b, e);
        auto ne = std::unique(b, e);
        std::fill(ne, e, -1);
    }
    std::vector<int> udeg(n);
    for (int r = 0; r < n; ++r) {
        int d = 0;
        for (long i = off[r]; i < off[r + 1] && adj[i] >= 0; ++i) ++d;
        udeg[r] = d;
    }
    for (int r = 0; r < n; ++r)
        std::sort(adj.begin() + off[r], adj.begin() + off[r] + udeg[r],
                  [&](int a, int b) { return udeg[a] < udeg[b] || (udeg[a] == udeg[b] && a < b); });
    std::vector<int> level(n, -1), order;
    order.reserve(n);
    std::vector<char> done(n, 0);
    // BFS from s; returns the vertices of the last level and the depth
    auto bfs = [&](int s, std::vector<int>& visit, std::vector<int>& lastlvl) -> int {
        visit.clear();
        visit.push_back(s);
        level[s] = 0;
        size_t h = 0;
        int depth = 0;
        while (h < visit.size()) {
            int u = visit[h++];
            depth = std::max(depth, level[u]);
            for (long i = off[u]; i < off[u] + udeg[u]; ++i) {
                int w = adj[i];
                if (level[w] < 0) {
                    level[w] = level[u] + 1;
                    visit.push_back(w);
                }
            }
        }
        lastlvl.clear();
        for (int u : visit)
            if (level[u] == depth) lastlvl.push_back(u);
        for (int u : visit) level[u] = -1;
        return depth;
    };
    std::vector<int> visit, last;
    for (int seed = 0; seed < n; ++seed) {
        if (done[seed]) continue;
        // pseudo-peripheral start (George-Liu): repeat BFS from a min-degree
        // vertex of the last level while the depth grows
        int s = seed;
        int depth = bfs(s, visit, last);
        for (int it = 0; it < 8; ++it) {
            int best = last[0];
            for (int u : last)
                if (udeg[u] < udeg[best] || (udeg[u] == udeg[best] && u < best)) best = u;
            std::vector<int> v2, l2;
            int d2 = bfs(best, v2, l2);
            if (d2 <= depth) break;
            s = best;
            depth = d2;
            last.swap(l2);
        }
        // Cuthill-McKee from s
        size_t h = order.size();
        order.push_back(s);
        done[s] = 1;
        while (h < order.size()) {
            int u = order[h++];
            for (long i = off[u]; i < off[u] + udeg[u]; ++i) {
                int w = adj[i];
                if (!done[w]) {
                    done[w] = 1;
                    order.push_back(w);
                }
            }
        }
    }
    for (int i = 0; i < n; ++i) perm[i] = order[n - 1 - i];
    return SPMV_SUCCESS;
}

extern "C" spmv_status_t muxi_permute_symmetric(int n, const int* rp, const int* ci, const void* vals,
                                                spmv_type_t vt, const int* perm, int* nrp, int* nci, void* nvals) {
    if (n < 0 || (n > 0 && (!rp || !perm || !nrp))) return SPMV_ERROR_INVALID_MATRIX;
    if (vt != SPMV_TYPE_FLOAT && vt != SPMV_TYPE_DOUBLE) return SPMV_ERROR_UNSUPPORTED_TYPE;
    const size_t es = vt == SPMV_TYPE_DOUBLE ? 8 : 4;
    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) inv[perm[i]] = i;
    nrp[0] = 0;
    std::vector<std::pair<int, int>> row;
    for (int i = 0; i < n; ++i) {
        int o = perm[i];
        row.clear();
        for (int j = rp[o]; j < rp[o + 1]; ++j) row.push_back({inv[ci[j]], j});
        std::sort(row.begin(), row.end());
        int pos = nrp[i];
        for (auto& pr : row) {
            nci[pos] = pr.first;
            std::memcpy((char*)nvals + es * pos, (const char*)vals + es * pr.second, es);
            ++pos;
        }
        nrp[i + 1] = pos;
    }
    return SPMV_SUCCESS;
}
