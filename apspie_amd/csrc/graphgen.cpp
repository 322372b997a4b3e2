// Host-side synthetic inputs and MatrixMarket reading.
//
// The reference ships no generator; these definitions belong to this
// library and are pinned by known-answer tests:
//   R-MAT(scale, ef, a, b, c, seed): ef * 2^scale edges; for each edge and
//     each of `scale` bit levels (first level = most significant bit) one
//     double r = genrand_res53 of std::mt19937(seed) picks the quadrant
//     r < a: (0,0), r < a+b: (0,1), r < a+b+c: (1,0), else (1,1).
//     No noise, no vertex permutation.  Then self-loops dropped, reverse
//     edges added, rows sorted, duplicates removed; isolated vertices kept.
//   Lattice(W, seed): W x W 4-neighbour grid, id = r*W + c, edges listed per
//     vertex as (right, down); a Fisher-Yates shuffle (j = draw % (i+1),
//     i from E-1 down to 1) removes the first floor(E/10) of them; then
//     floor(E/1000) random pairs (u = draw % n, v = draw % n, u != v, not
//     already adjacent) are added.
//   Weights(seed): for each pair u < v in CSR order, w = 1 + draw % 255,
//     mirrored onto (v, u).
// Sizes: call with rowptr/col NULL to get an upper bound on nnz, then call
// again with arrays of that capacity; *nnz then holds the exact count.
// genrand_res53 is the construction of mt19937ar.cpp:167-171.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <unordered_set>
#include <vector>

#include "../../include/apspie.h"

namespace {

inline double res53(std::mt19937 &g) {
    const uint32_t a = g() >> 5, b = g() >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

int n_threads() {
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

template <class F>
void parallel_for(int64_t n, F &&f) {
    const int T = n_threads();
    if (T == 1 || n < 4096) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const int64_t lo = n * t / T, hi = n * (t + 1) / T;
        th.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    for (auto &x : th) x.join();
}

// undirected edge list -> symmetric, sorted, dedup, loop-free CSR
struct Csr {
    std::vector<int32_t> rowptr, col;
};

Csr build_symmetric(int32_t n, const std::vector<int32_t> &u, const std::vector<int32_t> &v) {
    const int64_t ne = (int64_t)u.size();
    std::vector<int64_t> cnt((size_t)n + 1, 0);
    for (int64_t i = 0; i < ne; ++i) {
        if (u[i] == v[i]) continue;
        cnt[u[i] + 1]++;
        cnt[v[i] + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int32_t> tmp((size_t)cnt[n]);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < ne; ++i) {
        if (u[i] == v[i]) continue;
        tmp[fill[u[i]]++] = v[i];
        tmp[fill[v[i]]++] = u[i];
    }
    // sort + dedup each row in parallel, record unique counts
    std::vector<int32_t> uniq((size_t)n, 0);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t r = lo; r < hi; ++r) {
            int32_t *b = tmp.data() + cnt[r], *e = tmp.data() + cnt[r + 1];
            std::sort(b, e);
            uniq[r] = (int32_t)(std::unique(b, e) - b);
        }
    });
    Csr out;
    out.rowptr.assign((size_t)n + 1, 0);
    for (int32_t r = 0; r < n; ++r) out.rowptr[r + 1] = out.rowptr[r] + uniq[r];
    out.col.resize((size_t)out.rowptr[n]);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t r = lo; r < hi; ++r)
            std::memcpy(out.col.data() + out.rowptr[r], tmp.data() + cnt[r],
                        sizeof(int32_t) * uniq[r]);
    });
    return out;
}

int emit(const Csr &c, int32_t *rowptr, int32_t *col, int64_t *nnz) {
    *nnz = (int64_t)c.col.size();
    std::memcpy(rowptr, c.rowptr.data(), sizeof(int32_t) * c.rowptr.size());
    std::memcpy(col, c.col.data(), sizeof(int32_t) * c.col.size());
    return APSPIE_OK;
}

}  // namespace

extern "C" {

int apspie_gen_rmat(int scale, int ef, double a, double b, double c, uint32_t seed,
                    int32_t *rowptr, int32_t *col, int64_t *nnz) {
    if (!nnz || scale < 1 || scale > 30 || ef < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1)
        return APSPIE_ERR_INVALID;
    if ((int64_t)ef << scale > (int64_t)1 << 31) return APSPIE_ERR_UNSUPPORTED;
    const int64_t ne = (int64_t)ef << scale;
    if (!rowptr || !col) {  // size query: an upper bound on nnz
        *nnz = 2 * ne;
        return APSPIE_OK;
    }
    try {
        {
            std::vector<int32_t> u((size_t)ne), v((size_t)ne);
            std::mt19937 gen(seed);
            const double ab = a + b, abc = a + b + c;
            for (int64_t e = 0; e < ne; ++e) {
                int32_t x = 0, y = 0;
                for (int l = 0; l < scale; ++l) {
                    const double r = res53(gen);
                    const int xb = r >= ab;
                    const int yb = (r >= a && r < ab) || r >= abc;
                    x = (x << 1) | xb;
                    y = (y << 1) | yb;
                }
                u[e] = x;
                v[e] = y;
            }
            return emit(build_symmetric((int32_t)1 << scale, u, v), rowptr, col, nnz);
        }
    } catch (const std::bad_alloc &) {
        return APSPIE_ERR_OOM;
    }
}

int apspie_gen_lattice(int32_t W, uint32_t seed, int32_t *rowptr, int32_t *col, int64_t *nnz) {
    if (!nnz || W < 2 || (int64_t)W * W > INT32_MAX / 2) return APSPIE_ERR_INVALID;
    const int64_t n = (int64_t)W * W;
    const int64_t E = 2LL * W * (W - 1);
    if (!rowptr || !col) {  // size query: an upper bound on nnz
        *nnz = 2 * (E + E / 1000);
        return APSPIE_OK;
    }
    try {
        {
            std::vector<int32_t> lu((size_t)E), lv((size_t)E);
            int64_t k = 0;
            for (int64_t id = 0; id < n; ++id) {
                const int64_t r = id / W, cc = id % W;
                if (cc < W - 1) { lu[k] = (int32_t)id; lv[k] = (int32_t)(id + 1); ++k; }
                if (r < W - 1) { lu[k] = (int32_t)id; lv[k] = (int32_t)(id + W); ++k; }
            }
            std::mt19937 gen(seed);
            std::vector<int64_t> perm((size_t)E);
            for (int64_t i = 0; i < E; ++i) perm[i] = i;
            for (int64_t i = E - 1; i >= 1; --i) {
                const int64_t j = (int64_t)(gen() % (uint32_t)(i + 1));
                std::swap(perm[i], perm[j]);
            }
            std::vector<char> gone((size_t)E, 0);
            for (int64_t i = 0; i < E / 10; ++i) gone[perm[i]] = 1;
            std::vector<int32_t> u, v;
            u.reserve(E + E / 1000);
            v.reserve(E + E / 1000);
            std::unordered_set<uint64_t> have;
            have.reserve((size_t)(E + E / 1000) * 2);
            for (int64_t i = 0; i < E; ++i) {
                if (gone[i]) continue;
                u.push_back(lu[i]);
                v.push_back(lv[i]);
                have.insert(((uint64_t)(uint32_t)lu[i] << 32) | (uint32_t)lv[i]);
            }
            for (int64_t added = 0; added < E / 1000;) {
                const uint32_t x = gen() % (uint32_t)n, y = gen() % (uint32_t)n;
                if (x == y) continue;
                const uint32_t p = std::min(x, y), q = std::max(x, y);
                if (!have.insert(((uint64_t)p << 32) | q).second) continue;
                u.push_back((int32_t)p);
                v.push_back((int32_t)q);
                ++added;
            }
            return emit(build_symmetric((int32_t)n, u, v), rowptr, col, nnz);
        }
    } catch (const std::bad_alloc &) {
        return APSPIE_ERR_OOM;
    }
}

int apspie_gen_weights(int32_t m, const int32_t *rowptr, const int32_t *col, uint32_t seed,
                       float *val) {
    if (m < 0 || !rowptr || !val || (!col && rowptr[m])) return APSPIE_ERR_INVALID;
    // draws in CSR order for u < v, then mirrored in parallel
    std::vector<int64_t> up_start((size_t)m + 1, 0);
    for (int32_t u = 0; u < m; ++u) {
        int64_t k = 0;
        for (int32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) k += col[e] > u;
        up_start[u + 1] = up_start[u] + k;
    }
    std::vector<float> w((size_t)up_start[m]);
    std::mt19937 gen(seed);
    for (auto &x : w) x = (float)(1u + gen() % 255u);
    int bad = 0;
    parallel_for(m, [&](int64_t lo, int64_t hi) {
        for (int64_t u = lo; u < hi; ++u) {
            int64_t k = up_start[u];
            for (int32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
                const int32_t v = col[e];
                if (v > u) {
                    val[e] = w[k++];
                } else {
                    // weight of (v, u) with v < u: its rank among v's upper entries
                    const int32_t *b = col + rowptr[v], *en = col + rowptr[v + 1];
                    const int32_t *f = std::lower_bound(b, en, (int32_t)u);
                    if (f == en || *f != u) {
                        bad = 1;
                        continue;
                    }
                    const int32_t *ub = std::upper_bound(b, en, v);  // first col > v
                    val[e] = w[up_start[v] + (f - ub)];
                }
            }
        }
    });
    return bad ? APSPIE_ERR_INVALID : APSPIE_OK;
}

int apspie_read_mtx(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int32_t *row,
                    int32_t *col, float *val) {
    if (!path || !m || !n || !nnz) return APSPIE_ERR_INVALID;
    FILE *f = std::fopen(path, "r");
    if (!f) return APSPIE_ERR_INVALID;
    char line[4096];
    long long hm = 0, hn = 0, hz = 0;
    bool have_header = false;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '%') continue;  // readEdge skips comment lines
        if (std::sscanf(line, "%lld %lld %lld", &hm, &hn, &hz) == 3) {
            have_header = true;
            break;
        }
    }
    if (!have_header || hm < 0 || hn < 0 || hz < 0 || hm > INT32_MAX || hn > INT32_MAX) {
        std::fclose(f);
        return APSPIE_ERR_INVALID;
    }
    *m = (int32_t)hm;
    *n = (int32_t)hn;
    if (!row || !col) {  // size query
        *nnz = hz;
        std::fclose(f);
        return APSPIE_OK;
    }
    int64_t k = 0;
    while (k < hz && std::fgets(line, sizeof line, f)) {
        if (line[0] == '%' || line[0] == '\n') continue;
        long long a, b;
        double w;
        const int got = std::sscanf(line, "%lld %lld %lf", &a, &b, &w);
        if (got < 2 || a < 1 || b < 1 || a > hm || b > hn) {
            std::fclose(f);
            return APSPIE_ERR_INVALID;
        }
        // first index -> col, second -> row (readMtx, util.cuh:382-386)
        col[k] = (int32_t)(a - 1);
        row[k] = (int32_t)(b - 1);
        if (val) val[k] = got == 3 ? (float)w : 1.0f;
        ++k;
    }
    std::fclose(f);
    *nnz = k;
    return APSPIE_OK;
}

}  // extern "C"
