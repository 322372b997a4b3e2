// Autotuner: launch parameters of the traversals and the bin bounds of the
// hash machinery, timed on the graph at hand, and a persistent table of the
// winners per (GPU, graph class).
//
// No reference counterpart: the reference hard-codes NTHREADS 512, its mgpu
// LaunchBox tables (spgemm.cuh:181-185) and its hash SpGEMM's TABLE_SIZE,
// MAX_PROBES, THREADS and BLOCKS (spgemmKernel.cuh:9-26).  Here the tiers'
// table capacities are fixed per kernel, and what the tuner moves is which
// tier -- and so which table size and block shape -- serves a row:
//   apspie_autotune, batch == 0: single-source BFS over tile size x blocks
//     per CU, then the direction switch threshold alpha, and for a weighted
//     graph the SSSP near-far width (Bellman-Ford, and half, once and twice
//     the width chosen from the weights);
//   apspie_autotune, batch > 0: batched BFS over blocks per CU x levels per
//     hipGraph chunk, then the pull threshold, and for a weighted graph the
//     batched SSSP width (Bellman-Ford, 256, 1024);
//   apspie_autotune_hash: TC's medium / big tier bound (a 256-thread block
//     with a 4096-slot LDS table vs a 1024-thread block with 32768 slots),
//     then SpGEMM's wave / block tier bounds (512- and 8192-slot tables) and
//     whether rows beyond them use the dense accumulator or global tables
//     (A * A; skipped when the product does not fit one call).
// Every candidate is timed three times (median of the calls' device time);
// the buffers come from the context's scratch pool (its HBM budget applies)
// and every copy runs on its stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace apspie {

namespace {

// the options the tuners set: what a saved table entry holds
const char *const kTuned[] = {"bfs_tile",         "bfs_blocks_per_cu",  "bfs_alpha",
                              "sssp_delta",       "batch_blocks_per_cu", "bfs_levels_per_sync",
                              "batch_pull_div",   "batch_sssp_delta",    "tc_medium_row",
                              "spgemm_wave_ub",   "spgemm_block_ub",     "spgemm_dense"};

template <class Run>
double median_ms(Ctx *c, Run run) {
    std::vector<double> t;
    for (int rep = 0; rep < 3; ++rep) {
        const int st = run();
        if (st != APSPIE_OK) throw Error{st, "autotune run failed: " + c->last_error};
        t.push_back(c->stats.elapsed_ms);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

// the candidate among `values` of option `name` with the smallest time of run()
template <class Run>
void pick(apspie_ctx *c, const char *name, std::initializer_list<int64_t> values, Run run) {
    double best = 1e300;
    int64_t arg = opt(c, name, 0);
    for (int64_t v : values) {
        if (apspie_ctx_set_option(c, name, v) != APSPIE_OK) continue;
        const double t = median_ms(c, run);
        if (t < best) {
            best = t;
            arg = v;
        }
    }
    APSPIE_CHECK(apspie_ctx_set_option(c, name, arg) == APSPIE_OK, APSPIE_ERR_DEVICE,
                 "autotune could not set its pick");
}

int ilog2_ceil(int64_t x) {
    int l = 0;
    while ((int64_t(1) << l) < x) ++l;
    return l;
}

}  // namespace

// "gfx950_cu256": the architecture name up to its feature list, and the CU count
std::string tune_gpu_key(Ctx *c) {
    hipDeviceProp_t prop;
    APSPIE_HIP(hipGetDeviceProperties(&prop, c->device));
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    return arch + "_cu" + std::to_string(c->cus);
}

// "v22_d5_x18_sw": log2 of the vertex count (rounded up), of the mean row
// length (rounded), of the longest row (rounded up), symmetric or not (s/u)
// and weighted or not (w/n)
std::string tune_class_key(Ctx *c, Graph *g) {
    const double mean = g->m > 0 ? (double)g->nnz / g->m : 0.0;
    const int d = mean >= 1.0 ? (int)std::lround(std::log2(mean)) : 0;
    std::string k = "v" + std::to_string(ilog2_ceil(std::max<int64_t>(g->m, 1))) + "_d" +
                    std::to_string(d) + "_x" +
                    std::to_string(ilog2_ceil(std::max<int32_t>(graph_max_degree(c, g), 1))) +
                    "_";
    k += graph_is_symmetric(c, g) ? "s" : "u";
    k += g->val ? "w" : "n";
    return k;
}

}  // namespace apspie

using namespace apspie;

extern "C" int apspie_autotune(apspie_ctx *c, const apspie_graph *g, int32_t batch) {
    if (!c || !g || batch < 0) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        APSPIE_CHECK(g->m == g->n && g->m > 0, APSPIE_ERR_INVALID, "autotune needs a square graph");
        APSPIE_CHECK(batch <= 4096, APSPIE_ERR_INVALID, "autotune batches hold at most 4096 sources");
        set_device(c);
        std::vector<int32_t> rowptr(g->m + 1);
        copy_from_device(c, rowptr.data(), g->rowptr, sizeof(int32_t) * (g->m + 1),
                         APSPIE_MEM_HOST);
        sync(c);
        // probe sources: a fixed sample of vertices with edges
        std::mt19937 gen(12345);
        std::vector<int32_t> src;
        for (int tries = 0; tries < 100000 && (int)src.size() < std::max(3, batch); ++tries) {
            const int32_t v = (int32_t)(gen() % (uint32_t)g->m);
            if (rowptr[v + 1] > rowptr[v]) src.push_back(v);
        }
        APSPIE_CHECK(!src.empty(), APSPIE_ERR_INVALID, "graph has no edges");
        const size_t cols = batch > 0 ? (size_t)batch : 1;
        int32_t *out = (int32_t *)scratch(c, "tune_out", sizeof(int32_t) * (size_t)g->m * cols);
        const size_t n3 = std::min<size_t>(3, src.size());
        if (batch == 0) {
            auto bfs3 = [&] {
                for (size_t i = 0; i < n3; ++i) {
                    const int st = apspie_bfs(c, g, src[i], out, APSPIE_MEM_DEVICE);
                    if (st != APSPIE_OK) return st;
                }
                return (int)APSPIE_OK;
            };
            // tile x blocks per CU jointly, then alpha
            double best = 1e300;
            int64_t bt = 2048, bb = 8;
            for (int64_t tile : {1024, 2048, 4096})
                for (int64_t bpc : {2, 4, 8}) {
                    apspie_ctx_set_option(c, "bfs_tile", tile);
                    apspie_ctx_set_option(c, "bfs_blocks_per_cu", bpc);
                    double tot = 0;
                    for (size_t i = 0; i < n3; ++i)
                        tot += median_ms(c, [&] {
                            return apspie_bfs(c, g, src[i], out, APSPIE_MEM_DEVICE);
                        });
                    if (tot < best) { best = tot; bt = tile; bb = bpc; }
                }
            apspie_ctx_set_option(c, "bfs_tile", bt);
            apspie_ctx_set_option(c, "bfs_blocks_per_cu", bb);
            pick(c, "bfs_alpha", {7, 14, 28}, bfs3);
            if (g->val) {
                const int64_t d = (int64_t)std::min<double>(
                    std::max<double>(sssp_weight_delta(c, const_cast<apspie_graph *>(g)), 64),
                    1 << 29);
                pick(c, "sssp_delta", {(int64_t)0, d / 2, d, 2 * d}, [&] {
                    for (size_t i = 0; i < n3; ++i) {
                        const int st = apspie_sssp(c, g, src[i], (float *)out, APSPIE_MEM_DEVICE);
                        if (st != APSPIE_OK) return st;
                    }
                    return (int)APSPIE_OK;
                });
            }
        } else {
            src.resize((size_t)batch);
            auto bfsb = [&] {
                return apspie_bfs_batch(c, g, src.data(), batch, out, APSPIE_MEM_DEVICE);
            };
            double best = 1e300;
            int64_t bb = 4, bl = 12;
            for (int64_t bpc : {2, 4, 8})
                for (int64_t lps : {6, 12, 48}) {
                    apspie_ctx_set_option(c, "batch_blocks_per_cu", bpc);
                    apspie_ctx_set_option(c, "bfs_levels_per_sync", lps);
                    const double t = median_ms(c, bfsb);
                    if (t < best) { best = t; bb = bpc; bl = lps; }
                }
            apspie_ctx_set_option(c, "batch_blocks_per_cu", bb);
            apspie_ctx_set_option(c, "bfs_levels_per_sync", bl);
            pick(c, "batch_pull_div", {2, 4, 8}, bfsb);
            if (g->val)
                pick(c, "batch_sssp_delta", {0, 256, 1024}, [&] {
                    return apspie_sssp_batch(c, g, src.data(), batch, (float *)out,
                                             APSPIE_MEM_DEVICE);
                });
        }
    });
}

extern "C" int apspie_autotune_hash(apspie_ctx *c, const apspie_graph *g) {
    if (!c || !g) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        APSPIE_CHECK(g->m == g->n && g->m > 0, APSPIE_ERR_INVALID, "autotune needs a square graph");
        set_device(c);
        // TC (rows sorted): the medium / big tier bound
        uint64_t total = 0;
        if (apspie_tc(c, g, &total, nullptr, APSPIE_MEM_HOST) == APSPIE_OK)
            pick(c, "tc_medium_row", {256, 512, 1024, 2048},
                 [&] { return apspie_tc(c, g, &total, nullptr, APSPIE_MEM_HOST); });
        // SpGEMM A * A: the wave / block bounds jointly, then dense or global
        auto product = [&] {
            apspie_graph *p = nullptr;
            const int st = apspie_spgemm(c, g, g, &p);
            if (p) apspie_graph_destroy(p);
            return st;
        };
        // a product too large for one call (more than 2^31 entries, or beyond
        // the budget) is not tuned, nor one of more than 2^31 multiplications
        // (sum over rows of the squared row length: exact for a symmetric
        // A, the scale otherwise), whose every run would take about a second
        {
            std::vector<int32_t> rp((size_t)g->m + 1);
            copy_from_device(c, rp.data(), g->rowptr, sizeof(int32_t) * rp.size(),
                             APSPIE_MEM_HOST);
            sync(c);
            double prods = 0;
            for (int32_t v = 0; v < g->m; ++v) {
                const double d = (double)rp[(size_t)v + 1] - rp[v];
                prods += d * d;
            }
            if (prods > 2147483648.0) return;
        }
        const int first = product();
        if (first == APSPIE_ERR_OVERFLOW || first == APSPIE_ERR_OOM || first == APSPIE_ERR_BUDGET)
            return;
        APSPIE_CHECK(first == APSPIE_OK, first, "autotune run failed: " + c->last_error);
        double best = 1e300;
        int64_t bw = 256, bk = 4096;
        for (int64_t wub : {64, 128, 256})
            for (int64_t bub : {1024, 2048, 4096}) {
                apspie_ctx_set_option(c, "spgemm_wave_ub", wub);
                apspie_ctx_set_option(c, "spgemm_block_ub", bub);
                const double t = median_ms(c, product);
                if (t < best) { best = t; bw = wub; bk = bub; }
            }
        apspie_ctx_set_option(c, "spgemm_wave_ub", bw);
        apspie_ctx_set_option(c, "spgemm_block_ub", bk);
        if (g->n <= 32768) pick(c, "spgemm_dense", {0, 1}, product);
    });
}

// One line per (GPU, graph class): "gpu=<key> class=<key> name=value ...".
extern "C" int apspie_tune_save(apspie_ctx *c, const apspie_graph *g, const char *path) {
    if (!c || !g || !path) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        const std::string gpu = tune_gpu_key(c), cls = tune_class_key(c, const_cast<Graph *>(
                                                       static_cast<const Graph *>(g)));
        std::vector<std::string> keep;
        {
            std::ifstream in(path);
            std::string line;
            while (std::getline(in, line)) {
                std::istringstream ls(line);
                std::string a, b;
                ls >> a >> b;
                if (a == "gpu=" + gpu && b == "class=" + cls) continue;  // replaced below
                if (!line.empty()) keep.push_back(line);
            }
        }
        std::string mine = "gpu=" + gpu + " class=" + cls;
        for (const char *name : kTuned) {
            int64_t v = 0;
            APSPIE_CHECK(apspie_ctx_get_option(c, name, &v) == APSPIE_OK, APSPIE_ERR_INVALID,
                         std::string("no option ") + name);
            mine += std::string(" ") + name + "=" + std::to_string(v);
        }
        keep.push_back(mine);
        const std::string tmp = std::string(path) + ".tmp";
        {
            std::ofstream out(tmp, std::ios::trunc);
            APSPIE_CHECK(out.good(), APSPIE_ERR_INVALID, std::string("cannot write ") + tmp);
            for (const auto &l : keep) out << l << "\n";
            APSPIE_CHECK(out.good(), APSPIE_ERR_INVALID, std::string("write failed: ") + tmp);
        }
        APSPIE_CHECK(std::rename(tmp.c_str(), path) == 0, APSPIE_ERR_INVALID,
                     std::string("cannot replace ") + path);
    });
}

extern "C" int apspie_tune_load(apspie_ctx *c, const apspie_graph *g, const char *path,
                                int32_t *found) {
    if (!c || !g || !path || !found) return APSPIE_ERR_INVALID;
    *found = 0;
    return guard(c, [&] {
        set_device(c);
        std::ifstream in(path);
        if (!in.good()) return;  // no table yet: nothing applied
        const std::string gpu = tune_gpu_key(c), cls = tune_class_key(c, const_cast<Graph *>(
                                                       static_cast<const Graph *>(g)));
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string a, b, kv;
            ls >> a >> b;
            if (a != "gpu=" + gpu || b != "class=" + cls) continue;
            std::vector<std::pair<std::string, int64_t>> opts;
            while (ls >> kv) {
                const size_t eq = kv.find('=');
                APSPIE_CHECK(eq != std::string::npos && eq > 0, APSPIE_ERR_INVALID,
                             "malformed tuning entry: " + kv);
                char *end = nullptr;
                const long long v = std::strtoll(kv.c_str() + eq + 1, &end, 10);
                APSPIE_CHECK(end && *end == '\0', APSPIE_ERR_INVALID,
                             "malformed tuning value: " + kv);
                opts.emplace_back(kv.substr(0, eq), (int64_t)v);
            }
            // validate every value before applying any
            for (const auto &o : opts) {
                int64_t cur = 0;
                APSPIE_CHECK(apspie_ctx_get_option(c, o.first.c_str(), &cur) == APSPIE_OK,
                             APSPIE_ERR_INVALID, "unknown option in tuning table: " + o.first);
            }
            std::vector<std::pair<std::string, int64_t>> before;
            for (const auto &o : opts) {
                int64_t cur = 0;
                apspie_ctx_get_option(c, o.first.c_str(), &cur);
                before.emplace_back(o.first, cur);
                if (apspie_ctx_set_option(c, o.first.c_str(), o.second) != APSPIE_OK) {
                    for (const auto &b0 : before) apspie_ctx_set_option(c, b0.first.c_str(), b0.second);
                    throw Error{APSPIE_ERR_INVALID, "invalid value in tuning table: " + o.first};
                }
            }
            *found = 1;
            return;
        }
    });
}
