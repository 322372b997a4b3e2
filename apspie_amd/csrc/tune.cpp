// Autotuner for the traversal launch parameters.
//
// No reference counterpart (the reference hard-codes NTHREADS 512 and the
// mgpu LaunchBox tables, spgemm.cuh:181-185).  For one resident graph it times
// the candidate settings on a few sources and keeps the fastest in the
// context's option table (readable with apspie_ctx_get_option, so callers can
// persist them per GPU and graph class):
//   batch == 0: single-source BFS over tile size x blocks per CU, then the
//               direction switch threshold alpha (push -> pull), and for a
//               weighted graph the SSSP near-far width sssp_delta (Bellman-Ford,
//               and half, once and twice the width chosen from the weights);
//   batch  > 0: batched BFS over blocks per CU x levels per hipGraph chunk,
//               then the pull threshold batch_pull_div, and for a weighted
//               graph the batched SSSP threshold step batch_sssp_delta
//               (Bellman-Ford, 256, 1024).
#include <algorithm>
#include <random>
#include <vector>

#include "common.hpp"

extern "C" int apspie_autotune(apspie_ctx *c, const apspie_graph *g, int32_t batch) {
    if (!c || !g || batch < 0) return APSPIE_ERR_INVALID;
    return apspie::guard(c, [&] {
        APSPIE_CHECK(g->m == g->n && g->m > 0, APSPIE_ERR_INVALID, "autotune needs a square graph");
        std::vector<int32_t> rowptr(g->m + 1);
        APSPIE_HIP(hipMemcpy(rowptr.data(), g->rowptr, sizeof(int32_t) * (g->m + 1),
                             hipMemcpyDeviceToHost));
        // probe sources: vertices with the largest degrees among a fixed sample
        std::mt19937 gen(12345);
        std::vector<int32_t> src;
        for (int tries = 0; tries < 100000 && (int)src.size() < std::max(3, batch); ++tries) {
            const int32_t v = (int32_t)(gen() % (uint32_t)g->m);
            if (rowptr[v + 1] > rowptr[v]) src.push_back(v);
        }
        APSPIE_CHECK(!src.empty(), APSPIE_ERR_INVALID, "graph has no edges");
        auto time_of = [&](auto run) {
            std::vector<double> t;
            for (int rep = 0; rep < 3; ++rep) {
                int st = run();
                if (st != APSPIE_OK) throw apspie::Error{st, "autotune run failed"};
                t.push_back(c->stats.elapsed_ms);
            }
            std::sort(t.begin(), t.end());
            return t[t.size() / 2];
        };
        int32_t *out = nullptr;
        const size_t cols = batch > 0 ? (size_t)batch : 1;
        APSPIE_HIP(hipMalloc(&out, sizeof(int32_t) * (size_t)g->m * cols));
        try {
            if (batch == 0) {
                double best = 1e300;
                int64_t bt = 2048, bb = 4;
                for (int64_t tile : {1024, 2048, 4096})
                    for (int64_t bpc : {2, 4, 8}) {
                        apspie_ctx_set_option(c, "bfs_tile", tile);
                        apspie_ctx_set_option(c, "bfs_blocks_per_cu", bpc);
                        double tot = 0;
                        for (size_t i = 0; i < std::min<size_t>(3, src.size()); ++i)
                            tot += time_of([&] {
                                return apspie_bfs(c, g, src[i], out, APSPIE_MEM_DEVICE);
                            });
                        if (tot < best) { best = tot; bt = tile; bb = bpc; }
                    }
                apspie_ctx_set_option(c, "bfs_tile", bt);
                apspie_ctx_set_option(c, "bfs_blocks_per_cu", bb);
                int64_t ba = 14;
                best = 1e300;
                for (int64_t alpha : {7, 14, 28}) {
                    apspie_ctx_set_option(c, "bfs_alpha", alpha);
                    double tot = 0;
                    for (size_t i = 0; i < std::min<size_t>(3, src.size()); ++i)
                        tot += time_of([&] {
                            return apspie_bfs(c, g, src[i], out, APSPIE_MEM_DEVICE);
                        });
                    if (tot < best) { best = tot; ba = alpha; }
                }
                apspie_ctx_set_option(c, "bfs_alpha", ba);
                if (g->val) {
                    const int64_t d = (int64_t)std::min<double>(
                        std::max<double>(
                            apspie::sssp_weight_delta(c, const_cast<apspie_graph *>(g)), 64),
                        1 << 29);
                    int64_t bs = -1;
                    best = 1e300;
                    for (int64_t delta : {(int64_t)0, d / 2, d, 2 * d}) {
                        apspie_ctx_set_option(c, "sssp_delta", delta);
                        double tot = 0;
                        for (size_t i = 0; i < std::min<size_t>(3, src.size()); ++i)
                            tot += time_of([&] {
                                return apspie_sssp(c, g, src[i], (float *)out,
                                                   APSPIE_MEM_DEVICE);
                            });
                        if (tot < best) { best = tot; bs = delta; }
                    }
                    apspie_ctx_set_option(c, "sssp_delta", bs);
                }
            } else {
                double best = 1e300;
                int64_t bb = 4, bl = 12;
                src.resize((size_t)batch);
                for (int64_t bpc : {2, 4, 8})
                    for (int64_t lps : {6, 12, 48}) {
                        apspie_ctx_set_option(c, "batch_blocks_per_cu", bpc);
                        apspie_ctx_set_option(c, "bfs_levels_per_sync", lps);
                        const double t = time_of([&] {
                            return apspie_bfs_batch(c, g, src.data(), batch, out, APSPIE_MEM_DEVICE);
                        });
                        if (t < best) { best = t; bb = bpc; bl = lps; }
                    }
                apspie_ctx_set_option(c, "batch_blocks_per_cu", bb);
                apspie_ctx_set_option(c, "bfs_levels_per_sync", bl);
                int64_t bd = 4;
                best = 1e300;
                for (int64_t div : {2, 4, 8}) {
                    apspie_ctx_set_option(c, "batch_pull_div", div);
                    const double t = time_of([&] {
                        return apspie_bfs_batch(c, g, src.data(), batch, out, APSPIE_MEM_DEVICE);
                    });
                    if (t < best) { best = t; bd = div; }
                }
                apspie_ctx_set_option(c, "batch_pull_div", bd);
                if (g->val) {
                    int64_t bs = -1;
                    best = 1e300;
                    for (int64_t delta : {0, 256, 1024}) {
                        apspie_ctx_set_option(c, "batch_sssp_delta", delta);
                        const double t = time_of([&] {
                            return apspie_sssp_batch(c, g, src.data(), batch, (float *)out,
                                                     APSPIE_MEM_DEVICE);
                        });
                        if (t < best) { best = t; bs = delta; }
                    }
                    apspie_ctx_set_option(c, "batch_sssp_delta", bs);
                }
            }
        } catch (...) {
            (void)hipFree(out);
            throw;
        }
        APSPIE_HIP(hipFree(out));
    });
}
