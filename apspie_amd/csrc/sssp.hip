// Single-source shortest paths as repeated (min,+) SpMSpV.
//
// Semantics: sssp() (sssp.cuh:27-115) with spmspvSssp() (spmspvSssp.cuh:91-291):
// dist[src] = 0, dist[v] = min over paths of the FP32 weight sums, unreached
// = 1.70141e+38 (sssp.cuh:48-49).  The reference multiplies with unit weights
// after the first round (sssp.cuh:83,91) and stops after a fixed depth; here
// every round uses the real weights and the loop runs to the fixed point, so
// the result equals the CPU Dijkstra of testSssp.cpp:33-115.
//
// One kernel per round: the frontier holds the vertices whose distance went
// down in the previous round; each expanded edge relaxes its head with an
// atomic min on the distance bits (non-negative floats order like their bit
// patterns), and the thread whose relaxation lowered a vertex first in this
// round appends it to the next frontier (per-vertex round stamp).  With
// integer-valued weights every sum is exact, so the fixed point does not
// depend on the relaxation order.
#include "common.hpp"
#include "device.hpp"
#include "frontier.hpp"
#include "timing.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;
constexpr float kInf = APSPIE_SSSP_INF;

struct SsspOp {
    const float *w;
    int32_t *dist;  // float bits
    int32_t *stamp;
    int32_t round;  // index of the frontier being built
    __device__ __forceinline__ bool edge(int32_t u, int32_t e, int32_t v) const {
        const float du = dev::bitsf(dist[u]);
        const int32_t nd = dev::fbits(du + w[e]);
        if (nd >= dist[v]) return false;  // not better than what we see
        const int32_t old = dev::atomic_min_dev(dist + v, nd);
        if (nd >= old) return false;
        return dev::atomic_exch_dev(stamp + v, round) != round;
    }
};

template <int TILE>
__global__ void k_sssp_seed(const int32_t *rowptr, int32_t src, int32_t *dist, int32_t *stamp,
                            FrontierBufs f0, uint64_t *ctl) {
    const int32_t rs = rowptr[src];
    const int32_t d = rowptr[src + 1] - rs;
    if (threadIdx.x == 0) {
        dist[src] = dev::fbits(0.0f);
        stamp[src] = 0;
        ctl[0] = d > 0 ? ((uint64_t(1) << 32) | uint64_t(d)) : 0;
        ctl[1] = 0;
        ctl[2] = 0;
        for (int i = 4; i < 12; ++i) ctl[i] = 0;
        f0.v[0] = src;
        f0.start[0] = rs;
        f0.off[0] = 0;
    }
    for (int32_t t = threadIdx.x; (int64_t)t * TILE < d; t += blockDim.x) f0.tfirst[t] = 0;
}

template <int TILE>
void sssp_rounds(Ctx *c, const Graph *g, int32_t source, float *dist_out) {
    FrontierSet fs = frontier_alloc(c, "sssp", g->m, g->nnz, TILE, true);
    int32_t *dist = (int32_t *)dist_out;
    int32_t *stamp = (int32_t *)scratch(c, "sssp_stamp", sizeof(int32_t) * g->m);
    fill_i32(c, dist, __builtin_bit_cast(int32_t, kInf), g->m);
    fill_i32(c, stamp, -1, g->m);
    k_sssp_seed<TILE><<<1, 256, 0, c->stream>>>(g->rowptr, source, dist, stamp, fs.f[0], fs.ctl);
    APSPIE_LAUNCH_CHECK();

    const int grid = fill_grid(c, kBlock, (int)opt(c, "bfs_blocks_per_cu", 4));
    int chunk = (int)std::max<int64_t>(1, opt(c, "bfs_levels_per_sync", 8));
    KernelTimer kt(c);
    int round = 0;
    for (;;) {
        kt.begin();
        for (int k = 0; k < chunk; ++k, ++round) {
            const int cur = round & 1;
            k_expand<kBlock, TILE, true, true, true><<<grid, kBlock, 0, c->stream>>>(
                SsspOp{g->val, dist, stamp, round + 1}, g->rowptr, g->col, (int32_t)g->nnz,
                fs.f[cur], fs.ctl + (round % 3), fs.f[cur ^ 1], fs.ctl + ((round + 1) % 3),
                fs.ctl + ((round + 2) % 3), fs.ctl + 4);
            APSPIE_LAUNCH_CHECK();
        }
        kt.end(chunk);
        uint64_t *h = (uint64_t *)c->pinned;
        APSPIE_HIP(hipMemcpyAsync(h, fs.ctl + (round % 3), sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        APSPIE_HIP(hipMemcpyAsync(h + 1, fs.ctl + 4, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        sync(c);
        c->stats.host_syncs++;
        if (cnt_entries(h[0]) == 0) break;
        // each round lowers at least one distance; integer weights bound the rounds
        APSPIE_CHECK(round <= 4 * g->m + 64, APSPIE_ERR_DEVICE, "SSSP did not converge");
        chunk = std::min(chunk * 2, 256);
    }
    uint64_t *h = (uint64_t *)c->pinned;
    c->stats.edges_traversed = (int64_t)h[1];
    c->stats.vertices_reached = (int64_t)h[2];
    c->stats.iterations = (int32_t)h[3];
    kt.collect();
}

}  // namespace

void sssp_run(Ctx *c, const Graph *g, int32_t source, float *dist) {
    switch (opt(c, "bfs_tile", 2048)) {
    case 1024: sssp_rounds<1024>(c, g, source, dist); break;
    case 4096: sssp_rounds<4096>(c, g, source, dist); break;
    default: sssp_rounds<2048>(c, g, source, dist); break;
    }
}

}  // namespace apspie
