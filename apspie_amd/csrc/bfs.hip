// Breadth-first search as repeated Boolean SpMSpV with a visited mask.
//
// Semantics follow bfs() (bfs.cuh:32-125) / spmspvBfs() (spmspvBfs.cuh:96-243):
// depth[src] = 0, depth[v] = level of first discovery, -1 if unreached.  The
// reference runs expand -> sort -> reduce -> scatter -> addResult per level
// with host round trips; here one kernel per level does the whole SpMSpV:
// load-balanced expansion of the frontier, visited filter (Boolean semiring
// plus mask: the first writer of a bit wins), depth write and the append of
// the discovered vertices into the next frontier with their degree offsets.
#include "common.hpp"
#include "device.hpp"
#include "frontier.hpp"
#include "timing.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;

struct BfsOp {
    uint32_t *bitmap;
    int32_t *depth;
    int32_t nd;
    __device__ __forceinline__ bool edge(int32_t, int32_t, int32_t v) const {
        uint32_t *w = bitmap + (v >> 5);
        const uint32_t bit = 1u << (v & 31);
        if (*w & bit) return false;  // seen (possibly stale: the atomic decides)
        const uint32_t old = dev::atomic_or_dev(w, bit);
        if (old & bit) return false;
        depth[v] = nd;
        return true;
    }
};

template <int TILE>
__global__ void k_bfs_seed(const int32_t *rowptr, int32_t src, int32_t *depth,
                           uint32_t *bitmap, FrontierBufs f0, uint64_t *ctl) {
    const int32_t rs = rowptr[src];
    const int32_t d = rowptr[src + 1] - rs;
    if (threadIdx.x == 0) {
        depth[src] = 0;
        bitmap[src >> 5] = 1u << (src & 31);
        ctl[0] = d > 0 ? ((uint64_t(1) << 32) | uint64_t(d)) : 0;
        ctl[1] = 0;
        ctl[2] = 0;
        for (int i = 4; i < 12; ++i) ctl[i] = 0;
        f0.start[0] = rs;
        f0.off[0] = 0;
        if (f0.v) f0.v[0] = src;
    }
    for (int32_t t = threadIdx.x; (int64_t)t * TILE < d; t += blockDim.x) f0.tfirst[t] = 0;
}

template <int TILE>
void bfs_levels(Ctx *c, const Graph *g, int32_t source, int32_t *depth) {
    FrontierSet fs = frontier_alloc(c, "bfs", g->m, g->nnz, TILE, false);
    uint32_t *bitmap = (uint32_t *)scratch(c, "bfs_bitmap", sizeof(uint32_t) * (g->m / 32 + 1));
    fill_i32(c, depth, -1, g->m);
    fill_i32(c, (int32_t *)bitmap, 0, g->m / 32 + 1);
    k_bfs_seed<TILE><<<1, 256, 0, c->stream>>>(g->rowptr, source, depth, bitmap, fs.f[0], fs.ctl);
    APSPIE_LAUNCH_CHECK();

    const int grid = fill_grid(c, kBlock, (int)opt(c, "bfs_blocks_per_cu", 4));
    int chunk = (int)std::max<int64_t>(1, opt(c, "bfs_levels_per_sync", 8));
    KernelTimer kt(c);
    int level = 0;
    for (;;) {
        for (int k = 0; k < chunk; ++k, ++level) {
            const int cur = level & 1;
            kt.begin();
            k_expand<kBlock, TILE, false, false, true>
                <<<grid, kBlock, 0, c->stream>>>(BfsOp{bitmap, depth, level + 1}, g->rowptr, g->col,
                                                 (int32_t)g->nnz, fs.f[cur], fs.ctl + (level % 3), fs.f[cur ^ 1],
                                                 fs.ctl + ((level + 1) % 3),
                                                 fs.ctl + ((level + 2) % 3), fs.ctl + 4);
            kt.end();
            APSPIE_LAUNCH_CHECK();
        }
        uint64_t *h = (uint64_t *)c->pinned;
        APSPIE_HIP(hipMemcpyAsync(h, fs.ctl + (level % 3), sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        APSPIE_HIP(hipMemcpyAsync(h + 1, fs.ctl + 4, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        sync(c);
        c->stats.host_syncs++;
        if (cnt_entries(h[0]) == 0) break;
        APSPIE_CHECK(level <= g->m + 1, APSPIE_ERR_DEVICE, "BFS did not converge");
        chunk = std::min(chunk * 2, 256);
    }
    uint64_t *h = (uint64_t *)c->pinned;
    c->stats.edges_traversed = (int64_t)h[1];
    c->stats.vertices_reached = (int64_t)h[2];
    c->stats.iterations = (int32_t)h[3];
    kt.collect((int64_t)h[3]);
}

}  // namespace

void bfs_run(Ctx *c, const Graph *g, int32_t source, int32_t *depth) {
    switch (opt(c, "bfs_tile", 2048)) {
    case 1024: bfs_levels<1024>(c, g, source, depth); break;
    case 4096: bfs_levels<4096>(c, g, source, depth); break;
    default: bfs_levels<2048>(c, g, source, depth); break;
    }
}

}  // namespace apspie
