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

// ---------------------------------------------------------------- near-far
// Near-far rounds (Davidson et al.'s two-pile form of delta-stepping): a
// relaxation that lowers dist[v] below the threshold T puts v in the next
// frontier as above; one that leaves it at or above T only marks v pending.
// When a round finds the frontier empty it is a split round instead: T grows
// by delta and a sweep over all vertices moves the pending ones now below T
// into the frontier.  Vertices are expanded fewer times (mostly once, near
// their final distance); the fixed point is the same.
//
// ctl words (every word a round reads was written by the previous round, and
// no word a round reads is written in that round): [0..2] frontier counter
// ring, [4..6] stats, [12 + (r & 1)] T of round r (float bits), [14 + (r & 1)]
// 1 if round r - 1 was a split, and the pending-vertex count left by the
// split of round r in the spread slots [16 + 16 i + r % 3].
constexpr int kNfBlock = 1024;
constexpr int kNfPer = 8;  // vertices per thread per sweep pass

struct SsspNearOp {
    const float *w;
    int32_t *dist;  // float bits
    int32_t *stamp;
    uint8_t *pend;
    int32_t round;  // index of the frontier being built
    int32_t tbits;  // threshold T (float bits; distances are non-negative)
    __device__ __forceinline__ bool edge(int32_t u, int32_t e, int32_t v) const {
        const float du = dev::bitsf(dist[u]);
        const int32_t nd = dev::fbits(du + w[e]);
        if (nd >= dist[v]) return false;
        const int32_t old = dev::atomic_min_dev(dist + v, nd);
        if (nd >= old) return false;
        if (nd >= tbits) {  // far: expanded by a later split round
            pend[v] = 1;
            return false;
        }
        pend[v] = 0;  // near now (a racing far store only costs a re-expansion)
        return dev::atomic_exch_dev(stamp + v, round) != round;
    }
};

__device__ __forceinline__ uint64_t *pend_slot(uint64_t *ctl, int r) {
    return ctl + 16 + (blockIdx.x % kSpread) * 16 + r % 3;
}

// two 1024-thread blocks per CU resident: at most 64 VGPRs
template <int TILE>
__global__ __launch_bounds__(kNfBlock) __attribute__((amdgpu_waves_per_eu(8))) void
k_sssp_nf_round(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                const float *w, int32_t nnz, int32_t m, int32_t *dist, int32_t *stamp,
                uint8_t *pend, float delta, FrontierBufs cur, FrontierBufs nxt, uint64_t *ctl,
                int32_t r) {
    __shared__ ExpandLds<kNfBlock, TILE, true> lds;
    __shared__ uint64_t s_part[kNfBlock / 64];
    const uint64_t cc = ctl[r % 3];
    const int32_t F = (int32_t)cnt_entries(cc), E = (int32_t)cnt_edges(cc);
    const int32_t tbits = (int32_t)ctl[12 + (r & 1)];
    const bool split_last = ctl[14 + (r & 1)] != 0;
    uint64_t left_before = 0;  // pending count left by the previous round's split
    for (int i = 0; i < kSpread; ++i) left_before += ctl[16 + i * 16 + (r + 2) % 3];
    const bool split = F == 0;
    const bool done = split && split_last && left_before == 0;
    const int32_t tnext = split && !done ? dev::fbits(dev::bitsf(tbits) + delta) : tbits;
    __syncthreads();  // every thread has read the control words
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[(r + 2) % 3] = 0;
        ctl[12 + ((r + 1) & 1)] = (uint64_t)(uint32_t)tnext;
        ctl[14 + ((r + 1) & 1)] = split ? 1 : 0;
        for (int i = 0; i < kSpread; ++i) ctl[16 + i * 16 + (r + 1) % 3] = 0;
        if (F > 0) {
            ctl[4] += (uint64_t)E;
            ctl[5] += (uint64_t)F;
            ctl[6] += 1;
        }
    }
    if (done) return;
    if (!split) {
        SsspNearOp op{w, dist, stamp, pend, r + 1, tbits};
        expand_tiles<kNfBlock, TILE, true, true, true>(lds, op, rowptr, col, nnz, cur, F, E, nxt,
                                                       ctl + ((r + 1) % 3));
        return;
    }
    // split: move the pending vertices below the new threshold into the frontier
    uint64_t left = 0;
    for (int64_t pbase = (int64_t)blockIdx.x * kNfBlock * kNfPer; pbase < m;
         pbase += (int64_t)gridDim.x * kNfBlock * kNfPer) {
        int32_t cv[kNfPer], crs[kNfPer], cd[kNfPer];
        bool take[kNfPer];
#pragma unroll
        for (int k = 0; k < kNfPer; ++k) {
            const int64_t v = pbase + (int64_t)k * kNfBlock + threadIdx.x;
            cv[k] = (int32_t)v;
            crs[k] = 0;
            cd[k] = 0;
            take[k] = false;
            if (v < m && pend[v]) {
                if (dist[v] < tnext) take[k] = true;
                else ++left;
            }
        }
#pragma unroll
        for (int k = 0; k < kNfPer; ++k)
            if (take[k]) {
                pend[cv[k]] = 0;
                stamp[cv[k]] = r + 1;
                crs[k] = rowptr[cv[k]];
                cd[k] = rowptr[cv[k] + 1] - crs[k];
            }
        produce<kNfBlock, kNfPer, TILE, true>(cv, crs, cd, nxt, ctl + ((r + 1) % 3), lds.scan,
                                              &lds.base);
    }
    const uint64_t ws = dev::wave_sum_u64(left);
    if (dev::lane_id() == 0) s_part[threadIdx.x / 64] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int q = 0; q < kNfBlock / 64; ++q) t += s_part[q];
        if (t) dev::atomic_add_dev(pend_slot(ctl, r), t);
    }
}

template <int TILE>
void sssp_near_far(Ctx *c, const Graph *g, int32_t source, float *dist_out, float delta) {
    FrontierSet fs = frontier_alloc(c, "sssp", g->m, g->nnz, TILE, true);
    int32_t *dist = (int32_t *)dist_out;
    int32_t *stamp = (int32_t *)scratch(c, "sssp_stamp", sizeof(int32_t) * g->m);
    uint8_t *pend = (uint8_t *)scratch(c, "sssp_pend", (size_t)g->m + 4);
    fill_i32(c, dist, __builtin_bit_cast(int32_t, kInf), g->m);
    fill_i32(c, stamp, -1, g->m);
    fill_i32(c, (int32_t *)pend, 0, ceil_div(g->m + 4, 4));
    fill_i32(c, (int32_t *)fs.ctl, 0, 2 * kCtlWords);
    k_sssp_seed<TILE><<<1, 256, 0, c->stream>>>(g->rowptr, source, dist, stamp, fs.f[0], fs.ctl);
    fill_i32(c, (int32_t *)(fs.ctl + 12), __builtin_bit_cast(int32_t, delta), 1);  // T of round 0
    APSPIE_LAUNCH_CHECK();
    const int grid = fill_grid(
        c, kNfBlock, (int)std::max<int64_t>(1, opt(c, "bfs_blocks_per_cu", 8) * 256 / kNfBlock));
    int chunk = (int)std::max<int64_t>(1, opt(c, "bfs_levels_per_sync", 8));
    KernelTimer kt(c);
    uint64_t *h = (uint64_t *)c->pinned;
    int r = 0;
    for (;;) {
        kt.begin();
        for (int k = 0; k < chunk; ++k, ++r) {
            k_sssp_nf_round<TILE><<<grid, kNfBlock, 0, c->stream>>>(
                g->rowptr, g->col, g->val, (int32_t)g->nnz, g->m, dist, stamp, pend, delta,
                fs.f[r & 1], fs.f[(r + 1) & 1], fs.ctl, r);
            APSPIE_LAUNCH_CHECK();
        }
        kt.end(chunk);
        APSPIE_HIP(hipMemcpyAsync(h, fs.ctl, kCtlWords * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        sync(c);
        c->stats.host_syncs++;
        // finished when the last round was a split that left nothing pending
        // and found nothing below the new threshold
        uint64_t left = 0;
        for (int i = 0; i < kSpread; ++i) left += h[16 + i * 16 + (r + 2) % 3];
        if (cnt_entries(h[r % 3]) == 0 && h[14 + (r & 1)] != 0 && left == 0) break;
        APSPIE_CHECK(r <= 8 * g->m + 1024, APSPIE_ERR_DEVICE, "SSSP did not converge");
        chunk = std::min(chunk * 2, 256);
    }
    c->stats.edges_traversed = (int64_t)h[4];
    c->stats.vertices_reached = (int64_t)h[5];
    c->stats.iterations = (int32_t)h[6];
    kt.collect();
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
    const int64_t delta = opt(c, "sssp_delta", 0);
    if (delta > 0) {
        switch (opt(c, "bfs_tile", 2048)) {
        case 1024: sssp_near_far<1024>(c, g, source, dist, (float)delta); return;
        case 4096: sssp_near_far<4096>(c, g, source, dist, (float)delta); return;
        default: sssp_near_far<2048>(c, g, source, dist, (float)delta); return;
        }
    }
    switch (opt(c, "bfs_tile", 2048)) {
    case 1024: sssp_rounds<1024>(c, g, source, dist); break;
    case 4096: sssp_rounds<4096>(c, g, source, dist); break;
    default: sssp_rounds<2048>(c, g, source, dist); break;
    }
}

}  // namespace apspie
