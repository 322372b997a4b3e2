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
// round appends it to the next frontier (the atomic OR that sets its bit in
// the next frontier's bitmap).
// With integer-valued weights every sum is exact, so the fixed point does not
// depend on the relaxation order.
#include "common.hpp"
#include "device.hpp"
#include "frontier.hpp"
#include "timing.hpp"

namespace apspie {

namespace {

constexpr float kInf = APSPIE_SSSP_INF;

template <int TILE>
__global__ void k_sssp_seed(const int32_t *rowptr, int32_t src, int32_t *dist, FrontierBufs f0,
                            uint64_t *ctl) {
    const int32_t rs = rowptr[src];
    const int32_t d = rowptr[src + 1] - rs;
    if (threadIdx.x == 0) {
        dist[src] = dev::fbits(0.0f);
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

// -------------------------------------------------------- pull rounds
// Direction-optimising Bellman-Ford for a CSR equal to its transpose (values
// included): a round whose frontier holds more than nnz / sssp_pull_div edges
// is pulled instead of pushed -- every vertex takes the minimum of
// dist[u] + w(v, u) over its neighbours u in the frontier (a bitmap), which
// is the same set of relaxations as pushing the frontier, with far fewer
// atomics.  Rows of at most kLight entries are scanned by one lane (8
// vertices per thread, 8 loads in flight) and written by that lane alone.
// Longer rows are walked in tiles of kHvTile entries over a list of all such
// vertices built once per graph: each thread takes kHvPer consecutive
// entries (at most two rows, since every long row has more than kHvPer
// entries), the minimum per row is reduced in LDS and applied with one atomic
// min per tile and row.  The next frontier's bitmap doubles as the claim: the
// thread whose atomic OR sets a vertex's bit appends it to the next queue.
constexpr int kDoBlock = 1024;
constexpr int kDoTile = 2048;  // tiles of the pushed queue
constexpr int32_t kLight = 32;
constexpr int kDoPer = 8;
constexpr int kDoIlp = 8;
constexpr int kHvPer = 8;
constexpr int kHvTile = kDoBlock * kHvPer;
constexpr int kHvEnt = kHvTile / (kLight + 1) + 3;  // rows meeting one tile, with slack
static_assert(kHvEnt < kDoBlock, "one claim per thread");
static_assert(kHvPer <= kLight, "a thread's entries span at most two rows");

__device__ __forceinline__ bool claim_bit(uint32_t *bits, int32_t v) {
    const uint32_t bit = 1u << (v & 31);
    return !(dev::atomic_or_dev(bits + (v >> 5), bit) & bit);
}

// A relaxation that beats the distance it read applies its atomic min without
// waiting for the result: if a racing thread of the same round lowered v
// further, v still improved in this round and belongs to the next queue, so
// the claim (the bitmap OR) decides the append alone.
struct SsspPushOp {
    static constexpr bool kStreamColumns = true;  // columns and weights stream past the L2
    const float *w;
    int32_t *dist;  // float bits
    uint32_t *fb_next;
    __device__ __forceinline__ bool edge(int32_t u, int32_t e, int32_t v) const {
        const int32_t nd = dev::fbits(dev::bitsf(dist[u]) + __builtin_nontemporal_load(w + e));
        if (nd >= dist[v]) return false;
        __hip_atomic_fetch_min(dist + v, nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return claim_bit(fb_next, v);
    }
};

struct HeavyLds {
    int32_t off[kHvEnt + 1], start[kHvEnt], v[kHvEnt], best[kHvEnt];
    uint64_t scan[kDoBlock / 64 + 1];
    uint64_t base;
};

union DoLds {
    ExpandLds<kDoBlock, kDoTile, true> ex;
    HeavyLds hv;
};

struct HeavyOp {
    const int32_t *rowptr;
    __device__ __forceinline__ bool keep(int32_t v) const {
        return rowptr[v + 1] - rowptr[v] > kLight;
    }
};

__global__ void k_do_seed(const int32_t *rowptr, int32_t src, int32_t *dist, uint32_t *fb0,
                          FrontierBufs f0, uint64_t *ctl) {
    const int32_t rs = rowptr[src];
    const int32_t d = rowptr[src + 1] - rs;
    if (threadIdx.x == 0) {
        dist[src] = dev::fbits(0.0f);
        fb0[src >> 5] = 1u << (src & 31);
        ctl[0] = d > 0 ? ((uint64_t(1) << 32) | uint64_t(d)) : 0;
        f0.v[0] = src;
        f0.start[0] = rs;
        f0.off[0] = 0;
    }
    for (int32_t t = threadIdx.x; (int64_t)t * kDoTile < d; t += blockDim.x) f0.tfirst[t] = 0;
}

// two 1024-thread blocks per CU resident: at most 64 VGPRs
__global__ __launch_bounds__(kDoBlock) __attribute__((amdgpu_waves_per_eu(8))) void
k_sssp_do_round(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                const float *__restrict__ w, int32_t nnz, int32_t m, int32_t *dist,
                const uint32_t *__restrict__ fb_cur, uint32_t *fb_next, uint32_t *fb_clear,
                FrontierBufs cur, FrontierBufs nxt, FrontierBufs heavy, uint64_t heavy_cnt,
                int32_t pull_div, uint64_t *ctl, int32_t r) {
    __shared__ DoLds lds;
    const uint64_t cc = ctl[r % 3];
    const int32_t F = (int32_t)cnt_entries(cc), E = (int32_t)cnt_edges(cc);
    const bool pull = pull_div > 0 && (int64_t)E * pull_div > (int64_t)nnz;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[(r + 2) % 3] = 0;  // the counter of round r + 1's output
        if (F > 0) {
            ctl[4] += (uint64_t)E;
            ctl[5] += (uint64_t)F;
            ctl[6] += 1;
            ctl[7] += pull ? (uint64_t)nnz : (uint64_t)E;  // CSR entries read
        }
    }
    if (F == 0) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m / 32 + 1;
         i += (int64_t)gridDim.x * blockDim.x)
        fb_clear[i] = 0;
    uint64_t *cnt_next = ctl + (r + 1) % 3;
    if (!pull) {
        SsspPushOp op{w, dist, fb_next};
        expand_tiles<kDoBlock, kDoTile, true, true, true>(lds.ex, op, rowptr, col, nnz, cur, F,
                                                          E, nxt, cnt_next);
        return;
    }
    const int lane = dev::lane_id();
    // short rows: one lane per vertex
    for (int64_t pbase = (int64_t)blockIdx.x * kDoBlock * kDoPer; pbase < m;
         pbase += (int64_t)gridDim.x * kDoBlock * kDoPer) {
        int32_t crs[kDoPer], cd[kDoPer], best[kDoPer], d0[kDoPer];
        uint32_t pend = 0;
#pragma unroll
        for (int k = 0; k < kDoPer; ++k) {
            const int64_t v = pbase + (int64_t)k * kDoBlock + threadIdx.x;
            const bool in = v < m;
            crs[k] = in ? rowptr[v] : 0;
            cd[k] = in ? rowptr[v + 1] - crs[k] : 0;
            d0[k] = in ? dist[v] : 0;
            best[k] = d0[k];
            if (cd[k] > 0 && cd[k] <= kLight) pend |= 1u << k;
        }
        // each lane walks its short rows one after another, kDoIlp loads in flight
        {
            int32_t e = 0, lim = 0, kk = -1, bk = 0;
            for (;;) {
                if (e >= lim) {
                    if (kk >= 0) {
#pragma unroll
                        for (int k = 0; k < kDoPer; ++k)
                            if (k == kk) best[k] = bk;
                    }
                    if (!pend) break;
                    kk = __ffs(pend) - 1;
                    pend &= pend - 1;
#pragma unroll
                    for (int k = 0; k < kDoPer; ++k)
                        if (k == kk) {
                            e = crs[k];
                            lim = crs[k] + cd[k];
                            bk = best[k];
                        }
                }
                int32_t u[kDoIlp];
#pragma unroll
                for (int q = 0; q < kDoIlp; ++q) u[q] = e + q < lim ? col[e + q] : -1;
                bool f[kDoIlp];
#pragma unroll
                for (int q = 0; q < kDoIlp; ++q)
                    f[q] = u[q] >= 0 && ((fb_cur[u[q] >> 5] >> (u[q] & 31)) & 1u);
#pragma unroll
                for (int q = 0; q < kDoIlp; ++q)
                    if (f[q]) bk = min(bk, dev::fbits(dev::bitsf(dist[u[q]]) + w[e + q]));
                e += kDoIlp;
            }
        }
        int32_t pc[kDoPer], prs[kDoPer], pcd[kDoPer];
#pragma unroll
        for (int k = 0; k < kDoPer; ++k) {
            const int32_t v = (int32_t)(pbase + (int64_t)k * kDoBlock + threadIdx.x);
            const bool imp = cd[k] > 0 && cd[k] <= kLight && best[k] < d0[k];
            pc[k] = v;
            prs[k] = imp ? crs[k] : 0;
            pcd[k] = imp ? cd[k] : 0;
            if (imp) dist[v] = best[k];
            const uint64_t bm = __ballot(imp);
            const int64_t wb = (pbase + (int64_t)k * kDoBlock + (threadIdx.x & ~63)) >> 5;
            if (lane == 0 && (uint32_t)bm) dev::atomic_or_dev(fb_next + wb, (uint32_t)bm);
            if (lane == 32 && (uint32_t)(bm >> 32))
                dev::atomic_or_dev(fb_next + wb + 1, (uint32_t)(bm >> 32));
        }
        produce<kDoBlock, kDoPer, kDoTile, true>(pc, prs, pcd, nxt, cnt_next, lds.hv.scan,
                                                 &lds.hv.base);
    }
    // long rows: tiles over the long-row list
    const int32_t HF = (int32_t)cnt_entries(heavy_cnt), HE = (int32_t)cnt_edges(heavy_cnt);
    if (HF == 0) return;
    HeavyLds &L = lds.hv;
    const int32_t ntiles = (HE + kHvTile - 1) / kHvTile;
    for (int32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int32_t i0 = min(max(heavy.tfirst[t], 0), HF - 1);
        const int32_t i1 = (t + 1 < ntiles) ? min(max(heavy.tfirst[t + 1], i0), HF - 1) : HF - 1;
        const int32_t nv = min(i1 - i0 + 1, kHvEnt);
        // off[nv]: where the row after the tile's last one starts (or HE)
        for (int32_t j = threadIdx.x; j <= nv; j += kDoBlock) {
            L.off[j] = i0 + j < HF ? heavy.off[i0 + j] : HE;
            if (j < nv) {
                L.start[j] = heavy.start[i0 + j];
                L.v[j] = heavy.v[i0 + j];
                L.best[j] = 0x7fffffff;
            }
        }
        __syncthreads();
        const int32_t x0 = t * kHvTile + threadIdx.x * kHvPer;
        if (x0 < HE) {
            const int j = dev::upper_bound_m1(L.off, 0, nv, x0);
            const int32_t split = L.off[j + 1];  // first entry of row j + 1
            int32_t idx[kHvPer], u[kHvPer];
#pragma unroll
            for (int k = 0; k < kHvPer; ++k) {
                const int32_t x = x0 + k;
                const int jj = x >= split ? j + 1 : j;
                idx[k] = -1;
                if (x < HE && jj < nv) idx[k] = L.start[jj] + (x - L.off[jj]);
                if ((uint32_t)idx[k] >= (uint32_t)nnz) idx[k] = -1;  // guard only
                u[k] = idx[k] >= 0 ? col[idx[k]] : -1;
            }
            bool f[kHvPer];
#pragma unroll
            for (int k = 0; k < kHvPer; ++k)
                f[k] = u[k] >= 0 && ((fb_cur[u[k] >> 5] >> (u[k] & 31)) & 1u);
            int32_t b0 = 0x7fffffff, b1 = 0x7fffffff;
#pragma unroll
            for (int k = 0; k < kHvPer; ++k)
                if (f[k]) {
                    const int32_t c = dev::fbits(dev::bitsf(dist[u[k]]) + w[idx[k]]);
                    if (x0 + k >= split) b1 = min(b1, c);
                    else b0 = min(b0, c);
                }
            if (b0 != 0x7fffffff) atomicMin(&L.best[j], b0);
            if (b1 != 0x7fffffff) atomicMin(&L.best[j + 1], b1);
        }
        __syncthreads();
        int32_t hc[1] = {0}, hrs[1] = {0}, hcd[1] = {0};
        if (threadIdx.x < nv && L.best[threadIdx.x] != 0x7fffffff) {
            const int32_t j = threadIdx.x, v = L.v[j], b = L.best[j];
            if (b < dev::atomic_min_dev(dist + v, b) && claim_bit(fb_next, v)) {
                hc[0] = v;
                hrs[0] = L.start[j];
                hcd[0] = L.off[j + 1] - L.off[j];
            }
        }
        produce<kDoBlock, 1, kDoTile, true>(hc, hrs, hcd, nxt, cnt_next, L.scan, &L.base);
        __syncthreads();  // the tile's LDS is refilled next
    }
}

// The long-row list of a graph (rows longer than kLight, with tile owners
// every kHvTile entries), built once and kept with the graph.
void graph_prepare_heavy(Ctx *c, Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    if (g->heavy_mem || g->m == 0) return;
    const int64_t ent = g->nnz / (kLight + 1) + 2, tiles = g->nnz / kHvTile + 2;
    const size_t bytes = sizeof(int32_t) * (3 * ent + tiles) + 64;
    // owned by the graph (freed with it), like its other derived arrays
    APSPIE_HIP(hipMalloc(&g->heavy_mem, bytes));
    int32_t *p = (int32_t *)g->heavy_mem;
    g->heavy_v = p;
    g->heavy_start = p + ent;
    g->heavy_off = p + 2 * ent;
    g->heavy_tfirst = p + 3 * ent;
    uint64_t *cnt = (uint64_t *)scratch(c, "sssp_heavy_cnt", 64);
    fill_i32(c, (int32_t *)cnt, 0, 2);
    FrontierBufs hb{g->heavy_v, g->heavy_start, g->heavy_off, g->heavy_tfirst};
    const int grid = fill_grid(c, kDoBlock, 2);
    k_build<kDoBlock, 8, kHvTile, true><<<grid, kDoBlock, 0, c->stream>>>(
        HeavyOp{g->rowptr}, g->rowptr, nullptr, g->m, nullptr, hb, cnt);
    APSPIE_LAUNCH_CHECK();
    APSPIE_HIP(hipMemcpyAsync(&g->heavy_cnt, cnt, 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
}

void sssp_do(Ctx *c, const Graph *g0, int32_t source, float *dist_out, int32_t pull_div) {
    Graph *g = const_cast<Graph *>(g0);
    if (pull_div > 0) graph_prepare_heavy(c, g);  // only pulled rounds walk the long-row list
    FrontierSet fs = frontier_alloc(c, "sssp", g->m, g->nnz, kDoTile, true);
    int32_t *dist = (int32_t *)dist_out;
    const int64_t words = g->m / 32 + 1;
    uint32_t *fb = (uint32_t *)scratch(c, "sssp_fb", sizeof(uint32_t) * 3 * words);
    fill_i32(c, dist, __builtin_bit_cast(int32_t, kInf), g->m);
    fill_i32(c, (int32_t *)fb, 0, 3 * words);
    fill_i32(c, (int32_t *)fs.ctl, 0, 2 * kCtlWords);
    k_do_seed<<<1, 256, 0, c->stream>>>(g->rowptr, source, dist, fb, fs.f[0], fs.ctl);
    APSPIE_LAUNCH_CHECK();
    const FrontierBufs heavy{g->heavy_v, g->heavy_start, g->heavy_off, g->heavy_tfirst};
    const int grid = fill_grid(
        c, kDoBlock, (int)std::max<int64_t>(1, opt(c, "bfs_blocks_per_cu", 8) * 256 / kDoBlock));
    // rounds between host checks: the option first, then 4, 8, 16, ... (the
    // idle tail of the last chunk is a launch per round)
    int chunk = (int)std::max<int64_t>(1, opt(c, "bfs_levels_per_sync", 8));
    int grow = 4;
    KernelTimer kt(c);
    uint64_t *h = (uint64_t *)c->pinned;
    int r = 0;
    for (;;) {
        kt.begin();
        for (int k = 0; k < chunk; ++k, ++r) {
            k_sssp_do_round<<<grid, kDoBlock, 0, c->stream>>>(
                g->rowptr, g->col, g->val, (int32_t)g->nnz, g->m, dist, fb + (r % 3) * words,
                fb + ((r + 1) % 3) * words, fb + ((r + 2) % 3) * words, fs.f[r & 1],
                fs.f[(r + 1) & 1], heavy, g->heavy_cnt, pull_div, fs.ctl, r);
            APSPIE_LAUNCH_CHECK();
        }
        kt.end(chunk);
        APSPIE_HIP(hipMemcpyAsync(h, fs.ctl, 16 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        sync(c);
        c->stats.host_syncs++;
        if (cnt_entries(h[r % 3]) == 0) break;
        APSPIE_CHECK(r <= 4 * g->m + 64, APSPIE_ERR_DEVICE, "SSSP did not converge");
        chunk = grow;
        grow = std::min(grow * 2, 256);
    }
    c->stats.edges_traversed = (int64_t)h[4];
    c->stats.vertices_reached = (int64_t)h[5];
    c->stats.iterations = (int32_t)h[6];
    c->stats.edges_checked = (int64_t)h[7];
    kt.collect();
}

// ---------------------------------------------------------------- near-far
// Near-far rounds (Davidson et al.'s two-pile form of delta-stepping): a
// relaxation that lowers dist[v] below the threshold T puts v in the next
// frontier as above; one that leaves it at or above T only marks v pending.
// When a round finds the frontier empty it is a split round instead: T grows
// by delta and a sweep over all vertices moves the pending ones now below T
// into the frontier.  A pending vertex whose distance is already below the
// old T was lowered again by a near relaxation, which queued it with that
// distance: the sweep drops it (so near relaxations never clear the mark,
// and a far relaxation that lost the race to a near one costs nothing).  Vertices are expanded fewer times (mostly once, near
// their final distance); the fixed point is the same.  Relaxations and
// claims are those of the Bellman-Ford rounds above (a non-returning atomic
// min, the next bitmap's OR as the claim); a split's sweep meets every vertex
// once and appends without a claim.
//
// Light / heavy mode (delta-stepping's split of the edges, Meyer and
// Sanders): each row holds its light entries (value below delta) first, in
// a copy of the CSR kept with the graph.  A vertex claimed for the frontier
// expands only its light entries and is marked; the next split adds the
// marked vertices' heavy entries to the frontier (and clears the marks)
// while it moves the pending vertices, each with its light and its heavy
// entries.  A vertex is re-expanded with every improvement in its bucket,
// but its heavy entries, whose relaxations land at or above T whenever the
// step is delta, only once per bucket, with its distance at the bucket's
// end (R-MAT-22, width 64: 273M -> 165M expanded entries per source).  The
// fixed point is the same: every vertex whose distance drops is marked or
// moved, and a split sweeps every mark before the rounds can end.
//
// ctl words (every word a round reads was written by the previous round, and
// no word a round reads is written in that round): [0..2] frontier counter
// ring, [4..7] stats, [10 + (r & 1)] the split step of round r (float bits),
// [12 + (r & 1)] T of round r (float bits), [14 + (r & 1)]
// 1 if round r - 1 was a split, and, in the spread slots, the pending-vertex
// count left by the split of round r ([16 + 16 i + r % 3]) and the smallest
// distance among them ([16 + 16 i + 3 + r % 3], float bits, ~0 if none).
//
// The threshold of a split is T + step, at least the next float above T (T +
// step rounds to T once T exceeds step * 2^24); the step starts at delta.  A
// split that follows a split that expanded nothing doubles the step and
// jumps at least to just above the smallest pending distance the earlier
// sweep left: with weights much larger than delta, T would otherwise climb in
// maxdist / delta empty splits.  Any threshold above T gives the same fixed
// point; the jump and the doubling only bound the rounds (every second
// consecutive split moves a vertex, and sparse distances widen the step).
constexpr int kNfBlock = 1024;
constexpr int kNfPer = 8;  // vertices per thread per sweep pass

struct SsspNearOp {
    static constexpr bool kStreamColumns = true;  // columns and weights stream past the L2
    const float *w;
    int32_t *dist;  // float bits
    uint32_t *fb_next;
    uint8_t *pend;
    int32_t tbits;  // threshold T (float bits; distances are non-negative)
    const int32_t *rend;  // end of the row part a claimed vertex expands next
    uint32_t *hb;         // light / heavy mode: bit per vertex whose heavy entries are due
    __device__ __forceinline__ int32_t row_end(int32_t v) const { return rend[v]; }
    __device__ __forceinline__ bool edge(int32_t u, int32_t e, int32_t v) const {
        const int32_t nd = dev::fbits(dev::bitsf(dist[u]) + __builtin_nontemporal_load(w + e));
        if (nd >= dist[v]) return false;
        __hip_atomic_fetch_min(dist + v, nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (nd >= tbits) {  // far: expanded by a later split round
            pend[v] = 1;
            return false;
        }
        if (!claim_bit(fb_next, v)) return false;
        // its light entries come next; the heavy ones once the frontier empties
        if (hb) __hip_atomic_fetch_or(hb + (v >> 5), 1u << (v & 31), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
};

// two 1024-thread blocks per CU resident: at most 64 VGPRs
template <int TILE>
__global__ __launch_bounds__(kNfBlock) __attribute__((amdgpu_waves_per_eu(8))) void
k_sssp_nf_round(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                const float *w, int32_t nnz, int32_t m, int32_t *dist, uint32_t *fb_next,
                uint32_t *fb_clear, uint8_t *pend, FrontierBufs cur,
                FrontierBufs nxt, uint64_t *ctl, int32_t r, const int32_t *rend,
                uint32_t *hb) {
    __shared__ ExpandLds<kNfBlock, TILE, true> lds;
    __shared__ uint64_t s_part[kNfBlock / 64];
    __shared__ uint32_t s_min[kNfBlock / 64];
    const uint64_t cc = ctl[r % 3];
    const int32_t F = (int32_t)cnt_entries(cc), E = (int32_t)cnt_edges(cc);
    const int32_t tbits = (int32_t)ctl[12 + (r & 1)];
    const bool split_last = ctl[14 + (r & 1)] != 0;
    const uint64_t left_before = nf_left(ctl, r + 2);  // left by the previous round's split
    const bool split = F == 0;
    const bool done = split && split_last && left_before == 0;
    const uint64_t min_before = nf_min_left(ctl, r + 2);
    const bool jump = split && !done && split_last && min_before != ~0ull;
    const float step = nf_next_step(dev::bitsf((int32_t)ctl[10 + (r & 1)]), jump);
    const int32_t tnext =
        split && !done ? nf_next_threshold(tbits, step, split_last, min_before) : tbits;
    __syncthreads();  // every thread has read the control words
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[(r + 2) % 3] = 0;
        ctl[10 + ((r + 1) & 1)] = (uint64_t)(uint32_t)dev::fbits(step);
        ctl[12 + ((r + 1) & 1)] = (uint64_t)(uint32_t)tnext;
        ctl[14 + ((r + 1) & 1)] = split ? 1 : 0;
        nf_reset(ctl, r + 1);
        if (F > 0) {
            ctl[4] += (uint64_t)E;
            ctl[5] += (uint64_t)F;
            ctl[6] += 1;
            ctl[7] += (uint64_t)E;  // CSR entries read
        } else if (!done) {
            ctl[6] += 1;  // split rounds count as rounds
        }
    }
    if (done) return;
    // the bitmap round r + 1 claims into (round r claims into another)
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m / 32 + 1;
         i += (int64_t)gridDim.x * blockDim.x)
        fb_clear[i] = 0;
    if (F > 0) {
        SsspNearOp op{w, dist, fb_next, pend, tbits, rend, hb};
        expand_tiles<kNfBlock, TILE, true, true, true>(lds, op, rowptr, col, nnz, cur, F, E, nxt,
                                                       ctl + ((r + 1) % 3));
        return;
    }
    // split: move the pending vertices below the new threshold into the
    // frontier, and in light / heavy mode add the heavy entries of the
    // vertices marked since the last split (a moved vertex brings its light
    // and its heavy entries).  The 32 vertices of a mark word sit in 32
    // consecutive lanes, and the word's first lane clears it after the wave
    // has read it (nothing marks during a split).
    uint64_t left = 0;
    uint32_t min_left = ~0u;  // smallest distance left pending (bits)
    for (int64_t pbase = (int64_t)blockIdx.x * kNfBlock * kNfPer; pbase < m;
         pbase += (int64_t)gridDim.x * kNfBlock * kNfPer) {
        int32_t cv[kNfPer], crs[kNfPer], cd[kNfPer], hrs[kNfPer], hd[kNfPer];
        bool take[kNfPer], drop[kNfPer], marked[kNfPer];
#pragma unroll
        for (int k = 0; k < kNfPer; ++k) {
            const int64_t v = pbase + (int64_t)k * kNfBlock + threadIdx.x;
            cv[k] = (int32_t)v;
            crs[k] = cd[k] = hrs[k] = hd[k] = 0;
            take[k] = drop[k] = marked[k] = false;
            if (v < m && pend[v]) {
                const int32_t dv = dist[v];
                if (dv < tbits) drop[k] = true;  // lowered below T since: already expanded
                else if (dv < tnext) take[k] = true;
                else {
                    ++left;
                    min_left = min(min_left, (uint32_t)dv);
                }
            }
            if (hb && v < m) {
                const uint32_t word = hb[v >> 5];
                marked[k] = (word >> (v & 31)) & 1u;
                if ((v & 31) == 0 && word) hb[v >> 5] = 0;
            }
        }
#pragma unroll
        for (int k = 0; k < kNfPer; ++k) {
            if (take[k] || drop[k]) pend[cv[k]] = 0;
            if (take[k]) {  // the sweep meets each vertex once: no claim needed
                crs[k] = rowptr[cv[k]];
                cd[k] = rend[cv[k]] - crs[k];
            }
            if (take[k] || marked[k]) {
                hrs[k] = rend[cv[k]];
                hd[k] = hb ? rowptr[cv[k] + 1] - hrs[k] : 0;
            }
        }
        produce<kNfBlock, kNfPer, TILE, true>(cv, crs, cd, nxt, ctl + ((r + 1) % 3), lds.scan,
                                              &lds.base);
        if (hb)
            produce<kNfBlock, kNfPer, TILE, true>(cv, hrs, hd, nxt, ctl + ((r + 1) % 3),
                                                  lds.scan, &lds.base);
    }
    const uint64_t ws = dev::wave_sum_u64(left);
    const uint32_t wm = dev::wave_min_u32(min_left);
    if (dev::lane_id() == 0) {
        s_part[threadIdx.x / 64] = ws;
        s_min[threadIdx.x / 64] = wm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        uint32_t mn = ~0u;
        for (int q = 0; q < kNfBlock / 64; ++q) {
            t += s_part[q];
            mn = min(mn, s_min[q]);
        }
        nf_report(ctl, r, t, mn);
    }
}

// Rows split by value for the light / heavy mode: one wave per row counts
// its light entries (value below delta), then writes them first and the
// heavy ones after them, each part in CSR order (ballot ranks), and records
// where the light part ends.
constexpr int kLfBlock = 256;
__global__ __launch_bounds__(kLfBlock) void k_light_first(const int32_t *__restrict__ rowptr,
                                                          const int32_t *__restrict__ col,
                                                          const float *__restrict__ val,
                                                          int32_t m, float delta, int32_t *lcol,
                                                          float *lval, int32_t *lend) {
    const int lane = dev::lane_id();
    const uint64_t below = (1ull << lane) - 1;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t v = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; v < m;
         v += nwaves) {
        const int32_t rs = rowptr[v], re = rowptr[v + 1];
        uint64_t nl = 0;
        for (int32_t e = rs + lane; e < re; e += 64) nl += val[e] < delta ? 1 : 0;
        nl = dev::wave_sum_u64(nl);
        int32_t lpos = rs, hpos = rs + (int32_t)nl;
        for (int32_t e0 = rs; e0 < re; e0 += 64) {
            const int32_t e = e0 + lane;
            const bool in = e < re;
            const float x = in ? val[e] : 0.0f;
            const bool light = in && x < delta;
            const uint64_t bl = __ballot(light), bh = __ballot(in && !light);
            if (in) {
                const int32_t p = light ? lpos + __popcll(bl & below) : hpos + __popcll(bh & below);
                lcol[p] = col[e];
                lval[p] = x;
            }
            lpos += __popcll(bl);
            hpos += __popcll(bh);
        }
        if (lane == 0) lend[v] = rs + (int32_t)nl;
    }
}

// The light-first copy of a graph's rows for width delta, built on first
// use and kept with the graph until it is destroyed: none is freed while the
// graph lives, since a call on another context may still be launching
// rounds over the copy it took (of this width or another).
Graph::LightFirst graph_light_first(Ctx *c, Graph *g, float delta) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    for (const Graph::LightFirst &x : g->lf)
        if (x.delta == delta) return x;
    const int64_t E = std::max<int64_t>(g->nnz, 1);
    const size_t ecol = ((size_t)E * 4 + 255) / 256 * 256;
    const size_t bytes = 2 * ecol + sizeof(int32_t) * ((size_t)g->m + 1);
    void *mem = nullptr;
    // owned by the graph (freed with it), like its other derived arrays
    APSPIE_HIP(hipMalloc(&mem, bytes));
    int32_t *lcol = (int32_t *)mem;
    float *lval = (float *)((char *)mem + ecol);
    int32_t *lend = (int32_t *)((char *)mem + 2 * ecol);
    try {
        if (g->m > 0) {
            // a wave per row and every row its own wave: the two passes over
            // a row are a few dependent loads, so the grid covers the rows
            const int grid = (int)std::min<int64_t>(ceil_div(g->m, kLfBlock / 64), 1 << 24);
            k_light_first<<<grid, kLfBlock, 0, c->stream>>>(g->rowptr, g->col, g->val, g->m,
                                                            delta, lcol, lval, lend);
            APSPIE_LAUNCH_CHECK();
        }
        sync(c);
    } catch (...) {
        (void)hipFree(mem);
        throw;
    }
    g->lf.push_back(Graph::LightFirst{delta, mem, lcol, lend, lval});
    return g->lf.back();
}

template <int TILE>
void sssp_near_far(Ctx *c, const Graph *g, int32_t source, float *dist_out, float delta) {
    FrontierSet fs = frontier_alloc(c, "sssp", g->m, g->nnz, TILE, true);
    int32_t *dist = (int32_t *)dist_out;
    const int64_t words = g->m / 32 + 1;
    uint32_t *fb = (uint32_t *)scratch(c, "sssp_fb", sizeof(uint32_t) * 3 * words);
    uint8_t *pend = (uint8_t *)scratch(c, "sssp_pend", (size_t)g->m + 4);
    // light / heavy mode (option sssp_light_heavy, on by default) when the
    // weights fall on both sides of delta: a claimed vertex expands its light
    // entries (value below delta), and its heavy ones once per bucket, at
    // the next split, with its distance by then -- a heavy relaxation lands
    // at or above the threshold whenever the step is delta, so re-expanding
    // it with every improvement only repeats far marks
    float lo = 0.0f, hi = 0.0f;
    const bool lh = opt(c, "sssp_light_heavy", 1) != 0 &&
                    graph_weight_range(c, const_cast<Graph *>(g), &lo, &hi) && lo < delta &&
                    hi >= delta;
    const int32_t *col = g->col, *rend = g->rowptr + 1;
    const float *val = g->val;
    uint32_t *hb = nullptr;
    if (lh) {
        const Graph::LightFirst lf = graph_light_first(c, const_cast<Graph *>(g), delta);
        col = lf.col;
        val = lf.val;
        rend = lf.end;
        hb = (uint32_t *)scratch(c, "sssp_hb", sizeof(uint32_t) * words);
        fill_i32(c, (int32_t *)hb, 0, words);
    }
    fill_i32(c, dist, __builtin_bit_cast(int32_t, kInf), g->m);
    fill_i32(c, (int32_t *)fb, 0, 3 * words);
    fill_i32(c, (int32_t *)pend, 0, ceil_div(g->m + 4, 4));
    fill_i32(c, (int32_t *)fs.ctl, 0, 2 * kCtlWords);
    k_sssp_seed<TILE><<<1, 256, 0, c->stream>>>(g->rowptr, source, dist, fs.f[0], fs.ctl);
    fill_i32(c, (int32_t *)(fs.ctl + 12), __builtin_bit_cast(int32_t, delta), 1);  // T of round 0
    fill_i32(c, (int32_t *)(fs.ctl + 10), __builtin_bit_cast(int32_t, delta), 1);  // its step
    APSPIE_LAUNCH_CHECK();
    const int grid = fill_grid(
        c, kNfBlock, (int)std::max<int64_t>(1, opt(c, "bfs_blocks_per_cu", 8) * 256 / kNfBlock));
    // rounds between host checks: the option first, then 4, 8, 16, ... (the
    // idle tail of the last chunk is a launch per round)
    int chunk = (int)std::max<int64_t>(1, opt(c, "bfs_levels_per_sync", 8));
    int grow = 4;
    KernelTimer kt(c);
    uint64_t *h = (uint64_t *)c->pinned;
    int r = 0;
    for (;;) {
        kt.begin();
        for (int k = 0; k < chunk; ++k, ++r) {
            k_sssp_nf_round<TILE><<<grid, kNfBlock, 0, c->stream>>>(
                g->rowptr, col, val, (int32_t)g->nnz, g->m, dist,
                fb + ((r + 1) % 3) * words, fb + ((r + 2) % 3) * words, pend, fs.f[r & 1],
                fs.f[(r + 1) & 1], fs.ctl, r, rend, hb);
            APSPIE_LAUNCH_CHECK();
        }
        kt.end(chunk);
        APSPIE_HIP(hipMemcpyAsync(h, fs.ctl, kCtlWords * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        sync(c);
        c->stats.host_syncs++;
        // finished when the last round was a split that left nothing pending
        // and found nothing below the new threshold
        if (cnt_entries(h[r % 3]) == 0 && h[14 + (r & 1)] != 0 && nf_left(h, r + 2) == 0) break;
        APSPIE_CHECK(r <= 8 * g->m + 1024, APSPIE_ERR_DEVICE, "SSSP did not converge");
        chunk = grow;
        grow = std::min(grow * 2, 256);
    }
    c->stats.edges_traversed = (int64_t)h[4];
    c->stats.vertices_reached = (int64_t)h[5];
    c->stats.iterations = (int32_t)h[6];
    c->stats.edges_checked = (int64_t)h[7];
    kt.collect();
}

}  // namespace

// The near-far split width a graph's weights suggest: ten times the largest
// weight over the mean row length, rounded to the nearest power of two of at
// least 16 -- 64 on R-MAT-22 at weights 1..255 (mean row 30; 48-64 measured
// best, 8 % under Bellman-Ford) and 512 on 2D lattices (mean row 3.6; 1.35x
// and 2.5x under Bellman-Ford at 2048^2 and 4096^2).  No upper cap: the width
// follows the weights' scale (2^32 at weights up to 1e10 on R-MAT), so the
// number of splits does not grow with the weights.  0 when the weights are
// not positive or lie within a factor of two of each other (a vertex then
// settles in the round that first reaches it).
double sssp_weight_delta(Ctx *c, Graph *g) {
    float lo = 0.0f, hi = 0.0f;
    if (g->m == 0 || !graph_weight_range(c, g, &lo, &hi) || !(lo > 0.0f) || !(hi >= 2.0f * lo))
        return 0;
    const double want = 10.0 * hi / std::max(1.0, (double)g->nnz / g->m);
    double d = 16;
    while (d < 0x1p100 && d * 1.4142135623730951 < want) d *= 2;
    return d;
}

// The width used when sssp_delta is left at -1: the width the weights
// suggest from 2^22 vertices on, and from 2^20 on where that width splits
// the weights into light and heavy entries (light / heavy rounds on);
// Bellman-Ford rounds (0) otherwise, where the rounds' fixed cost outweighs
// the re-expansions near-far saves.  Measured at weights 1..255, ms per
// source, Bellman-Ford / width 64 with light / heavy rounds: R-MAT-18
// 0.58 / 0.94, R-MAT-20 1.42 / 1.33, R-MAT-21 2.21 / 1.91, R-MAT-22 4.03 /
// 2.90; the lattices' width (512) is above every weight, and 512^2-1024^2
// run faster without splits.
double sssp_auto_delta(Ctx *c, Graph *g) {
    if (g->m >= (1 << 22)) return sssp_weight_delta(c, g);
    if (g->m < (1 << 20) || opt(c, "sssp_light_heavy", 1) == 0) return 0;
    const double d = sssp_weight_delta(c, g);
    float lo = 0.0f, hi = 0.0f;
    return d > 0 && graph_weight_range(c, g, &lo, &hi) && lo < d && hi >= d ? d : 0;
}

// The batched SSSP width when batch_sssp_delta is left at -1: the weights'
// width where it is at least 256 (sparse rows, long paths: 512 on the
// lattices, marginally ahead of 256 at 128 and 512 sources), Bellman-Ford
// rounds otherwise (R-MAT, width 64: with many sources in flight the splits
// cost more than the re-expansions they save -- 32 / 128 sources of R-MAT-22
// measured 2.64 / 2.09 ms per query without splits, 2.68 / 2.24 at width
// 256, 4.57 / 3.05 at 64).
double sssp_batch_auto_delta(Ctx *c, Graph *g) {
    const double d = sssp_weight_delta(c, g);
    return d >= 256 ? d : 0;
}

void sssp_run(Ctx *c, const Graph *g, int32_t source, float *dist) {
    double delta = (double)opt(c, "sssp_delta", -1);
    if (delta < 0) delta = sssp_auto_delta(c, const_cast<Graph *>(g));
    if (delta > 0) {
        switch (opt(c, "bfs_tile", 2048)) {
        case 1024: sssp_near_far<1024>(c, g, source, dist, (float)delta); return;
        case 4096: sssp_near_far<4096>(c, g, source, dist, (float)delta); return;
        default: sssp_near_far<2048>(c, g, source, dist, (float)delta); return;
        }
    }
    // pulled rounds read w(v, u) for the edge u -> v: symmetric values only
    int32_t pull_div = (int32_t)opt(c, "sssp_pull_div", 0);
    if (pull_div > 0 && !graph_is_symmetric_weighted(c, const_cast<Graph *>(g))) pull_div = 0;
    sssp_do(c, g, source, dist, pull_div);
}

}  // namespace apspie
