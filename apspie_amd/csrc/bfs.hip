// Breadth-first search as repeated Boolean SpMSpV with a visited mask.
//
// Semantics follow bfs() (bfs.cuh:32-125) / spmspvBfs() (spmspvBfs.cuh:96-243):
// depth[src] = 0, depth[v] = level of first discovery, -1 if unreached.  The
// reference runs expand -> sort -> reduce -> scatter -> addResult per level
// with host round trips; here one kernel per level does the whole SpMSpV.
//
// Two directions of the same product y = A^T x (Boolean semiring, masked by
// the visited set):
//   push (SpMSpV, "top-down"): load-balanced expansion of the frontier's
//     rows; a fetch-or on the visited bitmap decides the first writer of each
//     bit, which writes the depth and appends the vertex to the next frontier.
//   pull (masked SpMV, "bottom-up", the dense-vector direction the reference
//     has as SpmvCsrBinary, bfs.cuh:11-14): every unvisited vertex scans its
//     row until it finds a frontier vertex.  A wave owns 64 consecutive
//     vertices, i.e. two bitmap words, so the visited and next-frontier
//     bitmaps are written without atomics.
// Every level leaves the next frontier in both forms (compacted queue with
// degree offsets for push, bitmap for pull), and both kernels of a level read
// the same device counters to pick the direction (Beamer's rule: pull when
// the frontier's edges exceed the unexplored edges / alpha while growing,
// push again when the frontier falls under m / beta while shrinking), so no
// host round trip is needed per level.  The result does not depend on the
// direction: a vertex's depth is its level of first discovery either way.
#include "common.hpp"
#include "device.hpp"
#include "frontier.hpp"
#include "timing.hpp"

namespace apspie {

namespace {

// 1024-thread blocks, two per CU: a level's appends to the next frontier
// are one 64-bit atomic per block and pass on a single counter, which the L2
// serialises (~10 ns each), so fewer, larger blocks shorten every level.
constexpr int kBlock = 1024;
constexpr uint32_t kPush = 0, kPull = 1;
constexpr int32_t kPullLight = 32;  // rows up to this length are scanned by one lane
constexpr int kPullPer = 8;         // vertices per thread per pull pass
constexpr int kPullIlp = 8;         // row entries a lane loads at once
constexpr int32_t kSmallPush = 1 << 18;  // push levels up to this many edges prefetch rows

// ctl layout (uint64 words): [0..2] frontier counters (ring), [4] edges
// traversed, [5] entries expanded, [6] levels, [16 + 16 i] edges scanned by
// pull (spread over 16 accumulators),
// [8..9] (mode << 32 | entries) per level parity, [10..11] explored edges
// before the level, per level parity, [12] row entries read by push levels.
struct Ctl {
    uint64_t *w;
    __device__ uint64_t *cnt(int l) const { return w + (l % 3); }
    __device__ uint64_t *slot(int l) const { return w + 8 + (l & 1); }
    __device__ uint64_t *expl(int l) const { return w + 10 + (l & 1); }
};

struct DoParams {
    int32_t direction;  // 0 push only, 1 direction-optimising, 2 pull whenever possible
    int32_t alpha, beta;
    int64_t nnz;
    int32_t m;
};

// Direction of level l, computed identically by every block from values
// written by the previous level: prev = its (mode << 32 | entries), explored =
// frontier edges of all levels before l.
__device__ __forceinline__ uint32_t decide(int l, const DoParams &p, uint64_t prev,
                                           uint64_t explored, int32_t F, int32_t E) {
    if (p.direction == 0) return kPush;
    if (p.direction == 2) return l == 0 ? kPush : kPull;
    const uint32_t pmode = uint32_t(prev >> 32), pF = uint32_t(prev);
    const int64_t unexplored = p.nnz - (int64_t)explored - E;
    if (pmode == kPush)
        return ((int64_t)E * p.alpha > unexplored && (uint32_t)F > pF) ? kPull : kPush;
    return ((int64_t)F * p.beta < p.m && (uint32_t)F < pF) ? kPush : kPull;
}

// bookkeeping done once per level by block 0 of the active kernel
__device__ __forceinline__ void level_books(const Ctl &c, int l, uint32_t mode, int32_t F,
                                            int32_t E, uint64_t explored) {
    *c.cnt(l + 2) = 0;
    *c.slot(l) = (uint64_t(mode) << 32) | uint32_t(F);
    *c.expl(l) = explored + (uint64_t)E;
    c.w[4] += (uint64_t)E;
    c.w[5] += (uint64_t)F;
    c.w[6] += 1;
}

__device__ __forceinline__ void clear_slice(uint32_t *p, int64_t words) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

struct BfsOp {
    uint32_t *vis;
    int32_t *depth;
    uint32_t *fb_next;  // next-frontier bitmap (null when only pushing)
    int32_t nd;
    __device__ __forceinline__ bool edge(int32_t, int32_t, int32_t v) const {
        uint32_t *w = vis + (v >> 5);
        const uint32_t bit = 1u << (v & 31);
        if (*w & bit) return false;  // seen (possibly stale: the atomic decides)
        const uint32_t old = dev::atomic_or_dev(w, bit);
        if (old & bit) return false;
        depth[v] = nd;
        if (fb_next)
            __hip_atomic_fetch_or(fb_next + (v >> 5), bit, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
};

// Everything a traversal needs before level 0, in one launch:
//   depth[v] = -1 (0 at the source);
//   visited bitmap = {source}, plus, when pulling is possible (symmetric CSR),
//     every vertex without neighbours: it can never be reached, and pull
//     levels then skip it after one bitmap read;
//   frontier bitmaps fb0 = {source}, fb1 = {} (fb2 is cleared by level 0);
//   the counter block and the level-0 frontier (the source's row).
// Depths -1 (the source 0) with 16-byte stores where the output is aligned,
// then the bitmap words: visited = the graph's isolated-vertex bits (given
// for a symmetric CSR, where such a vertex is never reached) | the source,
// the level-0 frontier bitmap = the source, the next one cleared.
template <int TILE>
__global__ void k_bfs_init(const int32_t *__restrict__ rowptr, const uint32_t *__restrict__ iso,
                           int32_t m, int32_t src, int32_t *depth, uint32_t *vis, uint32_t *fb0,
                           uint32_t *fb1, FrontierBufs f0, uint64_t *ctl) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t done = 0;
    if (((uintptr_t)depth & 15) == 0) {
        int4 *d4 = reinterpret_cast<int4 *>(depth);
        const int64_t n4 = m / 4;
        for (int64_t q = tid; q < n4; q += stride) {
            int4 x = make_int4(-1, -1, -1, -1);
            if (q == src / 4) {
                const int r = src % 4;
                x.x = r == 0 ? 0 : -1;
                x.y = r == 1 ? 0 : -1;
                x.z = r == 2 ? 0 : -1;
                x.w = r == 3 ? 0 : -1;
            }
            d4[q] = x;
        }
        done = n4 * 4;
    }
    for (int64_t v = done + tid; v < m; v += stride) depth[v] = v == src ? 0 : -1;
    const int64_t words = ((int64_t)m + 31) / 32;
    for (int64_t wi = tid; wi < words; wi += stride) {
        const uint32_t sbit = (src >> 5) == wi ? 1u << (src & 31) : 0u;
        vis[wi] = (iso ? iso[wi] : 0u) | sbit;
        if (fb0) {
            fb0[wi] = sbit;
            fb1[wi] = 0;
        }
    }
    if (blockIdx.x == 0) {
        const int32_t rs = rowptr[src];
        const int32_t d = rowptr[src + 1] - rs;
        for (int i = threadIdx.x; i < kCtlWords; i += blockDim.x) ctl[i] = 0;
        __syncthreads();
        if (threadIdx.x == 0) {
            ctl[0] = d > 0 ? ((uint64_t(1) << 32) | uint64_t(d)) : 0;
            ctl[9] = uint64_t(kPush) << 32;  // "level -1": pushing, empty
            f0.start[0] = rs;
            f0.off[0] = 0;
            if (f0.v) f0.v[0] = src;
        }
        for (int32_t t = threadIdx.x; (int64_t)t * TILE < d; t += blockDim.x) f0.tfirst[t] = 0;
    }
}

// One level of the direction-optimising loop: every block reads the same
// counters, picks the same direction and runs either the push expansion or
// the pull sweep.
// The level is latency-bound (a chain of dependent loads per wave), so the
// whole default grid (two blocks per CU) must be resident at once: 8 waves
// per SIMD, i.e. at most 64 VGPRs.
template <int TILE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_bfs_level(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int32_t *depth,
    uint32_t *vis, const uint32_t *__restrict__ fb_cur, uint32_t *fb_next, uint32_t *fb_clear,
    FrontierBufs cur, FrontierBufs nxt, uint64_t *ctlw, int32_t level, DoParams p) {
    __shared__ uint64_t s_scan[kBlock / 64 + 1];
    __shared__ uint64_t s_base;
    __shared__ uint64_t s_scanned[kBlock / 64];
    const Ctl c{ctlw};
    // the three control words are loaded together (one round trip)
    const uint64_t cc = *c.cnt(level);
    const uint64_t prev = *c.slot(level + 1), explored = *c.expl(level + 1);
    const int32_t F = (int32_t)cnt_entries(cc), E = (int32_t)cnt_edges(cc);
    if (F == 0) {
        // keep the counter ring clean after the last level: the entry two
        // slots ahead still holds an old frontier and would revive it
        if (blockIdx.x == 0 && threadIdx.x == 0) *c.cnt(level + 2) = 0;
        return;
    }
    const uint32_t mode = decide(level, p, prev, explored, F, E);
    __syncthreads();  // every thread has read the shared state before block 0 writes
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        level_books(c, level, mode, F, E, explored);
        if (mode == kPush) c.w[12] += (uint64_t)E;  // entries read by push levels
    }
    clear_slice(fb_clear, p.m / 32 + 1);
    if (mode == kPush) {
        // small levels are latency-bound: row extents loaded with the columns
        const bool small = E <= kSmallPush;
        BfsOp op{vis, depth, fb_next, level + 1};
        expand_tiles<kBlock, TILE, false, false, true>(op, rowptr, col, (int32_t)p.nnz, cur, F,
                                                       E, nxt, c.cnt(level + 1), small);
        return;
    }
    // pull: each thread takes kPullPer vertices of a pass; for every k the
    // block's 256 threads cover 256 consecutive vertices, so a wave owns 64
    // consecutive vertices (two bitmap words) and writes them without atomics.
    // All loads of a pass are issued before any of them is needed, and the
    // claims of the pass are appended with one block-wide reservation.
    const int lane = dev::lane_id();
    uint64_t scanned = 0;
    for (int64_t pbase = (int64_t)blockIdx.x * kBlock * kPullPer; pbase < p.m;
         pbase += (int64_t)gridDim.x * kBlock * kPullPer) {
        int32_t cv[kPullPer], crs[kPullPer], cd[kPullPer];
        int32_t rs[kPullPer], re[kPullPer];
        bool open[kPullPer], hit[kPullPer];
#pragma unroll
        for (int k = 0; k < kPullPer; ++k) {
            const int64_t v = pbase + (int64_t)k * kBlock + threadIdx.x;
            cv[k] = (int32_t)v;
            crs[k] = 0;
            cd[k] = 0;
            hit[k] = false;
            // row extents are loaded with the visited bits, not after them
            // (coalesced, and one dependent round trip less)
            const bool in = v < p.m;
            open[k] = in && !((vis[v >> 5] >> (v & 31)) & 1u);
            rs[k] = in ? rowptr[v] : 0;
            re[k] = in ? rowptr[v + 1] : 0;
        }
        // first entry of every open row (the lowest-id neighbour: on R-MAT
        // usually a hub, and most hits happen here)
        int32_t u0[kPullPer];
#pragma unroll
        for (int k = 0; k < kPullPer; ++k) u0[k] = open[k] && re[k] > rs[k] ? col[rs[k]] : -1;
#pragma unroll
        for (int k = 0; k < kPullPer; ++k) {
            if (u0[k] >= 0 && ((fb_cur[u0[k] >> 5] >> (u0[k] & 31)) & 1u)) hit[k] = true;
            if (open[k]) scanned += (uint64_t)min(re[k] - rs[k], 1);
        }
        // the next kPullLight - 1 entries of every row still open: each lane
        // walks its own rows one after another, eight loads in flight, so a
        // wave costs the longest lane's total rather than the sum over k of
        // the longest row of each k
        uint32_t pend = 0, hitm = 0, longm = 0;
#pragma unroll
        for (int k = 0; k < kPullPer; ++k)
            if (open[k] && !hit[k] && re[k] - rs[k] > 1) pend |= 1u << k;
        {
            int32_t e = 0, lim = 0, rend = 0;
            uint32_t kbit = 0;
            for (;;) {
                if (e >= lim) {
                    if (!pend) break;
                    const int kk = __ffs(pend) - 1;
                    pend &= pend - 1;
                    kbit = 1u << kk;
                    int32_t r0 = 0;
#pragma unroll
                    for (int k = 0; k < kPullPer; ++k)
                        if (k == kk) {
                            r0 = rs[k];
                            rend = re[k];
                        }
                    e = r0 + 1;
                    lim = min(rend, r0 + kPullLight);
                }
                int32_t u[kPullIlp];
#pragma unroll
                for (int q = 0; q < kPullIlp; ++q) u[q] = e + q < lim ? col[e + q] : -1;
                uint32_t f[kPullIlp];
#pragma unroll
                for (int q = 0; q < kPullIlp; ++q)
                    f[q] = u[q] >= 0 ? (fb_cur[u[q] >> 5] >> (u[q] & 31)) & 1u : 0u;
                int at = kPullIlp;
#pragma unroll
                for (int q = kPullIlp - 1; q >= 0; --q)
                    if (f[q]) at = q;
                if (at < kPullIlp) {
                    hitm |= kbit;
                    scanned += (uint64_t)(at + 1);
                    e = lim;
                } else {
                    scanned += (uint64_t)min(kPullIlp, lim - e);
                    e += kPullIlp;
                    if (e >= lim && lim < rend) longm |= kbit;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kPullPer; ++k) hit[k] = hit[k] || ((hitm >> k) & 1u);
        // rows still open after kPullLight entries: the whole wave scans one
        // row 64 entries at a time, coalesced, stopping at the first hit
#pragma unroll
        for (int k = 0; k < kPullPer; ++k) {
            const int32_t lim = min(re[k], rs[k] + kPullLight);
            uint64_t heavy = __ballot((longm >> k) & 1u);
            while (heavy) {
                const int h = __ffsll((unsigned long long)heavy) - 1;
                heavy &= heavy - 1;
                const int32_t hrs = __shfl(lim, h, 64), hre = __shfl(re[k], h, 64);
                bool found = false;
                for (int32_t e0 = hrs; e0 < hre; e0 += 64) {
                    const int32_t e = e0 + lane;
                    bool f = false;
                    if (e < hre) {
                        const int32_t u = col[e];
                        f = (fb_cur[u >> 5] >> (u & 31)) & 1u;
                    }
                    const uint64_t fm = __ballot(f);
                    if (fm) {
                        found = true;
                        if (lane == 0)
                            scanned += (uint64_t)(e0 - hrs + __ffsll((unsigned long long)fm));
                        break;
                    }
                }
                if (!found && lane == 0) scanned += (uint64_t)(hre - hrs);
                if (lane == h) hit[k] = found;
            }
        }
#pragma unroll
        for (int k = 0; k < kPullPer; ++k) {
            if (hit[k]) {
                depth[cv[k]] = level + 1;
                crs[k] = rs[k];
                cd[k] = re[k] - rs[k];
            }
            const uint64_t bm = __ballot(hit[k]);
            const int64_t wbase = (pbase + (int64_t)k * kBlock + (threadIdx.x & ~63)) >> 5;
            if (lane == 0 && (uint32_t)bm) {
                fb_next[wbase] = (uint32_t)bm;
                vis[wbase] |= (uint32_t)bm;
            }
            if (lane == 32 && (uint32_t)(bm >> 32)) {
                fb_next[wbase + 1] = (uint32_t)(bm >> 32);
                vis[wbase + 1] |= (uint32_t)(bm >> 32);
            }
        }
        produce<kBlock, kPullPer, TILE, false>(cv, crs, cd, nxt, c.cnt(level + 1), s_scan,
                                               &s_base);
    }
    const uint64_t ws = dev::wave_sum_u64(scanned);
    if (lane == 0) s_scanned[threadIdx.x / 64] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int q = 0; q < kBlock / 64; ++q) t += s_scanned[q];
        if (t) dev::atomic_add_dev(spread_slot(ctlw), t);
    }
}

// The device memory of one BFS call: the frontier pair and its control
// block, the visited bitmap and, when pulling, the three frontier bitmaps and
// the isolated-vertex bits.  bfs_run takes it from the context's scratch pool;
// bfs_run_in carves it from a block the caller provides.
struct BfsBufs {
    FrontierSet fs;
    uint32_t *vis = nullptr;
    uint32_t *fb = nullptr;         // null: every level pushes
    const uint32_t *iso = nullptr;  // with fb
    bool kernel_timer = true;       // per-chunk HIP events (created per call)
};

template <int TILE>
void bfs_levels(Ctx *c, const Graph *g, int32_t source, int32_t *depth, const BfsBufs &b) {
    const int direction = b.fb ? (int)opt(c, "bfs_direction", 1) : 0;
    const bool do_bfs = b.fb != nullptr;
    const FrontierSet &fs = b.fs;
    const int64_t words = g->m / 32 + 1;
    uint32_t *vis = b.vis, *fb = b.fb;
    const uint32_t *iso = b.iso;
    k_bfs_init<TILE><<<(int)std::min<int64_t>(ceil_div(g->m, 4 * 256), 4 * c->cus), 256, 0,
                       c->stream>>>(g->rowptr, iso, g->m, source, depth, vis, fb,
                                    fb ? fb + words : nullptr, fs.f[0], fs.ctl);
    APSPIE_LAUNCH_CHECK();

    // bfs_blocks_per_cu counts 256-thread groups per CU
    const int grid = fill_grid(
        c, kBlock, (int)std::max<int64_t>(1, opt(c, "bfs_blocks_per_cu", 8) * 256 / kBlock));
    const DoParams dp{direction, (int32_t)opt(c, "bfs_alpha", 14), (int32_t)opt(c, "bfs_beta", 24),
                      g->nnz, g->m};
    int chunk = (int)std::max<int64_t>(1, opt(c, "bfs_levels_per_sync", 8));
    KernelTimer kt(c, b.kernel_timer);
    uint64_t *h = (uint64_t *)c->pinned;  // copy of the control block
    int level = 0;
    for (;;) {
        kt.begin();
        for (int k = 0; k < chunk; ++k, ++level) {
            const int cur = level & 1;
            if (!do_bfs) {
                k_expand<kBlock, TILE, false, false, true><<<grid, kBlock, 0, c->stream>>>(
                    BfsOp{vis, depth, nullptr, level + 1}, g->rowptr, g->col, (int32_t)g->nnz,
                    fs.f[cur], fs.ctl + (level % 3), fs.f[cur ^ 1], fs.ctl + ((level + 1) % 3),
                    fs.ctl + ((level + 2) % 3), fs.ctl + 4);
            } else {
                uint32_t *fcur = fb + (level % 3) * words, *fnext = fb + ((level + 1) % 3) * words,
                         *fclr = fb + ((level + 2) % 3) * words;
                k_bfs_level<TILE><<<grid, kBlock, 0, c->stream>>>(
                    g->rowptr, g->col, depth, vis, fcur, fnext, fclr, fs.f[cur], fs.f[cur ^ 1],
                    fs.ctl, level, dp);
            }
            APSPIE_LAUNCH_CHECK();
        }
        kt.end(chunk);
        // the whole control block in one copy: counters and statistics
        APSPIE_HIP(hipMemcpyAsync(h, fs.ctl, kCtlWords * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, c->stream));
        // wait on the context's end-of-call event: when this chunk was the
        // last, the call's timing ends here without a second round trip
        APSPIE_HIP(hipEventRecord(c->ev1, c->stream));
        APSPIE_HIP(hipEventSynchronize(c->ev1));
        c->stats.host_syncs++;
        if (cnt_entries(h[level % 3]) == 0) {
            c->ev1_recorded = true;
            break;
        }
        APSPIE_CHECK(level <= g->m + 1, APSPIE_ERR_DEVICE, "BFS did not converge");
        chunk = std::min(chunk * 2, 256);
    }
    c->stats.edges_traversed = (int64_t)h[4];
    c->stats.vertices_reached = (int64_t)h[5];
    c->stats.iterations = (int32_t)h[6];
    c->stats.edges_checked = do_bfs ? (int64_t)(spread_sum(h) + h[12]) : (int64_t)h[4];
    kt.collect();  // every launch, the empty trailing ones included
}

template <int TILE>
void bfs_levels_pool(Ctx *c, const Graph *g, int32_t source, int32_t *depth) {
    // pulling reads a row as the in-edges of its vertex: only valid if A == A^T
    const bool pull = opt(c, "bfs_direction", 1) != 0 &&
                      graph_is_symmetric(c, const_cast<Graph *>(g));
    const int64_t words = g->m / 32 + 1;
    BfsBufs b;
    b.fs = frontier_alloc(c, "bfs", g->m, g->nnz, TILE, false);
    b.vis = (uint32_t *)scratch(c, "bfs_bitmap", sizeof(uint32_t) * words);
    if (pull) {
        b.fb = (uint32_t *)scratch(c, "bfs_fb", sizeof(uint32_t) * 3 * words);
        // symmetric: a vertex without neighbours is never reached, so it
        // starts out visited; fb[2 * words ..] is cleared by level 0 itself
        b.iso = graph_iso_bits(c, const_cast<Graph *>(g));
    }
    bfs_levels<TILE>(c, g, source, depth, b);
}

// Pieces of a caller-provided block, each 256-byte aligned.
struct Carver {
    char *base;
    int64_t used = 0;
    void *take(int64_t bytes) {
        void *p = base ? base + used : nullptr;
        used += (bytes + 255) / 256 * 256;
        return p;
    }
};

// Lays the buffers of one call out in `block` (null: only counts the bytes).
// The tile partition is sized for the smallest tile (1024), so the count
// holds for every bfs_tile.
template <int TILE>
int64_t bfs_carve(const Graph *g, bool pull, bool iso_cached, char *block, BfsBufs *b) {
    Carver cv{block};
    const int64_t ent = std::max<int64_t>(g->m, 1) + 1;
    const int64_t tiles = g->nnz / 1024 + 2;
    const int64_t words = g->m / 32 + 1;
    FrontierSet fs;
    fs.ctl = (uint64_t *)cv.take(sizeof(uint64_t) * kCtlWords);
    for (int i = 0; i < 2; ++i) {
        fs.f[i].v = nullptr;
        fs.f[i].start = (int32_t *)cv.take(sizeof(int32_t) * ent);
        fs.f[i].off = (int32_t *)cv.take(sizeof(int32_t) * ent);
        fs.f[i].tfirst = (int32_t *)cv.take(sizeof(int32_t) * tiles);
    }
    fs.cap_entries = ent;
    fs.cap_tiles = tiles;
    uint32_t *vis = (uint32_t *)cv.take(sizeof(uint32_t) * words);
    uint32_t *fb = pull ? (uint32_t *)cv.take(sizeof(uint32_t) * 3 * words) : nullptr;
    uint32_t *iso = pull && !iso_cached ? (uint32_t *)cv.take(sizeof(uint32_t) * words) : nullptr;
    if (b) {
        b->fs = fs;
        b->vis = vis;
        b->fb = fb;
        b->iso = iso;
    }
    return cv.used;
}

}  // namespace

void bfs_run(Ctx *c, const Graph *g, int32_t source, int32_t *depth) {
    switch (opt(c, "bfs_tile", 2048)) {
    case 1024: bfs_levels_pool<1024>(c, g, source, depth); break;
    case 4096: bfs_levels_pool<4096>(c, g, source, depth); break;
    default: bfs_levels_pool<2048>(c, g, source, depth); break;
    }
}

// Pulls only on a graph already known to be symmetric (built symmetrized, or
// checked by an earlier call): checking needs a transpose, which allocates.
static bool bfs_in_pulls(Ctx *c, const Graph *g) {
    return opt(c, "bfs_direction", 1) != 0 && graph_symmetry_known(const_cast<Graph *>(g)) == 1;
}

int64_t bfs_block_bytes(Ctx *c, const Graph *g) {
    const bool pull = bfs_in_pulls(c, g);
    return bfs_carve<1024>(g, pull, graph_iso_bits_cached(const_cast<Graph *>(g)) != nullptr,
                           nullptr, nullptr);
}

void bfs_run_in(Ctx *c, const Graph *g, int32_t source, int32_t *depth, void *block,
                int64_t bytes) {
    const bool pull = bfs_in_pulls(c, g);
    const uint32_t *iso_cached = graph_iso_bits_cached(const_cast<Graph *>(g));
    BfsBufs b;
    const int64_t need = bfs_carve<1024>(g, pull, iso_cached != nullptr, (char *)block, &b);
    APSPIE_CHECK(need <= bytes, APSPIE_ERR_OVERFLOW,
                 "BFS needs " + std::to_string(need) + " bytes of scratch, the block holds " +
                     std::to_string(bytes));
    if (pull && iso_cached) b.iso = iso_cached;
    else if (pull) iso_bits_into(c, g, const_cast<uint32_t *>(b.iso));
    b.kernel_timer = false;  // no events created inside the call
    switch (opt(c, "bfs_tile", 2048)) {
    case 1024: bfs_levels<1024>(c, g, source, depth, b); break;
    case 4096: bfs_levels<4096>(c, g, source, depth, b); break;
    default: bfs_levels<2048>(c, g, source, depth, b); break;
    }
}

}  // namespace apspie
