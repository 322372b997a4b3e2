// Batched multi-source BFS / SSSP: the frontier is a sparse m x B matrix
// (SpMSpM instead of SpMSpV).  No reference counterpart; per source the
// results equal apspie_bfs / apspie_sssp.
//
// Layout (vertex-major, HBM): W = ceil(B/32) words of bits per vertex for
// visited / current-frontier / next-frontier, B contiguous int32 depths (BFS)
// or FP32 distances (SSSP) per vertex, so one expanded edge touches
// contiguous words; the per-source [B][m] result is one tiled transpose at
// the end instead of a scattered write per discovery.
// A level is one load-balanced expansion over the vertices whose frontier row
// is non-empty (the same tile machinery as the single-source path):
//   BFS : new = front[u] & ~visited[v]; fetch-or into visited[v] decides which
//         source bits this thread discovered first -> depth writes; the
//         discovered bits go to next[v] and v joins the next frontier once.
//   SSSP: for each source bit s of front[u]: relax dist[v][s] with an atomic
//         min; improved bits go to next[v].
// then the current frontier rows are cleared for reuse.  Level loops run in
// chunks captured into a hipGraph (replayed until the frontier is empty), so
// the host pays one launch per chunk instead of two per level.
#include <algorithm>
#include <map>

#include "common.hpp"
#include "device.hpp"
#include "frontier.hpp"
#include "timing.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 2048;
constexpr int kMaxBatch = 512;  // sources per pass (larger batches are split)

// the W words of a vertex's bit row (16-byte loads when W is a multiple of 4;
// rows of W words are W * 4-byte aligned)
template <int W>
__device__ __forceinline__ void load_row(const uint32_t *p, uint32_t (&x)[W]) {
    if constexpr (W % 4 == 0) {
        const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
        for (int i = 0; i < W / 4; ++i) {
            const uint4 t = q[i];
            x[4 * i] = t.x;
            x[4 * i + 1] = t.y;
            x[4 * i + 2] = t.z;
            x[4 * i + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int w = 0; w < W; ++w) x[w] = p[w];
    }
}

// Push form of one batched BFS level: the edge u -> v carries the frontier
// bits of u not yet visited at v.  Every thread that finds such a bit sets it
// (visited, next) and writes its depth: racing threads of one level write the
// same values, so the ORs need no return value; one exchange on the vertex's
// stamp admits v to the next frontier once.
template <int W>
struct BfsBatchOp {
    const uint32_t *front;
    uint32_t *vis;
    uint32_t *next;
    int32_t *stamp;
    int32_t *depth;  // [m][32W], vertex-major (transposed at the end)
    int32_t m;
    const int32_t *lvl;  // device counter: depth of this level's discoveries
    int32_t nsrc;
    __device__ __forceinline__ bool edge(int32_t u, int32_t, int32_t v) const {
        uint32_t f[W];
        load_row<W>(front + (int64_t)u * W, f);
        uint32_t anyf = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) anyf |= f[w];
        if (!anyf) return false;
        uint32_t seen[W];
        load_row<W>(vis + (int64_t)v * W, seen);
        const int32_t level = *lvl;
        bool any = false;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            uint32_t cand = f[w] & ~seen[w];
            if (!cand) continue;
            any = true;
            __hip_atomic_fetch_or(vis + (int64_t)v * W + w, cand, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_or(next + (int64_t)v * W + w, cand, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
            while (cand) {
                const int b = __ffs(cand) - 1;
                cand &= cand - 1;
                const int s = w * 32 + b;
                if (s < nsrc) depth[(int64_t)v * (32 * W) + s] = level;
            }
        }
        return any && dev::atomic_exch_dev(stamp + v, level) != level;
    }
};

// Claims of a pull level are collected in an LDS queue and appended to the
// next frontier kQueue at a time: one 64-bit reservation per flush instead of
// one per pass (a pass covers only 256 / W * 4 vertices, and thousands of
// reservations on one counter serialise in the L2).
constexpr int kQueue = 1536;
// rows longer than kLongRow still open after their first probes are finished
// after the pass by a whole wave each (kLongCap per block and pass, every
// vertex of a pass at W = 1)
constexpr int32_t kLongRow = 64;
constexpr int kLongCap = 1024;
struct PullQueue {
    int32_t v[kQueue], rs[kQueue], d[kQueue];
    int32_t lv[kLongCap];
    uint32_t n, nl;
    uint32_t scan32[kBlock / 64 + 1];
    uint64_t scan[kBlock / 64 + 1];
    uint64_t base;
};

// LDS of the level kernel: the push branch's expansion or the pull's queue
union BatchLds {
    ExpandLds<kBlock, kTile, true> ex;
    PullQueue pq;
};

// append the queue to the next frontier (every thread of the block calls)
__device__ __forceinline__ void queue_flush(PullQueue &q, FrontierBufs nxt, uint64_t *cnt_next) {
    constexpr int QK = kQueue / kBlock;
    const uint32_t n = q.n;
    int32_t cv[QK], crs[QK], cd[QK];
#pragma unroll
    for (int k = 0; k < QK; ++k) {
        const uint32_t i = threadIdx.x + k * kBlock;
        cv[k] = i < n ? q.v[i] : 0;
        crs[k] = i < n ? q.rs[i] : 0;
        cd[k] = i < n ? q.d[i] : 0;
    }
    produce<kBlock, QK, kTile, true>(cv, crs, cd, nxt, cnt_next, q.scan, &q.base);
    __syncthreads();  // every thread has read q.n and its entries
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
}

// Writes one pulled vertex's word w: nw = the bits found and not yet visited
// (uv = the word's unvisited bits).  Returns whether anything was found.
template <int W>
__device__ __forceinline__ void pull_finish(int64_t i, uint32_t nw, uint32_t uv, uint32_t *vis,
                                            uint32_t *next, uint32_t *snap, int32_t *depth,
                                            int32_t level, int64_t v, int w) {
    // with a snapshot the level's discoveries are one full bit row per
    // vertex (the depths are read back from it at the end)
    if (snap) snap[i] = nw;
    next[i] = nw;  // every row: the buffer needs no clearing before a pull
    if (nw) {
        vis[i] = ~uv | nw;
        if (!snap)
            while (nw) {
                const int b = __ffs(nw) - 1;
                nw &= nw - 1;
                depth[v * (32 * W) + w * 32 + b] = level;
            }
    }
}

// The deferred long rows of a pull pass (q.lv[0 .. nl)), one wave per row:
// lane l ORs the frontier rows of entries l, l + 64, ...; every kLongStep
// entries the wave checks whether every unvisited bit of every word has been
// found.  Lane w < W then writes word w like pull_finish, and lane 0 appends
// a claim to the queue (room was made for nl claims).
constexpr int kLongStep = 256;
template <int W>
__device__ __noinline__ void pull_long(PullQueue &q, uint32_t nl,
                                       const int32_t *__restrict__ rowptr,
                                       const int32_t *__restrict__ col,
                                       const uint32_t *__restrict__ front, uint32_t *vis,
                                       uint32_t *next, uint32_t *snap, int32_t *depth,
                                       int32_t level) {
    const int lane = dev::lane_id();
    const int wave = threadIdx.x / 64, nwaves = kBlock / 64;
    for (uint32_t j = wave; j < nl; j += nwaves) {
        const int64_t v = q.lv[j];
        const int32_t rs = rowptr[v], re = rowptr[v + 1];
        uint32_t uv[W], acc[W];
        load_row<W>(vis + v * W, uv);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            uv[w] = ~uv[w];
            acc[w] = 0;
        }
        for (int32_t base = rs; base < re; base += kLongStep) {
#pragma unroll
            for (int t = 0; t < kLongStep / 64; ++t) {
                const int32_t e = base + t * 64 + lane;
                if (e < re) {
                    uint32_t f[W];
                    load_row<W>(front + (int64_t)col[e] * W, f);
#pragma unroll
                    for (int w = 0; w < W; ++w) acc[w] |= f[w];
                }
            }
            bool done = true;
#pragma unroll
            for (int w = 0; w < W; ++w) done = done && (dev::wave_or_u32(acc[w]) & uv[w]) == uv[w];
            if (done) break;
        }
        uint32_t mine = 0, my_uv = 0, any = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t nw = dev::wave_or_u32(acc[w]) & uv[w];
            any |= nw;
            if (lane == w) {
                mine = nw;
                my_uv = uv[w];
            }
        }
        if (lane < W) pull_finish<W>(v * W + lane, mine, my_uv, vis, next, snap, depth, level, v, lane);
        if (lane == 0 && any) {
            const uint32_t pos = atomicAdd(&q.n, 1u);
            q.v[pos] = (int32_t)v;
            q.rs[pos] = rs;
            q.d[pos] = re - rs;
        }
    }
}

// Pull form of one batched BFS level, for symmetric CSR (a row lists the
// vertex's in-neighbours): a group of W lanes owns vertex v, lane w its source
// word w.  The lane ORs word w of its neighbours' frontier rows (coalesced:
// the group reads a W-word row at once), keeps the bits not yet visited, and
// writes the visited / next words and the depths itself -- no atomics.  The
// loads of KP vertices per thread are issued stage by stage.  A row longer
// than kLongRow that is still open after its first Q neighbours (a hub whose
// sources are not all in reach at this level would otherwise be walked entry
// by entry by its own lanes: 30 ms levels at R-MAT-22) is queued in LDS and
// finished after the pass by one wave, its 64 lanes striding the row and
// checking every 256 entries whether every unvisited bit has been found.
template <int W>
__device__ __forceinline__ void batch_pull(PullQueue &q, const int32_t *__restrict__ rowptr,
                                           const int32_t *__restrict__ col, int32_t m,
                                           const uint32_t *__restrict__ front, uint32_t *vis,
                                           uint32_t *next, int32_t *depth, int32_t level,
                                           uint32_t *snap, FrontierBufs nxt, uint64_t *cnt_next) {
    constexpr int VPB = kBlock / W;  // vertices per block row
    constexpr int KP = 4;            // block rows per pass
    constexpr int Q = 4;             // neighbours loaded at once
    if (threadIdx.x == 0) {
        q.n = 0;
        q.nl = 0;
    }
    __syncthreads();
    const int g = threadIdx.x / W, w = threadIdx.x % W;
    const int lane = dev::lane_id();
    const uint64_t gmask = W == 64 ? ~0ull : (((1ull << W) - 1) << ((lane / W) * W));
    const int leader = (lane / W) * W;  // the group's first lane
    for (int64_t pbase = (int64_t)blockIdx.x * VPB * KP; pbase < m;
         pbase += (int64_t)gridDim.x * VPB * KP) {
        int32_t cv[KP], crs[KP], cd[KP], rs[KP], re[KP];
        uint32_t uv[KP], acc[KP];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int64_t v = pbase + (int64_t)k * VPB + g;
            cv[k] = (int32_t)v;
            crs[k] = 0;
            cd[k] = 0;
            acc[k] = 0;
            const bool in = v < m;
            uv[k] = in ? ~vis[v * W + w] : 0u;
            rs[k] = in ? rowptr[v] : 0;
            re[k] = in ? rowptr[v + 1] : 0;
        }
        // first Q neighbours of every row, then the rest row by row
        int32_t u[KP][Q];
#pragma unroll
        for (int k = 0; k < KP; ++k)
#pragma unroll
            for (int q = 0; q < Q; ++q)
                u[k][q] = uv[k] && rs[k] + q < re[k] ? col[rs[k] + q] : -1;
#pragma unroll
        for (int k = 0; k < KP; ++k)
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if (u[k][q] >= 0) acc[k] |= front[(int64_t)u[k][q] * W + w];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            // an open long row goes to the LDS list (the group's leader
            // reserves a slot); the pass then finishes it as if nothing was
            // found -- no more entries, no bits -- and the long-row pass after
            // the block's next barrier rewrites its words
            if (__ballot(re[k] - rs[k] > kLongRow) != 0) {  // uniform: a long row in this wave
                const bool open = (__ballot((acc[k] & uv[k]) != uv[k]) & gmask) != 0;
                int slot = -1;
                if (open && re[k] - rs[k] > kLongRow && w == 0)
                    slot = (int)atomicAdd(&q.nl, 1u);
                slot = __shfl(slot, leader, 64);
                if (slot >= 0 && slot < kLongCap) {
                    if (w == 0) q.lv[slot] = cv[k];
                    acc[k] = 0;
                    re[k] = rs[k] + Q;
                }
            }
            for (int32_t e = rs[k] + Q; e < re[k] && (acc[k] & uv[k]) != uv[k]; e += Q) {
                int32_t uu[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) uu[q] = e + q < re[k] ? col[e + q] : -1;
#pragma unroll
                for (int q = 0; q < Q; ++q)
                    if (uu[q] >= 0) acc[k] |= front[(int64_t)uu[q] * W + w];
            }
        }
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const uint32_t nw = acc[k] & uv[k];
            if (cv[k] < m)
                pull_finish<W>((int64_t)cv[k] * W + w, nw, uv[k], vis, next, snap, depth, level,
                               cv[k], w);
            const uint64_t bm = __ballot(nw != 0);
            if (w == 0 && (bm & gmask)) {
                crs[k] = rs[k];
                cd[k] = re[k] - rs[k];
            }
        }
        // append this pass's claims to the queue (flushing it first if full)
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < KP; ++k) c += cd[k] > 0 ? 1u : 0u;
        uint32_t total;
        uint32_t pos = dev::block_exclusive_add32<kBlock>(c, q.scan32, total);
        if (q.n + total > (uint32_t)kQueue) queue_flush(q, nxt, cnt_next);
        pos += q.n;
#pragma unroll
        for (int k = 0; k < KP; ++k)
            if (cd[k] > 0) {
                q.v[pos] = cv[k];
                q.rs[pos] = crs[k];
                q.d[pos] = cd[k];
                ++pos;
            }
        __syncthreads();  // every thread has read q.n
        if (threadIdx.x == 0) q.n += total;
        __syncthreads();
        // the pass's long rows, one wave each; their claims join the queue
        const uint32_t nl = min(q.nl, (uint32_t)kLongCap);
        if (nl > 0) {
            if (q.n + nl > (uint32_t)kQueue) queue_flush(q, nxt, cnt_next);
            pull_long<W>(q, nl, rowptr, col, front, vis, next, snap, depth, level);
            __syncthreads();  // every wave is done with q.lv and q.n
            if (threadIdx.x == 0) q.nl = 0;
            __syncthreads();
        }
    }
    if (q.n > 0) queue_flush(q, nxt, cnt_next);
}

// Snapshots of pulled levels: the j-th pulled level writes its discovered
// bits for every vertex to snap + j * stride and its level number to
// level_of[j], up to cap levels (later pulled levels write depths directly):
// one full bit row per vertex instead of a scattered 4-byte depth store per
// discovered (vertex, source) pair.  npull[l & 1] holds the number of
// snapshots taken before level l (two slots: a level reads one and block 0
// writes the other).
struct BatchSnap {
    uint32_t *snap;
    int64_t stride;
    int32_t cap;
    int32_t *level_of;
    uint64_t *npull;
};

// One batched BFS level: push (load-balanced expansion of the union
// frontier) or pull, chosen identically by every block from the union
// frontier's edge count: pull when E * pull_div > nnz (pull_div 0: never).
// The two forms are separate kernels (each level launches both; the one
// whose form was not chosen returns at once), so the pull kernel is not held
// to the push expansion's register count.
template <int W, bool PULL>
__device__ __forceinline__ void batch_level(BatchLds &lds, const int32_t *__restrict__ rowptr,
                                            const int32_t *__restrict__ col, int32_t nnz,
                                            int32_t m, BfsBatchOp<W> &op, uint32_t *next_w,
                                            int32_t pull_div, FrontierBufs cur,
                                            const uint64_t *cnt_cur, FrontierBufs nxt,
                                            uint64_t *cnt_next, uint64_t *cnt_clear,
                                            uint64_t *stats, BatchSnap sn) {
    const uint64_t cc = *cnt_cur;
    const int32_t F = (int32_t)cnt_entries(cc);
    const int32_t E = (int32_t)cnt_edges(cc);
    const bool pull = pull_div > 0 && (int64_t)E * pull_div > (int64_t)nnz;
    if (pull != PULL) return;
    const int32_t level = *op.lvl;
    const uint64_t j = sn.npull[level & 1];
    const bool take = pull && (int64_t)j < (int64_t)sn.cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *cnt_clear = 0;
        if (F > 0) {
            stats[0] += (uint64_t)E;  // union-frontier edges
            stats[1] += (uint64_t)F;  // union-frontier entries
            stats[2] += 1;            // levels with work
            if (pull) stats[3] += 1;  // of which pulled
            sn.npull[(level + 1) & 1] = j + (take ? 1 : 0);
            if (take) sn.level_of[j] = level;
        }
    }
    if (F == 0) return;
    if constexpr (PULL)
        batch_pull<W>(lds.pq, rowptr, col, m, op.front, op.vis, next_w, op.depth, level,
                      take ? sn.snap + (int64_t)j * sn.stride : nullptr, nxt, cnt_next);
    else
        expand_tiles<kBlock, kTile, true, true, true>(lds.ex, op, rowptr, col, nnz, cur, F, E,
                                                       nxt, cnt_next);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_batch_bfs_push(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int32_t nnz, int32_t m,
    BfsBatchOp<W> op, uint32_t *next_w, int32_t pull_div, FrontierBufs cur,
    const uint64_t *cnt_cur, FrontierBufs nxt, uint64_t *cnt_next, uint64_t *cnt_clear,
    uint64_t *stats, BatchSnap sn) {
    __shared__ BatchLds lds;
    batch_level<W, false>(lds, rowptr, col, nnz, m, op, next_w, pull_div, cur, cnt_cur, nxt,
                          cnt_next, cnt_clear, stats, sn);
}

template <int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_batch_bfs_pull(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int32_t nnz, int32_t m,
    BfsBatchOp<W> op, uint32_t *next_w, int32_t pull_div, FrontierBufs cur,
    const uint64_t *cnt_cur, FrontierBufs nxt, uint64_t *cnt_next, uint64_t *cnt_clear,
    uint64_t *stats, BatchSnap sn) {
    __shared__ BatchLds lds;
    batch_level<W, true>(lds, rowptr, col, nnz, m, op, next_w, pull_div, cur, cnt_cur, nxt,
                         cnt_next, cnt_clear, stats, sn);
}

// Near-far rounds for the batch (the single-source scheme with one threshold
// T shared by all sources): a relaxation that lowers dist[v][s] below T sets
// bit s of next[v]; one that leaves it at or above T sets bit s of pend[v].
// A round that finds the union frontier empty is a split round: T grows by
// the step (delta, doubled by each split after a split that moved nothing,
// see nf_next_step) and a sweep (a group of W lanes per vertex, lane w owning word w)
// moves the pending bits now below T into next, drops the ones that fell
// below the old T since (already expanded) and keeps the rest.
// ctl: [10 + (r & 1)] the step of round r, [12 + (r & 1)] T of round r,
// [14 + (r & 1)] round r - 1 was a split,
// pending bits left by the split of round r in the spread slots (r % 3).
template <int W>
struct SsspBatchNearOp {
    const uint32_t *front;
    uint32_t *next;
    uint32_t *pend;  // null: no threshold (every improvement is near)
    int32_t *stamp;
    int32_t *dist;  // [m][32W] float bits
    const float *wt;
    const int32_t *lvl;  // device round counter
    int32_t tbits;
    // the edge u -> v (e = CSR position): relax every source of front[u]
    // (the row read at once), four sources at a time with their distance
    // loads in flight.  An improvement is judged against the distance read,
    // and the atomic min is applied without waiting for it: a racing lower
    // write can leave a spurious frontier bit, which costs a re-expansion,
    // never a wrong result.
    __device__ __forceinline__ bool edge(int32_t u, int32_t e, int32_t v) const {
        constexpr int B = 32 * W;
        uint32_t fr[W];
        load_row<W>(front + (int64_t)u * W, fr);
        uint32_t anyf = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) anyf |= fr[w];
        if (!anyf) return false;
        const float we = wt[e];
        bool any = false;
        for (int w = 0; w < W; ++w) {
            uint32_t f = fr[w];
            if (!f) continue;
            const int32_t *du_row = dist + (int64_t)u * B + w * 32;
            int32_t *dv_row = dist + (int64_t)v * B + w * 32;
            uint32_t near = 0, far = 0;
            while (f) {
                int b[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    b[q] = f ? __ffs(f) - 1 : -1;
                    f &= f - 1;
                }
                int32_t du[4], dv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    du[q] = b[q] >= 0 ? du_row[b[q]] : 0;
                    dv[q] = b[q] >= 0 ? dv_row[b[q]] : 0;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (b[q] < 0) continue;
                    const int32_t nd = dev::fbits(dev::bitsf(du[q]) + we);
                    if (nd >= dv[q]) continue;
                    __hip_atomic_fetch_min(dv_row + b[q], nd, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    if (!pend || nd < tbits) near |= 1u << b[q];
                    else far |= 1u << b[q];
                }
            }
            if (far)
                __hip_atomic_fetch_or(pend + (int64_t)v * W + w, far, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            if (near) {
                __hip_atomic_fetch_or(next + (int64_t)v * W + w, near, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
                any = true;
            }
        }
        if (!any) return false;
        const int32_t round = *lvl;
        return dev::atomic_exch_dev(stamp + v, round) != round;
    }
};

// Pulled form of one batched SSSP round, for a CSR whose rows are all short
// and whose weights are symmetric (w(u, v) sits in v's row too): a group of W
// lanes owns vertex v, lane w its source word w.  The lane walks v's row; a
// neighbour u whose frontier word w has bits set relaxes each of those
// sources into v's own distances, which the lane reads and writes with
// plain stores -- nothing else writes them this round.  The improved bits
// go to next[v] (below the threshold, or always without one) or pend[v],
// written by the lane alone; a vertex with a near bit joins the next
// frontier through the block's LDS queue.  Every row of next is written, so
// the buffer needs no clearing before a pulled round.  A neighbour's
// distance read while its owner lowers it is either value, both upper
// bounds: the relaxations are those of the push, and so is the fixed point.
template <int W>
__device__ __forceinline__ void batch_sssp_pull(PullQueue &q, const int32_t *__restrict__ rowptr,
                                                const int32_t *__restrict__ col, int32_t m,
                                                const SsspBatchNearOp<W> &op, int32_t tbits,
                                                FrontierBufs nxt, uint64_t *cnt_next) {
    constexpr int B = 32 * W;
    constexpr int VPB = kBlock / W;
    // vertices per group and pass, measured on the 2048^2 lattice (ms per
    // query): one for rows of one word (32 sources: 0.826 with two, 0.812
    // with one), two for rows of two (64 sources: 0.685, 0.697 with four),
    // four from four words on (128 / 256 / 512 sources: 0.616 / 0.598 /
    // 0.597 with two, 0.611 / 0.591 / 0.579 with four)
    constexpr int KP = W == 1 ? 1 : W == 2 ? 2 : 4;
    static_assert(KP * kBlock <= kQueue, "one pass's claims fit the LDS queue");
    constexpr int Q = 4;  // row entries loaded at once
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
    const int g = threadIdx.x / W, w = threadIdx.x % W;
    const int lane = dev::lane_id();
    const uint64_t gmask = W == 64 ? ~0ull : (((1ull << W) - 1) << ((lane / W) * W));
    for (int64_t pbase = (int64_t)blockIdx.x * VPB * KP; pbase < m;
         pbase += (int64_t)gridDim.x * VPB * KP) {
        int32_t cv[KP], crs[KP], cd[KP];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int64_t v = pbase + (int64_t)k * VPB + g;
            cv[k] = (int32_t)v;
            crs[k] = 0;
            cd[k] = 0;
            uint32_t near = 0, far = 0;
            int32_t rs = 0, re = 0;
            if (v < m) {
                rs = rowptr[v];
                re = rowptr[v + 1];
                int32_t *dv_row = op.dist + v * B + w * 32;
                for (int32_t e0 = rs; e0 < re; e0 += Q) {
                    int32_t u[Q];
                    uint32_t f[Q];
#pragma unroll
                    for (int t = 0; t < Q; ++t) u[t] = e0 + t < re ? col[e0 + t] : -1;
#pragma unroll
                    for (int t = 0; t < Q; ++t) f[t] = u[t] >= 0 ? op.front[(int64_t)u[t] * W + w] : 0u;
#pragma unroll
                    for (int t = 0; t < Q; ++t) {
                        if (!f[t]) continue;
                        const float we = op.wt[e0 + t];
                        const int32_t *du_row = op.dist + (int64_t)u[t] * B + w * 32;
                        uint32_t x = f[t];
                        while (x) {  // four sources at a time, their loads in flight
                            int b[4];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                b[j] = x ? __ffs(x) - 1 : -1;
                                x &= x - 1;
                            }
                            int32_t du[4], dv[4];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                du[j] = b[j] >= 0 ? du_row[b[j]] : 0;
                                dv[j] = b[j] >= 0 ? dv_row[b[j]] : 0;
                            }
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                if (b[j] < 0) continue;
                                const int32_t nd = dev::fbits(dev::bitsf(du[j]) + we);
                                if (nd < dv[j]) {
                                    dv_row[b[j]] = nd;
                                    if (!op.pend || nd < tbits) near |= 1u << b[j];
                                    else far |= 1u << b[j];
                                }
                            }
                        }
                    }
                }
                far &= ~near;
                const int64_t i = v * W + w;
                op.next[i] = near;
                if (far) op.pend[i] |= far;
            }
            const uint64_t bm = __ballot(near != 0);
            if (w == 0 && (bm & gmask)) {
                crs[k] = rs;
                cd[k] = re - rs;
            }
        }
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < KP; ++k) c += cd[k] > 0 ? 1u : 0u;
        uint32_t total;
        uint32_t pos = dev::block_exclusive_add32<kBlock>(c, q.scan32, total);
        if (q.n + total > (uint32_t)kQueue) queue_flush(q, nxt, cnt_next);
        pos += q.n;
#pragma unroll
        for (int k = 0; k < KP; ++k)
            if (cd[k] > 0) {
                q.v[pos] = cv[k];
                q.rs[pos] = crs[k];
                q.d[pos] = cd[k];
                ++pos;
            }
        __syncthreads();  // every thread has read q.n
        if (threadIdx.x == 0) q.n += total;
        __syncthreads();
    }
    if (q.n > 0) queue_flush(q, nxt, cnt_next);
}

// Bellman-Ford rounds of the batch (no threshold): every improvement is near.
template <int W>
__global__ __launch_bounds__(kBlock) void k_batch_sssp_bf(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int32_t nnz,
    SsspBatchNearOp<W> op, FrontierBufs cur, const uint64_t *cnt_cur, FrontierBufs nxt,
    uint64_t *cnt_next, uint64_t *cnt_clear, uint64_t *stats) {
    __shared__ BatchLds lds;
    const uint64_t cc = *cnt_cur;
    const int32_t F = (int32_t)cnt_entries(cc), E = (int32_t)cnt_edges(cc);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *cnt_clear = 0;
        if (F > 0) {
            stats[0] += (uint64_t)E;
            stats[1] += (uint64_t)F;
            stats[2] += 1;
        }
    }
    if (F == 0) return;
    expand_tiles<kBlock, kTile, true, true, true>(lds.ex, op, rowptr, col, nnz, cur, F, E, nxt,
                                                   cnt_next);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_batch_sssp_nf(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int32_t nnz, int32_t m,
    SsspBatchNearOp<W> op, FrontierBufs cur, FrontierBufs nxt, uint64_t *ctl,
    int32_t r, int32_t pull_div) {
    constexpr int B = 32 * W;
    __shared__ BatchLds lds;
    __shared__ uint64_t s_part[kBlock / 64];
    __shared__ uint32_t s_min[kBlock / 64];
    const uint64_t cc = ctl[r % 3];
    const int32_t F = (int32_t)cnt_entries(cc), E = (int32_t)cnt_edges(cc);
    const int32_t tbits = (int32_t)ctl[12 + (r & 1)];
    const bool split_last = ctl[14 + (r & 1)] != 0;
    const uint64_t left_before = nf_left(ctl, r + 2);
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
        }
    }
    if (done) return;
    if (!split && pull_div > 0 && (int64_t)E * pull_div > (int64_t)nnz) {
        batch_sssp_pull<W>(lds.pq, rowptr, col, m, op, tbits, nxt, ctl + (r + 1) % 3);
        return;
    }
    if (!split) {
        op.tbits = tbits;
        expand_tiles<kBlock, kTile, true, true, true>(lds.ex, op, rowptr, col, nnz, cur, F, E,
                                                       nxt, ctl + (r + 1) % 3);
        return;
    }
    // split: one group of W lanes per vertex
    constexpr int VPB = kBlock / W;
    constexpr int KP = 4;
    PullQueue &q = lds.pq;
    if (threadIdx.x == 0) {
        q.n = 0;
        q.nl = 0;
    }
    __syncthreads();
    const int g = threadIdx.x / W, w = threadIdx.x % W;
    const int lane = dev::lane_id();
    const uint64_t gmask = W == 64 ? ~0ull : (((1ull << W) - 1) << ((lane / W) * W));
    const int32_t round = *op.lvl;
    uint64_t left = 0;
    uint32_t min_left = ~0u;  // smallest distance left pending (bits)
    for (int64_t pbase = (int64_t)blockIdx.x * VPB * KP; pbase < m;
         pbase += (int64_t)gridDim.x * VPB * KP) {
        int32_t cv[KP], crs[KP], cd[KP];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int64_t v = pbase + (int64_t)k * VPB + g;
            cv[k] = (int32_t)v;
            crs[k] = 0;
            cd[k] = 0;
            uint32_t moved = 0;
            if (v < m) {
                const int64_t i = v * W + w;
                uint32_t pb = op.pend[i];
                if (pb) {
                    uint32_t keep = 0, x = pb;
                    while (x) {
                        const int b = __ffs(x) - 1;
                        x &= x - 1;
                        const int32_t d = op.dist[v * B + w * 32 + b];
                        if (d < tbits) continue;  // lowered below T since: expanded
                        if (d < tnext) {
                            moved |= 1u << b;
                        } else {
                            keep |= 1u << b;
                            min_left = min(min_left, (uint32_t)d);
                        }
                    }
                    op.pend[i] = keep;
                    left += (uint64_t)__popc(keep);
                    if (moved) op.next[i] = moved;
                }
            }
            const uint64_t bm = __ballot(moved != 0);
            if (w == 0 && (bm & gmask)) {
                op.stamp[cv[k]] = round;
                const int32_t rs = rowptr[cv[k]];
                crs[k] = rs;
                cd[k] = rowptr[cv[k] + 1] - rs;
            }
        }
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < KP; ++k) c += cd[k] > 0 ? 1u : 0u;
        uint32_t total;
        uint32_t pos = dev::block_exclusive_add32<kBlock>(c, q.scan32, total);
        if (q.n + total > (uint32_t)kQueue) queue_flush(q, nxt, ctl + (r + 1) % 3);
        pos += q.n;
#pragma unroll
        for (int k = 0; k < KP; ++k)
            if (cd[k] > 0) {
                q.v[pos] = cv[k];
                q.rs[pos] = crs[k];
                q.d[pos] = cd[k];
                ++pos;
            }
        __syncthreads();
        if (threadIdx.x == 0) q.n += total;
        __syncthreads();
    }
    if (q.n > 0) queue_flush(q, nxt, ctl + (r + 1) % 3);
    const uint64_t ws = dev::wave_sum_u64(left);
    const uint32_t wm = dev::wave_min_u32(min_left);
    if (lane == 0) {
        s_part[threadIdx.x / 64] = ws;
        s_min[threadIdx.x / 64] = wm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        uint32_t mn = ~0u;
        for (int qq = 0; qq < kBlock / 64; ++qq) {
            t += s_part[qq];
            mn = min(mn, s_min[qq]);
        }
        nf_report(ctl, r, t, mn);
    }
}

// visited words at the start: bits of sources beyond nsrc count as visited
// (the last word of a batch), and with a symmetric CSR (comp given) a vertex
// counts as visited by every source of another connected component -- it is
// never reached from them, and a pulled level would otherwise scan its whole
// row for their bits at every pulled level (a vertex without neighbours is
// its own component).
template <int W>
__global__ void k_batch_vis_init(int32_t m, int32_t nsrc, const int32_t *__restrict__ comp,
                                 const int32_t *__restrict__ srcs, uint32_t *vis) {
    constexpr int B = 32 * W;
    __shared__ int32_t s_comp[B];
    for (int i = threadIdx.x; i < B; i += blockDim.x)
        s_comp[i] = comp && i < nsrc ? comp[srcs[i]] : -1;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)m * W;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i / W;
        const int w = (int)(i % W);
        const int valid = min(max(nsrc - w * 32, 0), 32);
        uint32_t x = valid == 32 ? 0u : ~((1u << valid) - 1u);
        if (comp) {
            const int32_t cv = comp[v];
            for (int b = 0; b < valid; ++b)
                if (s_comp[w * 32 + b] != cv) x |= 1u << b;
        }
        vis[i] = x;
    }
}

// zero the frontier rows of the entries just expanded; advance the level
// (every round when `every`: near-far rounds number their claims by round,
// split rounds included).  The rows are left alone when the next level pulls
// (pull_div > 0 and its union frontier has more than nnz / pull_div edges):
// that level writes every row of this buffer as its next frontier.
template <int W>
__global__ void k_clear_rows(FrontierBufs cur, const uint64_t *cnt, uint32_t *front,
                             int32_t *lvl, bool every, const uint64_t *cnt_next,
                             int32_t pull_div, int32_t nnz) {
    const int32_t F = (int32_t)cnt_entries(*cnt);
    if (blockIdx.x == 0 && threadIdx.x == 0 && (F > 0 || every)) *lvl += 1;
    if (pull_div > 0 && (int64_t)cnt_edges(*cnt_next) * pull_div > (int64_t)nnz) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)F * W;
         i += (int64_t)gridDim.x * blockDim.x)
        front[(int64_t)cur.v[i / W] * W + (i % W)] = 0;
}

// seed: source i -> bit i of its vertex; depth/dist 0 at the source
template <int W>
__global__ void k_batch_seed(const int32_t *srcs, int32_t nsrc, uint32_t *vis, uint32_t *front,
                             int32_t *depth, int32_t m, int32_t *dist) {
    for (int i = threadIdx.x; i < nsrc; i += blockDim.x) {
        const int32_t s = srcs[i];
        if (vis) atomicOr(vis + (int64_t)s * W + i / 32, 1u << (i % 32));
        atomicOr(front + (int64_t)s * W + i / 32, 1u << (i % 32));
        if (depth) depth[(int64_t)s * (32 * W) + i] = 0;
        if (dist) dist[(int64_t)s * (32 * W) + i] = 0;
    }
}

struct SeedOp {
    __device__ __forceinline__ bool keep(int32_t) const { return true; }
};

// dist [m][B] -> out [nsrc][m]
__global__ void k_transpose_dist(const int32_t *dist, int32_t m, int32_t B, int32_t nsrc,
                                 float *out) {
    __shared__ float tile[32][33];
    const int64_t v0 = (int64_t)blockIdx.x * 32, s0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int64_t v = v0 + r, s = s0 + tx;
        if (v < m && s < nsrc) tile[r][tx] = dev::bitsf(dist[v * B + s]);
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t s = s0 + r, v = v0 + tx;
        if (v < m && s < nsrc) out[s * m + v] = tile[tx][r];
    }
}

__global__ void k_batch_npull(uint64_t *npull, const int32_t *lvl) { npull[0] = npull[*lvl & 1]; }

// Depths of one pass into out [nsrc][m], T vertices per block through LDS
// (T * W = 256 words, one per thread): pairs discovered in a snapshotted
// level take that level's number; the others are reached (visited bit set)
// with the depth written where they were discovered (vertex-major dep
// [m][32W]), or unreached (-1).  With comp given, pairs of a vertex and a
// source in different components were marked visited up front and are
// unreached.  The snapshots are read kSnapRun at a time (independent loads
// in flight).
constexpr int kSnapRun = 16;
template <int W>
__global__ __launch_bounds__(256) void k_batch_depths(
    const int32_t *dep, const uint32_t *vis, const int32_t *comp, const int32_t *srcs,
    int32_t m, int32_t nsrc, const uint32_t *snap, int64_t stride, const uint64_t *npull,
    const int32_t *level_of, int32_t *out) {
    constexpr int B = 32 * W, T = 256 / W;
    __shared__ int32_t d[T][B + 1];
    __shared__ int32_t s_comp[B];
    for (int i = threadIdx.x; i < B; i += 256) s_comp[i] = comp && i < nsrc ? comp[srcs[i]] : -1;
    const int64_t v0 = (int64_t)blockIdx.x * T;
    const int t = threadIdx.x / W, w = threadIdx.x % W;
    const int64_t v = v0 + t;
    for (int b = 0; b < 32; ++b) d[t][w * 32 + b] = -2;
    const int32_t ns = (int32_t)npull[0];
    for (int32_t j0 = 0; j0 < ns; j0 += kSnapRun) {
        uint32_t bits[kSnapRun];
#pragma unroll
        for (int q = 0; q < kSnapRun; ++q)
            bits[q] = (j0 + q < ns && v < m) ? snap[(int64_t)(j0 + q) * stride + v * W + w] : 0u;
#pragma unroll
        for (int q = 0; q < kSnapRun; ++q) {
            if (!bits[q]) continue;
            const int32_t L = level_of[j0 + q];
            while (bits[q]) {
                const int b = __ffs(bits[q]) - 1;
                bits[q] &= bits[q] - 1;
                d[t][w * 32 + b] = L;
            }
        }
    }
    __syncthreads();
    for (int x = threadIdx.x; x < T * nsrc; x += 256) {
        const int s = x / T, tt = x % T;
        const int64_t vv = v0 + tt;
        if (vv >= m) continue;
        int32_t r = d[tt][s];
        if (r == -2) {
            r = -1;
            if ((vis[vv * W + s / 32] >> (s % 32)) & 1u)
                if (!comp || comp[vv] == s_comp[s]) r = dep[vv * B + s];
        }
        out[(int64_t)s * m + vv] = r;
    }
}

// One batched traversal pass for up to 32*W sources.
template <int W, bool SSSP>
void batch_pass(Ctx *c, const Graph *g, const int32_t *src_host, int32_t nsrc, int32_t *depth_out,
                float *dist_out) {
    constexpr int B = 32 * W;
    const int32_t m = g->m;
    const int64_t words = (int64_t)m * W;
    FrontierSet fs = frontier_alloc(c, "batch", m, g->nnz, kTile, true);
    uint32_t *vis = SSSP ? nullptr : (uint32_t *)scratch(c, "batch_vis", sizeof(uint32_t) * words);
    uint32_t *fa = (uint32_t *)scratch(c, "batch_fa", sizeof(uint32_t) * words);
    uint32_t *fb = (uint32_t *)scratch(c, "batch_fb", sizeof(uint32_t) * words);
    int32_t *stamp = (int32_t *)scratch(c, "batch_stamp", sizeof(int32_t) * m);
    int32_t *dist = SSSP ? (int32_t *)scratch(c, "batch_dist", sizeof(int32_t) * m * (int64_t)B)
                         : nullptr;
    int32_t *depth = SSSP ? nullptr
                          : (int32_t *)scratch(c, "batch_dep", sizeof(int32_t) * m * (int64_t)B);
    // snapshots of pulled BFS levels, within batch_snapshot_mb (and half the
    // room the context's HBM budget or arena leaves; a snapshot buffer kept
    // from an earlier batch counts as room)
    BatchSnap sn{};
    if (!SSSP) {
        int64_t bytes = opt(c, "batch_snapshot_mb", 24576) << 20;
        if (c->budget > 0 || c->arena) {
            auto kept = c->scratch.find("batch_snap");
            const int64_t have = kept != c->scratch.end() ? (int64_t)kept->second.bytes : 0;
            bytes = std::min<int64_t>(bytes, std::max<int64_t>(have, scratch_room(c) / 2));
        }
        sn.stride = words;
        sn.cap = (int32_t)std::min<int64_t>(bytes / (int64_t)(sizeof(uint32_t) * words), 4096);
        sn.snap = sn.cap > 0 ? (uint32_t *)scratch(c, "batch_snap",
                                                   sizeof(uint32_t) * words * (int64_t)sn.cap)
                             : nullptr;
        sn.level_of = (int32_t *)scratch(c, "batch_snap_lv", sizeof(int32_t) * 4096);
        sn.npull = fs.ctl + 12;
    }
    int32_t *srcs = (int32_t *)scratch(c, "batch_srcs", sizeof(int32_t) * (B + 1));
    APSPIE_HIP(hipMemcpyAsync(srcs, src_host, sizeof(int32_t) * nsrc, hipMemcpyHostToDevice,
                              c->stream));
    // distinct source vertices form the first frontier
    std::vector<int32_t> uniq(src_host, src_host + nsrc);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    int32_t *ulist = (int32_t *)scratch(c, "batch_ulist", sizeof(int32_t) * (B + 1));
    APSPIE_HIP(hipMemcpyAsync(ulist, uniq.data(), sizeof(int32_t) * uniq.size(),
                              hipMemcpyHostToDevice, c->stream));
    const int grid = fill_grid(c, kBlock, (int)opt(c, "batch_blocks_per_cu", 4));
    // near-far rounds for SSSP (threshold step batch_sssp_delta, 0 = off,
    // -1 = chosen per graph; passes of up to 128 sources take twice that
    // width: fewer rounds, whose fixed cost weighs more on a narrow row --
    // 2048^2 lattice, 32 / 64 / 128 sources: 0.876 -> 0.829 / 0.705 -> 0.682
    // / 0.625 -> 0.616 ms per query; 512 sources 0.596 -> 0.603)
    double bdelta = SSSP ? (double)opt(c, "batch_sssp_delta", -1) : 0;
    if (bdelta < 0) bdelta = sssp_batch_auto_delta(c, const_cast<Graph *>(g)) * (W <= 4 ? 2 : 1);
    const float delta = (float)bdelta;
    const bool near_far = SSSP && delta > 0;
    uint32_t *pend =
        near_far ? (uint32_t *)scratch(c, "batch_pend", sizeof(uint32_t) * words) : nullptr;
    // pulling reads a row as in-edges: only for a symmetric CSR
    const bool sym = !SSSP && graph_is_symmetric(c, const_cast<Graph *>(g));
    // components (built once per graph): pairs in different components start settled
    const int32_t *comp = sym ? graph_components(c, const_cast<Graph *>(g)) : nullptr;
    // A pulled level reads W words per vertex and a pushed one W words per
    // expanded edge, so narrow batches (W <= 4) on graphs whose rows are all
    // short gain from pulling earlier: their threshold divisor is four times
    // the option (2048^2 lattice: 32 sources 7.7x -> 10.1x the plain rate).
    // On a graph with long rows an early pull is a trap: a hub keeps scanning
    // its whole row for sources whose frontier has not reached it (RMAT-22, 64
    // sources: 0.29 -> 3.5 ms per query), so there the option applies as is.
    const int64_t div_opt = opt(c, "batch_pull_div", 4);
    const bool short_rows = sym && graph_max_degree(c, const_cast<Graph *>(g)) <= 64;
    const int32_t pull_div =
        sym ? (int32_t)std::min<int64_t>(div_opt * (W <= 4 && short_rows ? 4 : 1), 1 << 22)
            : 0;
    // near-far SSSP rounds pull on graphs whose rows are all short and whose
    // weights are symmetric (a pulled round walks every vertex's whole row);
    // rows of one or two words pull at twice the option's divisor (2048^2
    // lattice, 32 / 64 sources: 0.94 -> 0.88 / 0.73 -> 0.71 ms per query;
    // 128 sources are best at the option's 4)
    const int32_t sssp_pull_div =
        near_far && opt(c, "batch_sssp_pull_div", 4) > 0 &&
                graph_max_degree(c, const_cast<Graph *>(g)) <= 64 &&
                graph_is_symmetric_weighted(c, const_cast<Graph *>(g))
            ? (int32_t)std::min<int64_t>(opt(c, "batch_sssp_pull_div", 4) * (W <= 2 ? 2 : 1),
                                         1 << 22)
            : 0;
    if (!SSSP) {
        k_batch_vis_init<W><<<grid, kBlock, 0, c->stream>>>(m, nsrc, comp, srcs, vis);
    } else {
        fill_i32(c, dist, __builtin_bit_cast(int32_t, APSPIE_SSSP_INF), (int64_t)m * B);
    }
    fill_i32(c, (int32_t *)fa, 0, words);
    fill_i32(c, (int32_t *)fb, 0, words);
    fill_i32(c, stamp, -1, m);
    if (pend) fill_i32(c, (int32_t *)pend, 0, words);
    fill_i32(c, (int32_t *)fs.ctl, 0, 2 * kCtlWords);
    int32_t *lvl = (int32_t *)(fs.ctl + 8);  // level / round counter
    fill_i32(c, lvl, 1, 1);
    if (near_far) {  // T of round 0 and its split step
        fill_i32(c, (int32_t *)(fs.ctl + 12), __builtin_bit_cast(int32_t, delta), 1);
        fill_i32(c, (int32_t *)(fs.ctl + 10), __builtin_bit_cast(int32_t, delta), 1);
    }
    k_batch_seed<W><<<1, 256, 0, c->stream>>>(srcs, nsrc, vis, fa, depth, m, dist);
    k_build<kBlock, 1, kTile, true><<<1, kBlock, 0, c->stream>>>(
        SeedOp{}, g->rowptr, ulist, (int32_t)uniq.size(), nullptr, fs.f[0], fs.ctl + 0);
    APSPIE_LAUNCH_CHECK();

    // level l: expand f[l&1] (bits in fa/fb alternating), clear its rows
    auto level = [&](int l) {
        const int cur = l & 1;
        uint32_t *front = cur ? fb : fa, *next = cur ? fa : fb;
        if (near_far)
            k_batch_sssp_nf<W><<<grid, kBlock, 0, c->stream>>>(
                g->rowptr, g->col, (int32_t)g->nnz, m,
                SsspBatchNearOp<W>{front, next, pend, stamp, dist, g->val, lvl, 0}, fs.f[cur],
                fs.f[cur ^ 1], fs.ctl, l % 6, sssp_pull_div);
        else if (SSSP)
            k_batch_sssp_bf<W><<<grid, kBlock, 0, c->stream>>>(
                g->rowptr, g->col, (int32_t)g->nnz,
                SsspBatchNearOp<W>{front, next, nullptr, stamp, dist, g->val, lvl, 0}, fs.f[cur],
                fs.ctl + (l % 3), fs.f[cur ^ 1], fs.ctl + ((l + 1) % 3), fs.ctl + ((l + 2) % 3),
                fs.ctl + 4);
        else {
            const BfsBatchOp<W> op{front, vis, next, stamp, depth, m, lvl, nsrc};
            k_batch_bfs_push<W><<<grid, kBlock, 0, c->stream>>>(
                g->rowptr, g->col, (int32_t)g->nnz, m, op, next, pull_div, fs.f[cur],
                fs.ctl + (l % 3), fs.f[cur ^ 1], fs.ctl + ((l + 1) % 3), fs.ctl + ((l + 2) % 3),
                fs.ctl + 4, sn);
            if (pull_div > 0)  // six 256-thread blocks per CU fit its registers and LDS
                k_batch_bfs_pull<W><<<fill_grid(c, kBlock, 6), kBlock, 0, c->stream>>>(
                    g->rowptr, g->col, (int32_t)g->nnz, m, op, next, pull_div, fs.f[cur],
                    fs.ctl + (l % 3), fs.f[cur ^ 1], fs.ctl + ((l + 1) % 3),
                    fs.ctl + ((l + 2) % 3), fs.ctl + 4, sn);
        }
        k_clear_rows<W><<<grid, kBlock, 0, c->stream>>>(fs.f[cur], fs.ctl + (l % 3), front, lvl,
                                                        near_far, fs.ctl + ((l + 1) % 3),
                                                        SSSP ? sssp_pull_div : pull_div,
                                                        (int32_t)g->nnz);
    };

    // chunks of K levels captured once and replayed: K is a multiple of 6 so
    // the counter ring (period 3) and the buffer parity (period 2) realign;
    // level numbers come from the device counter, not from kernel arguments
    const int K = 6 * (int)std::max<int64_t>(1, ceil_div(opt(c, "bfs_levels_per_sync", 8), 6));
    const bool use_graph = opt(c, "batch_graph", 1) != 0;
    hipGraphExec_t exec = nullptr;
    int l = 0;
    uint64_t *h = (uint64_t *)c->pinned;
    KernelTimer kt(c);
    for (;;) {
        kt.begin();
        if (use_graph && exec) {
            APSPIE_HIP(hipGraphLaunch(exec, c->stream));
            l += K;
        } else if (use_graph) {
            hipGraph_t graph;
            APSPIE_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; ++k) level(l + k);
            APSPIE_HIP(hipStreamEndCapture(c->stream, &graph));
            APSPIE_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            APSPIE_HIP(hipGraphDestroy(graph));
            APSPIE_HIP(hipGraphLaunch(exec, c->stream));
            l += K;
        } else {
            for (int k = 0; k < K; ++k) level(l + k);
            l += K;
        }
        kt.end();
        APSPIE_LAUNCH_CHECK();
        APSPIE_HIP(hipMemcpyAsync(h, fs.ctl, kCtlWords * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        sync(c);
        c->stats.host_syncs++;
        if (!near_far && cnt_entries(h[l % 3]) == 0) break;
        if (near_far) {  // the last round was a split that moved and left nothing
            if (cnt_entries(h[l % 3]) == 0 && h[14 + (l & 1)] != 0 && nf_left(h, l + 2) == 0)
                break;
        }
        if (l > 8 * (int64_t)m + 4096) {
            if (exec) (void)hipGraphExecDestroy(exec);
            throw Error{APSPIE_ERR_DEVICE, "batched traversal did not converge"};
        }
    }
    if (exec) APSPIE_HIP(hipGraphExecDestroy(exec));
    c->stats.edges_traversed += (int64_t)h[4];
    c->stats.vertices_reached += (int64_t)h[5];
    c->stats.iterations = std::max(c->stats.iterations, (int32_t)h[6]);
    kt.collect();
    if (SSSP) {
        dim3 grd((unsigned)ceil_div(m, 32), (unsigned)ceil_div(nsrc, 32));
        k_transpose_dist<<<grd, 256, 0, c->stream>>>(dist, m, B, nsrc, dist_out);
    } else {
        // the snapshot count after the last level sits in the slot of the level
        // that found the frontier empty; copy it to slot 0 for the reader
        k_batch_npull<<<1, 1, 0, c->stream>>>(sn.npull, lvl);
        k_batch_depths<W><<<(unsigned)ceil_div(m, 256 / W), 256, 0, c->stream>>>(
            depth, vis, comp, srcs, m, nsrc, sn.snap, sn.stride, sn.npull, sn.level_of,
            depth_out);
    }
    APSPIE_LAUNCH_CHECK();
}

template <bool SSSP>
void batch_run(Ctx *c, const Graph *g, const int32_t *src, int32_t nsrc, int32_t *depth,
               float *dist) {
    // BFS on a graph with long rows (low diameter, skewed degrees) runs in
    // groups of batch_long_row_group sources (default 256): a pulled vertex
    // keeps scanning until every word of its row has met a frontier
    // neighbour, and a few far sources keep a whole group scanning.  R-MAT-22,
    // 512 sources of the giant component, four source sets: 0.159 / 0.138 /
    // 0.154 ms per query on average in groups of 128 / 256 / 512 (0.107-0.165
    // in groups of 256; plain 0.206)
    static_assert(kQueue % kBlock == 0, "queue flushes take whole block rows");
    const int32_t group =
        !SSSP && graph_max_degree(c, const_cast<Graph *>(g)) > 64
            ? (int32_t)std::min<int64_t>(std::max<int64_t>(opt(c, "batch_long_row_group", 256), 1),
                                         kMaxBatch)
            : kMaxBatch;
    for (int32_t s0 = 0; s0 < nsrc; s0 += group) {
        const int32_t k = std::min(group, nsrc - s0);
        int32_t *d = depth ? depth + (int64_t)s0 * g->m : nullptr;
        float *f = dist ? dist + (int64_t)s0 * g->m : nullptr;
        // the narrowest bit row that holds the pass's sources (W words of 32)
        if (k <= 32) batch_pass<1, SSSP>(c, g, src + s0, k, d, f);
        else if (k <= 64) batch_pass<2, SSSP>(c, g, src + s0, k, d, f);
        else if (k <= 128) batch_pass<4, SSSP>(c, g, src + s0, k, d, f);
        else if (k <= 256) batch_pass<8, SSSP>(c, g, src + s0, k, d, f);
        else batch_pass<16, SSSP>(c, g, src + s0, k, d, f);
    }
}

}  // namespace

void bfs_batch_run(Ctx *c, const Graph *g, const int32_t *src, int32_t nsrc, int32_t *depth) {
    batch_run<false>(c, g, src, nsrc, depth, nullptr);
}

void sssp_batch_run(Ctx *c, const Graph *g, const int32_t *src, int32_t nsrc, float *dist) {
    batch_run<true>(c, g, src, nsrc, nullptr, dist);
}

}  // namespace apspie
