// Sparse frontier machinery shared by the SpMSpV-based traversals.
//
// A frontier is the sparse vector x of one SpMSpV step, stored compacted in
// HBM.  Only entries with out-degree > 0 are kept (others expand to nothing).
// Per entry j:
//   v[j]      vertex id (the nonzero's index)
//   start[j]  rowptr[v]             -- the frontier column gather
//   off[j]    sum of degrees of entries before j   -- exclusive scan of
//             column lengths, produced on the fly: a block appends its
//             entries with ONE 64-bit atomic that reserves both the entry
//             slots (high word) and the edge range (low word), so offsets are
//             monotone in slot order without a separate scan pass.
//   tfirst[t] the entry that owns edge t*TILE: the merge-path partition of
//             the expanded edge list into TILE-sized tiles, written by the
//             producer for each tile boundary that falls inside its range.
// The consumer (k_expand) walks tiles of TILE consecutive expanded edges;
// every tile finds its entries from tfirst, stages their offsets in LDS and
// maps each edge slot to its entry by binary search -- a load-balanced
// expansion whose cost does not depend on the degree skew.
#pragma once

#include <type_traits>
#include <utility>

#include "common.hpp"
#include "device.hpp"

namespace apspie {

struct FrontierBufs {
    int32_t *v;
    int32_t *start;
    int32_t *off;
    int32_t *tfirst;
};

// counter word: (entries << 32) | expanded edges
__host__ __device__ inline uint32_t cnt_entries(uint64_t c) { return uint32_t(c >> 32); }
__host__ __device__ inline uint32_t cnt_edges(uint64_t c) { return uint32_t(c); }

constexpr uint64_t kLow40 = (uint64_t(1) << 40) - 1;

// Block-wide append of up to K entries per thread.  cd[k] == 0 means "no
// entry".  All threads of the block must call it.
template <int BLOCK, int K, int TILE, bool STORE_V>
__device__ __forceinline__ void produce(const int32_t (&cv)[K], const int32_t (&crs)[K],
                                        const int32_t (&cd)[K], FrontierBufs nxt,
                                        uint64_t *cnt_next, uint64_t *s_scan,
                                        uint64_t *s_base) {
    uint32_t c = 0, dsum = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        c += cd[k] > 0 ? 1u : 0u;
        dsum += (uint32_t)cd[k];
    }
    uint64_t packed = (uint64_t(c) << 40) | uint64_t(dsum);
    uint64_t total;
    uint64_t excl = dev::block_exclusive_add64<BLOCK>(packed, s_scan, total);
    if (total == 0) return;  // uniform across the block
    if (threadIdx.x == 0) {
        uint64_t add = ((total >> 40) << 32) | (total & kLow40);
        *s_base = dev::atomic_add_dev(cnt_next, add);
    }
    __syncthreads();
    uint64_t base = *s_base;
    __syncthreads();  // s_base is reused by the next call
    if (c == 0) return;
    uint32_t slot = cnt_entries(base) + uint32_t(excl >> 40);
    int64_t o = (int64_t)cnt_edges(base) + (int64_t)(excl & kLow40);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        int32_t d = cd[k];
        if (d > 0) {
            if (STORE_V) nxt.v[slot] = cv[k];
            nxt.start[slot] = crs[k];
            nxt.off[slot] = (int32_t)o;
            // tiles whose first edge lies in [o, o + d)
            for (int64_t t = (o + TILE - 1) / TILE; t * TILE < o + d; ++t)
                nxt.tfirst[t] = (int32_t)slot;
            ++slot;
            o += d;
        }
    }
}

// Generic load-balanced expansion over one frontier.  Op supplies:
//   __device__ bool edge(int32_t u, int32_t e, int32_t v)
//       per expanded edge (u -> v, e = CSR position); returns true if v is
//       claimed for the next frontier (each v at most once per level).
// When PRODUCE is false nothing is appended.
// The tile loop of the expansion, for kernels that add their own prologue.
// F / E: entries and expanded edges of `cur` (F > 0).
// LDS of one expansion: the tile's entries and the append scan.  A kernel
// that also does other work can overlay it with its own LDS in a union.
template <int BLOCK, int TILE, bool STORE_V_IN>
struct ExpandLds {
    int32_t off[TILE + 2];
    int32_t start[TILE + 2];
    int32_t v[STORE_V_IN ? TILE + 2 : 1];
    uint64_t scan[BLOCK / 64 + 1];
    uint64_t base;
};

// An Op may declare `static constexpr bool kStreamColumns = true`: the column
// indices are then read with non-temporal loads, so the streamed CSR does not
// push the Op's randomly probed per-vertex data out of the L2.
template <class Op, class = void>
struct streams_columns : std::false_type {};
template <class Op>
struct streams_columns<Op, std::void_t<decltype(Op::kStreamColumns)>>
    : std::bool_constant<Op::kStreamColumns> {};

// An Op may supply `__device__ int32_t row_end(int32_t v) const`: the end of
// the part of v's row that an appended entry covers (a prefix of the row,
// e.g. its light entries); without it the whole row, rowptr[v + 1].
template <class Op, class = void>
struct has_row_end : std::false_type {};
template <class Op>
struct has_row_end<Op, std::void_t<decltype(std::declval<const Op &>().row_end(0))>>
    : std::true_type {};
template <class Op>
__device__ __forceinline__ int32_t row_end_of(const Op &op, const int32_t *rowptr, int32_t v) {
    if constexpr (has_row_end<Op>::value) return op.row_end(v);
    else return rowptr[v + 1];
}

// prefetch loads the row extent of every edge's endpoint together with the
// column instead of after a successful claim: one dependent round trip less
// per level for 8 extra bytes per edge, worth it on small (latency-bound)
// frontiers only.  (A runtime flag: a second instantiation would duplicate
// the LDS.)
template <int BLOCK, int TILE, bool STORE_V_IN, bool STORE_V_OUT, bool PRODUCE, class Op>
__device__ __forceinline__ void expand_tiles(ExpandLds<BLOCK, TILE, STORE_V_IN> &L, Op &op,
                                             const int32_t *__restrict__ rowptr,
                                             const int32_t *__restrict__ col, int32_t nnz,
                                             FrontierBufs cur, int32_t F, int32_t E,
                                             FrontierBufs nxt, uint64_t *cnt_next,
                                             bool prefetch = false) {
    constexpr int K = TILE / BLOCK;
    int32_t *s_off = L.off, *s_start = L.start, *s_v = L.v;
    const int32_t ntiles = (int32_t)((E + TILE - 1) / TILE);
    for (int32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // entries owning this tile; every degree is >= 1, so at most TILE + 1.
        // The clamps only guard memory if the frontier were inconsistent.  A
        // frontier of one tile owns it whole (no partition lookup).
        const int32_t i0 = ntiles == 1 ? 0 : min(max(cur.tfirst[t], 0), F - 1);
        const int32_t i1 = (t + 1 < ntiles) ? min(max(cur.tfirst[t + 1], i0), F - 1) : F - 1;
        const int32_t nv = min(i1 - i0 + 1, TILE + 1);
        for (int32_t j = threadIdx.x; j < nv; j += BLOCK) {
            s_off[j] = cur.off[i0 + j];
            s_start[j] = cur.start[i0 + j];
            if (STORE_V_IN) s_v[j] = cur.v[i0 + j];
        }
        __syncthreads();
        const int32_t e0 = t * TILE;
        int32_t cv[K], crs[K], cd[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            cv[k] = 0;
            crs[k] = 0;
            cd[k] = 0;
            const int32_t e = e0 + threadIdx.x + k * BLOCK;
            if (e < E) {
                const int j = dev::upper_bound_m1(s_off, 0, nv, e);
                const int32_t idx = s_start[j] + (e - s_off[j]);
                if ((uint32_t)idx >= (uint32_t)nnz) continue;
                int32_t v;
                if constexpr (streams_columns<Op>::value) v = __builtin_nontemporal_load(col + idx);
                else v = col[idx];
                const int32_t u = STORE_V_IN ? s_v[j] : -1;
                int32_t prs = 0, pre = 0;
                if (PRODUCE && prefetch) {
                    prs = rowptr[v];
                    pre = row_end_of(op, rowptr, v);
                }
                if (op.edge(u, idx, v) && PRODUCE) {
                    if (!prefetch) {
                        prs = rowptr[v];
                        pre = row_end_of(op, rowptr, v);
                    }
                    cv[k] = v;
                    crs[k] = prs;
                    cd[k] = pre - prs;
                }
            }
        }
        if (PRODUCE)
            produce<BLOCK, K, TILE, STORE_V_OUT>(cv, crs, cd, nxt, cnt_next, L.scan, &L.base);
        __syncthreads();  // s_off / s_start are overwritten by the next tile
    }
}

template <int BLOCK, int TILE, bool STORE_V_IN, bool STORE_V_OUT, bool PRODUCE, class Op>
__device__ __forceinline__ void expand_tiles(Op &op, const int32_t *__restrict__ rowptr,
                                             const int32_t *__restrict__ col, int32_t nnz,
                                             FrontierBufs cur, int32_t F, int32_t E,
                                             FrontierBufs nxt, uint64_t *cnt_next,
                                             bool prefetch = false) {
    __shared__ ExpandLds<BLOCK, TILE, STORE_V_IN> L;
    expand_tiles<BLOCK, TILE, STORE_V_IN, STORE_V_OUT, PRODUCE>(L, op, rowptr, col, nnz, cur, F,
                                                                 E, nxt, cnt_next, prefetch);
}

template <int BLOCK, int TILE, bool STORE_V_IN, bool STORE_V_OUT, bool PRODUCE, class Op>
__global__ __launch_bounds__(BLOCK) void k_expand(Op op, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ col, int32_t nnz,
                                                  FrontierBufs cur, const uint64_t *cnt_cur,
                                                  FrontierBufs nxt, uint64_t *cnt_next,
                                                  uint64_t *cnt_clear, uint64_t *stats) {
    const uint64_t cc = *cnt_cur;
    const int32_t F = (int32_t)cnt_entries(cc);
    const int32_t E = (int32_t)cnt_edges(cc);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (cnt_clear) *cnt_clear = 0;
        if (stats && F > 0) {
            stats[0] += (uint64_t)E;  // edges traversed
            stats[1] += (uint64_t)F;  // entries expanded
            stats[2] += 1;            // levels with work
        }
    }
    if (F == 0) return;
    expand_tiles<BLOCK, TILE, STORE_V_IN, STORE_V_OUT, PRODUCE>(op, rowptr, col, nnz, cur, F, E,
                                                                 nxt, cnt_next);
}

// Build a frontier from a vertex predicate over [0, n) or from a vertex
// list.  Op supplies  __device__ bool keep(int32_t v).
template <int BLOCK, int K, int TILE, bool STORE_V, class Op>
__global__ __launch_bounds__(BLOCK) void k_build(Op op, const int32_t *__restrict__ rowptr,
                                                 const int32_t *list, int32_t n,
                                                 const uint64_t *n_from_counter,
                                                 FrontierBufs nxt, uint64_t *cnt_next) {
    __shared__ uint64_t s_scan[BLOCK / 64 + 1];
    __shared__ uint64_t s_base;
    if (n_from_counter) n = (int32_t)cnt_entries(*n_from_counter);
    const int64_t per_block = (int64_t)BLOCK * K;
    for (int64_t b0 = (int64_t)blockIdx.x * per_block; b0 < n; b0 += (int64_t)gridDim.x * per_block) {
        int32_t cv[K], crs[K], cd[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            cv[k] = 0;
            crs[k] = 0;
            cd[k] = 0;
            const int64_t i = b0 + threadIdx.x + (int64_t)k * BLOCK;
            if (i < n) {
                const int32_t v = list ? list[i] : (int32_t)i;
                if (op.keep(v)) {
                    const int32_t rs = rowptr[v];
                    cv[k] = v;
                    crs[k] = rs;
                    cd[k] = rowptr[v + 1] - rs;
                }
            }
        }
        produce<BLOCK, K, TILE, STORE_V>(cv, crs, cd, nxt, cnt_next, s_scan, &s_base);
    }
}

// Control block: 16 words of counters / statistics, then 16 accumulators
// 128 bytes apart, which blocks pick by block index so that a per-block
// statistic does not serialise thousands of atomics on one address.
constexpr int kCtlWords = 16 + 16 * 16;
constexpr int kSpread = 16;
__device__ __forceinline__ uint64_t *spread_slot(uint64_t *ctl) {
    return ctl + 16 + (blockIdx.x % kSpread) * 16;
}
inline uint64_t spread_sum(const uint64_t *h) {
    uint64_t t = 0;
    for (int i = 0; i < kSpread; ++i) t += h[16 + i * 16];
    return t;
}

// Near-far split bookkeeping in the spread slots: the split of round r adds
// the count of pending entries it leaves to slot r % 3 and lowers slot
// 3 + r % 3 to the smallest of their distances (non-negative float bits;
// ~0 = none).  Round r reads the slots of round r - 1 and resets those of
// round r + 1, so no round reads a word it writes.
__host__ __device__ inline uint64_t nf_left(const uint64_t *ctl, int r) {
    uint64_t t = 0;
    for (int i = 0; i < kSpread; ++i) t += ctl[16 + i * 16 + r % 3];
    return t;
}
__device__ __forceinline__ uint64_t nf_min_left(const uint64_t *ctl, int r) {
    uint64_t t = ~0ull;
    for (int i = 0; i < kSpread; ++i) t = min(t, ctl[16 + i * 16 + 3 + r % 3]);
    return t;
}
__device__ __forceinline__ void nf_reset(uint64_t *ctl, int r) {
    for (int i = 0; i < kSpread; ++i) {
        ctl[16 + i * 16 + r % 3] = 0;
        ctl[16 + i * 16 + 3 + r % 3] = ~0ull;
    }
}
// one block's share of a split's result (called by one thread)
__device__ __forceinline__ void nf_report(uint64_t *ctl, int r, uint64_t left, uint32_t min_bits) {
    uint64_t *row = ctl + 16 + (blockIdx.x % kSpread) * 16;
    if (left) {
        dev::atomic_add_dev(row + r % 3, left);
        __hip_atomic_fetch_min(row + 3 + r % 3, (uint64_t)min_bits, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}
// The step of a split: the width delta at first, doubled by every split that
// follows a split that expanded nothing (the bucket [T, T + step) was empty:
// the distances are sparse against the step), capped at 2^100.  ctl[10 +
// (r & 1)] holds the step of round r (float bits), like T.
__device__ __forceinline__ float nf_next_step(float step, bool jump) {
    return jump ? fminf(step * 2.0f, 0x1p100f) : step;
}
// The threshold a split round raises T (float bits, non-negative) to: T +
// step, at least the next float above T, and -- when the previous round was
// a split that expanded nothing -- above the smallest distance that split
// left pending, so that this split moves it.
__device__ __forceinline__ int32_t nf_next_threshold(int32_t tbits, float step, bool jump,
                                                     uint64_t min_left) {
    int32_t t = dev::fbits(dev::bitsf(tbits) + step);
    if (t <= tbits) t = tbits + 1;
    if (jump && min_left != ~0ull && (int64_t)min_left + 1 > (int64_t)t)
        t = (int32_t)min_left + 1;
    return t;
}

// Frontier storage sized for a graph: two ping-pong frontiers, a ring of
// three counters and a stats block.
struct FrontierSet {
    FrontierBufs f[2];
    uint64_t *ctl;  // [0..2] counters, [4..15] stats, [16..271] spread accumulators
    int64_t cap_entries = 0, cap_tiles = 0;
};

inline FrontierSet frontier_alloc(Ctx *c, const std::string &tag, int32_t m, int64_t nnz,
                                  int tile, bool with_v) {
    FrontierSet s;
    const int64_t ent = std::max<int64_t>(m, 1) + 1;
    const int64_t tiles = nnz / tile + 2;
    for (int i = 0; i < 2; ++i) {
        std::string p = tag + std::to_string(i);
        s.f[i].v = with_v ? (int32_t *)scratch(c, p + "_v", sizeof(int32_t) * ent) : nullptr;
        s.f[i].start = (int32_t *)scratch(c, p + "_s", sizeof(int32_t) * ent);
        s.f[i].off = (int32_t *)scratch(c, p + "_o", sizeof(int32_t) * ent);
        s.f[i].tfirst = (int32_t *)scratch(c, p + "_t", sizeof(int32_t) * tiles);
    }
    s.ctl = (uint64_t *)scratch(c, tag + "_ctl", sizeof(uint64_t) * kCtlWords);
    s.cap_entries = ent;
    s.cap_tiles = tiles;
    return s;
}

}  // namespace apspie
