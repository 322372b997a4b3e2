// Triangle counting as the masked sparse product L . L o L.
//
// Reference: the TC driver builds the strictly-lower and strictly-upper
// operands (buildLower, test/tc/main.cu:39-55), multiplies them with
// cuSPARSE (spgemm(), tc.cuh:62-165), intersects the product with the mask
// (ewiseMultTc, tc.cuh:168-199) and reduces the per-column counts in float
// (test/tc/main.cu:249).  Here the product is never materialised: for every
// row r of L (the columns of row r below r, a prefix of the sorted row)
//     per_row[r] = sum_{k in L_r} |L_r cap L_k|
// = triangles whose largest vertex is r; each intersection is evaluated
// from its shorter side (see the pair split below the table helpers).
// Rows are binned by |L_r|:
//   * small rows: one wave per row, L_r in a 256-slot LDS hash table;
//   * medium rows: one block per row, 4096-slot LDS table;
//   * large rows: one block per row, table in global memory (spill).
// The wedges (k, j), j in L_k, of a row are enumerated load-balanced: an
// exclusive scan of |L_k| over k in L_r (the symbolic row count of L.L) maps
// each lane to its wedge; each wedge is one probe of the hash table.  Counts
// are integers: per_row int32, the total an exact uint64.
#include <algorithm>

#include "common.hpp"
#include "device.hpp"
#include "timing.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;
constexpr int kSmallRow = 128;     // wave bin: |L_r| <= 128, table 256
constexpr int kSmallTable = 256;
constexpr int kMediumRow = 2048;   // block bin capacity: |L_r| <= 2048 (LDS table 4096); the
                                   // bound used is tc_medium_row, 512 by default
constexpr int kMediumTable = 4096;
constexpr int kMediumChunk = 1024;  // staging of the medium tier (LDS: 5 blocks per CU)
constexpr int kBigRow = 16384;     // 1024-thread block, LDS table 32768 (128 KB)
constexpr int kBigTable = 32768;
constexpr int kBigBlock = 1024;    // also for the longest rows (table in global memory)
constexpr int kBigChunk = 1024;    // entries per enumeration chunk in the big tiers
constexpr int kRun = 8;            // consecutive wedges per lane (one binary search)
constexpr int kBlockRun = 8;       // the same in the block tiers (16 measured no faster)

// lowcnt[r] = number of entries of row r below the diagonal (rows sorted)
__global__ void k_lowcnt(const int32_t *rowptr, const int32_t *col, int32_t m, int32_t *lowcnt) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = rowptr[r], hi = rowptr[r + 1];
        const int32_t b = lo;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (col[mid] < (int32_t)r) lo = mid + 1; else hi = mid;
        }
        lowcnt[r] = lo - b;
    }
}

// rows sorted and duplicate-free?  One thread per entry: entry e must exceed
// entry e - 1 unless e starts a row (row starts marked in a bitmap first, so
// a hub row is checked by many threads, not walked by one).
__global__ void k_row_starts(const int32_t *rowptr, int32_t m, int64_t nnz, uint32_t *bits) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = rowptr[r];
        if (e < nnz) atomicOr(bits + (e >> 5), 1u << (e & 31));
    }
}

__global__ void k_check_sorted(const int32_t *col, int64_t nnz, const uint32_t *bits,
                               int32_t *flag) {
    for (int64_t e = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x)
        if (!((bits[e >> 5] >> (e & 31)) & 1u) && col[e - 1] >= col[e]) flag[0] = 1;
}

// bins by |L_r|: [0] small, [1] medium, [2] big, [3] huge (lists[b] has m
// slots each); ctl[0..3] counts, ctl[4] max |L_r|
__global__ void k_tc_bin(const int32_t *lowcnt, int32_t m, int32_t medium_row, int32_t big_row,
                         int32_t *lists, uint32_t *ctl, int32_t *per_row) {
    for (int64_t r0 = blockIdx.x * (int64_t)blockDim.x; r0 < m; r0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = r0 + threadIdx.x;
        int32_t n = r < m ? lowcnt[r] : 0;
        const int bin = n < 2 ? -1
                              : (n <= kSmallRow ? 0
                                                : (n <= medium_row ? 1 : (n <= big_row ? 2 : 3)));
        if (r < m && per_row) per_row[r] = 0;  // rows add into each other's counts
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint64_t mask = __ballot(bin == b);
            if (!mask) continue;
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if (dev::lane_id() == leader) base = dev::atomic_add_dev(ctl + b, (uint32_t)__popcll(mask));
            base = __shfl(base, leader, 64);
            if (bin == b) {
                const uint32_t slot = base + __popcll(mask & dev::lanemask_lt());
                lists[(int64_t)b * m + slot] = (int32_t)r;
            }
        }
        if (bin == 3) atomicMax(ctl + 4, (uint32_t)n);
    }
}

__device__ __forceinline__ void table_insert(int32_t *table, uint32_t mask, int32_t key) {
    uint32_t h = dev::fib_slot((uint32_t)key, mask);
    for (;;) {
        const int32_t prev = atomicCAS(table + h, -1, key);
        if (prev == -1 || prev == key) return;
        h = (h + 1) & mask;
    }
}

// A bit filter beside each table: bit h of the key's multiplicative hash
// (its top BITS bits) is set on insert.  Most probes miss (85 % at RMAT-22),
// and a clear bit settles a miss with one LDS read instead of a probe chain.
template <int BITS>
__device__ __forceinline__ void bloom_insert(uint32_t *bloom, int32_t key) {
    const uint32_t h = ((uint32_t)key * 2654435769u) >> (32 - BITS);
    atomicOr(bloom + (h >> 5), 1u << (h & 31));
}

template <int BITS>
__device__ __forceinline__ bool bloom_may_have(const uint32_t *bloom, int32_t key) {
    const uint32_t h = ((uint32_t)key * 2654435769u) >> (32 - BITS);
    return (bloom[h >> 5] >> (h & 31)) & 1u;
}

__device__ __forceinline__ bool table_has(const int32_t *table, uint32_t mask, int32_t key) {
    uint32_t h = dev::fib_slot((uint32_t)key, mask);
    for (;;) {
        const int32_t t = table[h];
        if (t == key) return true;
        if (t == -1) return false;
        h = (h + 1) & mask;
    }
}

// Pairs (i, k), k in L_i, are split by size so that an intersection
// |L_i cap L_k| with a long L_k is counted by probing the shorter list into a
// table of the longer one (the work drops from sum |L_k| towards
// sum min(|L_i|, |L_k|)):
//   |L_k| <= max(|L_i|, kSmallRow): "own" pair of row i -- L_k is scanned
//     against the table of L_i;
//   otherwise: "upper" pair of row k (a block-bin row) -- row k holds L_k in
//     its table and scans L_i for every i in U_k (the part of row k above the
//     diagonal) with |L_i| < |L_k|; the count goes to row i.
// Small rows never take upper pairs (a low-id hub has a short L and a huge U).
// Both enumerations of a row share its table and are load-balanced the same
// way (exclusive scan of the scanned lengths, one lane per wedge).

// one wave per small row
__global__ __launch_bounds__(kBlock) void k_tc_small(const int32_t *rowptr, const int32_t *col,
                                                     int32_t nnz, const int32_t *lowcnt,
                                                     const int32_t *rows, const uint32_t *cnt,
                                                     int32_t *per_row, uint64_t *total,
                                                     uint64_t *wedges) {
    constexpr int NW = kBlock / 64;
    constexpr int kSmallBloom = 11;  // 2048 bits per wave
    __shared__ int32_t s_table[NW][kSmallTable];
    __shared__ uint32_t s_bloom[NW][(1 << kSmallBloom) / 32];
    __shared__ int32_t s_pre[NW][kSmallRow + 1];
    __shared__ int32_t s_rs[NW][kSmallRow];
    const int w = threadIdx.x / 64, lane = dev::lane_id();
    const int32_t nrows = (int32_t)cnt[0];
    uint64_t my_tri = 0, my_wedges = 0;
    for (int32_t ri = blockIdx.x * NW + w; ri < nrows; ri += gridDim.x * NW) {
        const int32_t r = rows[ri];
        const int32_t n = lowcnt[r];
        const int32_t base = rowptr[r];
        for (int s = lane; s < kSmallTable; s += 64) s_table[w][s] = -1;
        for (int s = lane; s < (1 << kSmallBloom) / 32; s += 64) s_bloom[w][s] = 0;
        dev::wave_sync();
        // own pairs: every k in L_r with |L_k| <= kSmallRow
        uint32_t carry = 0;
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int idx = c0 + lane;
            uint32_t lc = 0;
            if (idx < n) {
                const int32_t k = col[base + idx];
                table_insert(s_table[w], kSmallTable - 1, k);
                bloom_insert<kSmallBloom>(s_bloom[w], k);
                s_rs[w][idx] = rowptr[k];
                const int32_t l = lowcnt[k];
                lc = l <= kSmallRow ? (uint32_t)l : 0u;  // n <= kSmallRow
            }
            const uint32_t inc = dev::wave_inclusive_add(lc);
            if (idx < n) s_pre[w][idx] = (int32_t)(carry + inc - lc);
            carry += dev::wave_read_lane(inc, 63);
        }
        if (lane == 0) s_pre[w][n] = (int32_t)carry;
        dev::wave_sync();
        const int32_t W = (int32_t)carry;
        uint32_t c = 0;
        // kRun consecutive wedges per lane, their column loads issued together
        for (int32_t x0 = lane * kRun; x0 < W; x0 += 64 * kRun) {
            int j = dev::upper_bound_m1(s_pre[w], 0, n, x0);
            int32_t key[kRun];
#pragma unroll
            for (int q = 0; q < kRun; ++q) {
                const int32_t x = x0 + q;
                key[q] = -1;
                if (x < W) {
                    while (s_pre[w][j + 1] <= x) ++j;
                    const int32_t e = s_rs[w][j] + (x - s_pre[w][j]);
                    if ((uint32_t)e < (uint32_t)nnz) key[q] = col[e];  // guard only
                }
            }
#pragma unroll
            for (int q = 0; q < kRun; ++q)
                if (key[q] >= 0 && bloom_may_have<kSmallBloom>(s_bloom[w], key[q]))
                    c += table_has(s_table[w], kSmallTable - 1, key[q]) ? 1u : 0u;
        }
        const uint64_t rc = dev::wave_sum_u64(c);
        if (lane == 0) {
            if (rc && per_row) atomicAdd(per_row + r, (int32_t)rc);
            my_tri += rc;
            my_wedges += (uint64_t)W;
        }
        dev::wave_sync();
    }
    const uint64_t t = dev::wave_sum_u64(my_tri), wd = dev::wave_sum_u64(my_wedges);
    if (lane == 0 && (t || wd)) {
        dev::atomic_add_dev(total, t);
        dev::atomic_add_dev(wedges, wd);
    }
}

// The wedges of one chunk of entries j in [0, cn): entry j owns the wedges
// [pre[j], pre[j + 1]) starting at CSR position rsk[j]; pre[cn] = W.  A lane
// takes kRun consecutive wedges, finds the first one's entry by binary search
// and walks on from there, issuing the kRun column loads back to back before
// probing; f(j, key) is called for every wedge with the column it reaches.
template <int BLOCK, int RUN, class F>
__device__ __forceinline__ void for_wedges(const int32_t *col, const int32_t *pre,
                                           const int32_t *rsk, int32_t cn, int32_t W,
                                           int32_t nnz, F &&f) {
    for (int64_t x0 = (int64_t)threadIdx.x * RUN; x0 < W; x0 += (int64_t)BLOCK * RUN) {
        int j = dev::upper_bound_m1(pre, 0, cn, (int32_t)x0);
        int32_t key[RUN], jj[RUN];
#pragma unroll
        for (int q = 0; q < RUN; ++q) {
            const int32_t x = (int32_t)x0 + q;
            key[q] = -1;
            jj[q] = j;
            if (x < W) {
                while (pre[j + 1] <= x) ++j;
                const int32_t e = rsk[j] + (x - pre[j]);
                jj[q] = j;
                if ((uint32_t)e < (uint32_t)nnz) key[q] = col[e];  // guard only
            }
        }
#pragma unroll
        for (int q = 0; q < RUN; ++q)
            if (key[q] >= 0) f(jj[q], key[q]);
    }
}

// Work items of the block bins: a row's own pairs in chunks of CHUNK entries
// of L_r, then its upper pairs in chunks of CHUNK entries of U_r; item =
// (row, chunk index), so the chunks of one long row spread over many blocks
// instead of queueing behind each other in one.  Rows of list a come first.
template <int CHUNK>
__global__ void k_tc_items(const int32_t *rowptr, const int32_t *lowcnt, const int32_t *rows_a,
                           const uint32_t *cnt_a, const int32_t *rows_b, const uint32_t *cnt_b,
                           int2 *items, uint32_t *nitems) {
    const int32_t na = (int32_t)cnt_a[0], nr = na + (cnt_b ? (int32_t)cnt_b[0] : 0);
    const int lane = dev::lane_id();
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < nr;
         i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        uint32_t k = 0;
        int32_t r = 0;
        if (i < nr) {
            r = i < na ? rows_a[i] : rows_b[i - na];
            const int32_t n = lowcnt[r], u = rowptr[r + 1] - rowptr[r] - n;
            k = (uint32_t)((n + CHUNK - 1) / CHUNK + (u + CHUNK - 1) / CHUNK);
        }
        const uint32_t incl = dev::wave_inclusive_add(k);
        uint32_t base = 0;
        if (lane == 63 && incl) base = dev::atomic_add_dev(nitems, incl);
        base = dev::wave_read_lane(base, 63) + incl - k;
        for (uint32_t q = 0; q < k; ++q) items[base + q] = make_int2(r, (int32_t)q);
    }
}

// One block per work item of the block bins (see k_tc_items); consecutive
// items of one row in a block keep the row's table.  A row with
// |L_r| <= TABLE / 2 keeps its table in LDS; a longer one in the block's
// region of global memory (g_cap slots).  The scanned lengths are staged in
// LDS, CHUNK entries at a time, for both pair kinds.
template <int TABLE, int BLOCK, int CHUNK>
__global__ __launch_bounds__(BLOCK) void k_tc_block(const int32_t *rowptr, const int32_t *col,
                                                    int32_t nnz, const int32_t *lowcnt,
                                                    const int2 *items, const uint32_t *nitems,
                                                    int32_t *per_row, uint64_t *total,
                                                    uint64_t *wedges, int32_t *g_table,
                                                    int64_t g_cap) {
    constexpr int BLOOM = TABLE <= 4096 ? 14 : 16;  // 2 KB / 8 KB of filter bits
    __shared__ int32_t s_table[TABLE];
    __shared__ uint32_t s_bloom[(1 << BLOOM) / 32];
    __shared__ int32_t pre[CHUNK + 1];
    __shared__ int32_t rsk[CHUNK];
    __shared__ int32_t cntu[CHUNK];
    __shared__ uint32_t s_scan[BLOCK / 64 + 1];
    __shared__ uint64_t s_red[BLOCK / 64];
    const uint32_t nit = *nitems;
    uint64_t my_tri = 0, my_wedges = 0;
    int32_t table_row = -1;
    for (uint32_t it = blockIdx.x; it < nit; it += gridDim.x) {
        const int2 item = items[it];
        const int32_t r = item.x;
        const int32_t n = lowcnt[r];
        const int32_t base = rowptr[r], rend = rowptr[r + 1];
        const int32_t own_chunks = (n + CHUNK - 1) / CHUNK;
        const bool lds = 2 * n <= TABLE;
        int32_t *table = lds ? s_table : g_table + (int64_t)blockIdx.x * g_cap;
        uint32_t tsize = TABLE;
        if (!lds) {
            tsize = 1;
            while (tsize < 2u * (uint32_t)n) tsize <<= 1;
        }
        const uint32_t mask = tsize - 1;
        if (r != table_row) {  // uniform: every thread read the same item
            for (uint32_t s = threadIdx.x; s < tsize; s += BLOCK) table[s] = -1;
            for (int s = threadIdx.x; s < (1 << BLOOM) / 32; s += BLOCK) s_bloom[s] = 0;
            __syncthreads();
            for (int32_t idx = threadIdx.x; idx < n; idx += BLOCK) {
                const int32_t k = col[base + idx];
                table_insert(table, mask, k);
                bloom_insert<BLOOM>(s_bloom, k);
            }
            __syncthreads();
            table_row = r;
        }
        if (item.y < own_chunks) {
            // own pairs: k in L_r with |L_k| <= |L_r|, entries [k0, k0 + CHUNK) of L_r
            uint32_t c = 0;
            const int32_t k0 = item.y * CHUNK;
            const int32_t cn = min(CHUNK, n - k0);
            uint32_t carry = 0;
            for (int c0 = 0; c0 < cn; c0 += BLOCK) {
                const int idx = c0 + threadIdx.x;
                uint32_t lc = 0;
                if (idx < cn) {
                    const int32_t k = col[base + k0 + idx];
                    rsk[idx] = rowptr[k];
                    const int32_t l = lowcnt[k];
                    lc = l <= n ? (uint32_t)l : 0u;
                }
                uint32_t tot;
                const uint32_t ex = dev::block_exclusive_add32<BLOCK>(lc, s_scan, tot);
                if (idx < cn) pre[idx] = (int32_t)(carry + ex);
                carry += tot;
            }
            if (threadIdx.x == 0) pre[cn] = (int32_t)carry;
            __syncthreads();
            for_wedges<BLOCK, kBlockRun>(col, pre, rsk, cn, (int32_t)carry, nnz, [&](int, int32_t key) {
                c += (bloom_may_have<BLOOM>(s_bloom, key) && table_has(table, mask, key)) ? 1u
                                                                                       : 0u;
            });
            if (threadIdx.x == 0) my_wedges += carry;
            const uint64_t wc = dev::wave_sum_u64(c);
            if (dev::lane_id() == 0) s_red[threadIdx.x / 64] = wc;
            __syncthreads();
            if (threadIdx.x == 0) {
                uint64_t rc = 0;
                for (int q = 0; q < BLOCK / 64; ++q) rc += s_red[q];
                if (rc && per_row) atomicAdd(per_row + r, (int32_t)rc);
                my_tri += rc;
            }
            __syncthreads();  // pre / rsk / s_red are refilled by the next item
        } else {
            // upper pairs: i in U_r with |L_i| < |L_r|, entries [u0, u0 + CHUNK) of U_r
            const int32_t u0 = base + n + (item.y - own_chunks) * CHUNK;
            const int32_t cn = min(CHUNK, rend - u0);
            uint32_t carry = 0;
            for (int c0 = 0; c0 < cn; c0 += BLOCK) {
                const int idx = c0 + threadIdx.x;
                uint32_t li = 0;
                if (idx < cn) {
                    const int32_t i = col[u0 + idx];
                    rsk[idx] = rowptr[i];
                    cntu[idx] = 0;
                    const int32_t l = lowcnt[i];
                    li = l < n ? (uint32_t)l : 0u;
                }
                uint32_t tot;
                const uint32_t ex = dev::block_exclusive_add32<BLOCK>(li, s_scan, tot);
                if (idx < cn) pre[idx] = (int32_t)(carry + ex);
                carry += tot;
            }
            if (threadIdx.x == 0) pre[cn] = (int32_t)carry;
            __syncthreads();
            if (carry > 0) {  // uniform: most chunks of a hub's upper part scan nothing
                for_wedges<BLOCK, kBlockRun>(col, pre, rsk, cn, (int32_t)carry, nnz, [&](int j, int32_t key) {
                    if (bloom_may_have<BLOOM>(s_bloom, key) && table_has(table, mask, key))
                        atomicAdd(cntu + j, 1);
                });
                __syncthreads();
                uint64_t cu = 0;
                for (int idx = threadIdx.x; idx < cn; idx += BLOCK) {
                    const int32_t ci = cntu[idx];
                    if (ci) {
                        if (per_row) atomicAdd(per_row + col[u0 + idx], ci);
                        cu += (uint64_t)ci;
                    }
                }
                const uint64_t wcu = dev::wave_sum_u64(cu);
                if (dev::lane_id() == 0) s_red[threadIdx.x / 64] = wcu;
                __syncthreads();
                if (threadIdx.x == 0) {
                    for (int q = 0; q < BLOCK / 64; ++q) my_tri += s_red[q];
                    my_wedges += carry;
                }
            }
            __syncthreads();  // rsk / pre / cntu are refilled by the next chunk
        }
    }
    if (threadIdx.x == 0 && (my_tri || my_wedges)) {
        dev::atomic_add_dev(total, my_tri);
        dev::atomic_add_dev(wedges, my_wedges);
    }
}

}  // namespace

// Per-row count of entries below the diagonal, cached with the graph (built
// once, under the graph's lock, by whichever context needs it first).
void graph_prepare_tc(Ctx *c, Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    if (g->lowcnt || g->m == 0) return;
    const int32_t m = g->m;
    const int grid = fill_grid(c, kBlock, 4);
    uint32_t *ctl = (uint32_t *)scratch(c, "tc_prep", 64);
    fill_i32(c, (int32_t *)ctl, 0, 16);
    {
        uint32_t *bits = (uint32_t *)scratch(c, "tc_rowstarts", sizeof(uint32_t) * (g->nnz / 32 + 1));
        fill_i32(c, (int32_t *)bits, 0, g->nnz / 32 + 1);
        k_row_starts<<<grid, kBlock, 0, c->stream>>>(g->rowptr, m, g->nnz, bits);
        k_check_sorted<<<grid, kBlock, 0, c->stream>>>(g->col, g->nnz, bits, (int32_t *)ctl + 4);
        uint32_t flag = 0;
        APSPIE_HIP(hipMemcpyAsync(&flag, ctl + 4, 4, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        APSPIE_CHECK(flag == 0, APSPIE_ERR_INVALID, "TC needs sorted, duplicate-free rows");
        // owned by the graph (freed with it), not charged to this context
        int32_t *lc = nullptr;
        APSPIE_HIP(hipMalloc(&lc, sizeof(int32_t) * std::max(m, 1)));
        k_lowcnt<<<grid, kBlock, 0, c->stream>>>(g->rowptr, g->col, m, lc);
        APSPIE_LAUNCH_CHECK();
        sync(c);
        g->lowcnt = lc;
    }
}

void tc_run(Ctx *c, const Graph *g0, uint64_t *total_out, int32_t *per_row) {
    Graph *g = const_cast<Graph *>(g0);
    *total_out = 0;
    const int32_t m = g->m;
    if (m == 0) return;
    graph_prepare_tc(c, g);
    const int grid = fill_grid(c, kBlock, 4);
    uint32_t *ctl = (uint32_t *)scratch(c, "tc_ctl", 64);
    uint64_t *acc = (uint64_t *)(ctl + 8);  // [0] triangles, [1] wedges
    fill_i32(c, (int32_t *)ctl, 0, 16);
    int32_t *per = per_row;  // may be null: total only, no per-row atomics
    int32_t *lists = (int32_t *)scratch(c, "tc_lists", sizeof(int32_t) * 4 * (size_t)m);
    // the medium / big tier bound (tunable up to the medium tier's LDS table
    // capacity, kMediumRow): a row above it goes to the 1024-thread tier, whose
    // LDS table holds rows up to kBigRow (longer ones use global memory)
    const int32_t medium_row =
        (int32_t)std::min<int64_t>(std::max<int64_t>(opt(c, "tc_medium_row", 512), kSmallRow),
                                   kMediumRow);
    k_tc_bin<<<grid, kBlock, 0, c->stream>>>(g->lowcnt, m, medium_row, kBigRow, lists, ctl, per);
    APSPIE_LAUNCH_CHECK();
    uint32_t h[5];
    APSPIE_HIP(hipMemcpyAsync(h, ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    c->stats.host_syncs++;
    KernelTimer kt(c);
    kt.begin();
    // block-bin work items: ctl[5] counts those of the long rows (huge, then
    // big), ctl[6] those of the medium rows
    const int64_t nbig = (int64_t)h[2] + h[3];
    int2 *items_big = nullptr, *items_med = nullptr;
    if (nbig) {
        items_big = (int2 *)scratch(c, "tc_items_big",
                                    sizeof(int2) * (g->nnz / kBigChunk + 2 * nbig + 64));
        k_tc_items<kBigChunk><<<grid, kBlock, 0, c->stream>>>(
            g->rowptr, g->lowcnt, lists + 3 * (size_t)m, ctl + 3, lists + 2 * (size_t)m, ctl + 2,
            items_big, ctl + 5);
        APSPIE_LAUNCH_CHECK();
    }
    if (h[1]) {
        items_med = (int2 *)scratch(c, "tc_items_med",
                                    sizeof(int2) * (g->nnz / kMediumChunk + 2 * (int64_t)h[1] + 64));
        k_tc_items<kMediumChunk><<<grid, kBlock, 0, c->stream>>>(
            g->rowptr, g->lowcnt, lists + m, ctl + 1, nullptr, nullptr, items_med, ctl + 6);
        APSPIE_LAUNCH_CHECK();
    }
    if (nbig) {  // longest rows first: they set the tail
        int64_t cap = 1;
        while (cap < 2 * (int64_t)h[4]) cap <<= 1;
        const int blocks = c->cus;
        int32_t *gt = h[3] ? (int32_t *)scratch(c, "tc_gtable", sizeof(int32_t) * cap * blocks)
                           : nullptr;
        k_tc_block<kBigTable, kBigBlock, kBigChunk><<<blocks, kBigBlock, 0, c->stream>>>(
            g->rowptr, g->col, (int32_t)g->nnz, g->lowcnt, items_big, ctl + 5, per, acc, acc + 1,
            gt, cap);
        APSPIE_LAUNCH_CHECK();
    }
    if (h[1]) {
        k_tc_block<kMediumTable, kBlock, kMediumChunk><<<5 * c->cus, kBlock, 0, c->stream>>>(
            g->rowptr, g->col, (int32_t)g->nnz, g->lowcnt, items_med, ctl + 6, per, acc, acc + 1,
            nullptr, 0);
        APSPIE_LAUNCH_CHECK();
    }
    if (h[0]) {
        const int blocks = (int)std::min<int64_t>(ceil_div(h[0], kBlock / 64), 8 * c->cus);
        k_tc_small<<<blocks, kBlock, 0, c->stream>>>(g->rowptr, g->col, (int32_t)g->nnz,
                                                     g->lowcnt, lists, ctl + 0, per, acc, acc + 1);
        APSPIE_LAUNCH_CHECK();
    }
    kt.end(1);
    uint64_t res[2];
    APSPIE_HIP(hipMemcpyAsync(res, acc, sizeof res, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    c->stats.host_syncs++;
    *total_out = res[0];
    c->stats.edges_traversed = (int64_t)res[1];  // wedges probed
    c->stats.vertices_reached = (int64_t)h[0] + h[1] + h[2] + h[3];
    c->stats.iterations = 1;
    kt.collect();
}

}  // namespace apspie
