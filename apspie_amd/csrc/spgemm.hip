// General sparse product C = A * B over (+, x) by row-wise accumulation.
//
// Reference: the hash SpGEMM spgemm() (spgemm.cuh:131-342) with its insert
// kernels (spgemmKernel.cuh:179-343): open-addressing tables keyed by output
// coordinates, atomic value accumulation, a sort to order the output.  The
// reference tiles C in 2D over global-memory tables and drops products after
// 25 probes; here each output row gets its own accumulator sized from the
// row, so nothing is ever dropped:
//   0. u[i] = min(sum_{k in A_i} |B_k|, n): the products of row i, capped by
//      the number of columns (no row has more distinct outputs)
//   1. rows binned by u: wave + 512-slot LDS table (u <= 256); block +
//      8192-slot LDS table (u <= 4096); longer rows with n <= 32768 a dense
//      LDS accumulator (value per column + a column bitmap, no hashing and no
//      sort); longer rows of wider matrices a table in global memory
//   2. symbolic pass: distinct columns per row
//   3. exclusive scan -> C row offsets
//   4. numeric pass: accumulate, then write the row in column order (the
//      dense accumulator is in order already; a block table on a matrix of
//      at most 32768 columns writes each entry at its rank among the row's
//      columns, from a column bitmap; other tables compact their used slots
//      and sort only those)
// The block tiers walk a row's products load-balanced: an exclusive scan of
// |B_k| over a chunk of A_i's entries, then each lane takes 8 consecutive
// products (one binary search).  Values are exact for integer-valued inputs;
// float (+) order otherwise follows the atomic order.
#include <algorithm>

#include "common.hpp"
#include "device.hpp"
#include "primitives.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;
constexpr int kBig = 1024;          // threads of the dense and global tiers
constexpr int kWaveUb = 256, kWaveTable = 512;
constexpr int kBlockUb = 4096, kBlockTable = 8192;
constexpr int kDenseN = 32768;      // dense accumulator: value + bit per column
constexpr int kChunk = 1024;        // entries of A_i staged per enumeration chunk
constexpr int kRun = 8;             // consecutive products per lane
constexpr int32_t kEmpty = 0x7fffffff;  // sorts after every column id

// u[i] = min(products of row i, n)
// one wave per row of A (the lanes stride its entries): a hub row is not
// left to one thread
__global__ void k_row_bound(const int32_t *arp, const int32_t *acol, const int32_t *brp,
                            int32_t m, int32_t n, int64_t *ub) {
    const int lane = dev::lane_id();
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; i < m;
         i += waves) {
        uint64_t s = 0;
        for (int32_t e = arp[i] + lane; e < arp[i + 1]; e += 64) {
            const int32_t k = acol[e];
            s += (uint64_t)(brp[k + 1] - brp[k]);
        }
        s = dev::wave_sum_u64(s);
        if (lane == 0) ub[i] = s == 0 ? 0 : std::min<int64_t>((int64_t)s, n);
    }
}

// bins: 0 wave, 1 block-LDS, 2 dense, 3 block-global; ctl[0..3] counts;
// ctl[4] max u of the global bin
__global__ void k_bin_rows(const int64_t *ub, int32_t m, int32_t wave_ub, int32_t block_ub,
                           bool dense_ok, int32_t *lists, uint32_t *ctl, int32_t *row_nnz) {
    for (int64_t r0 = blockIdx.x * (int64_t)blockDim.x; r0 < m; r0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = r0 + threadIdx.x;
        const int64_t u = r < m ? ub[r] : 0;
        const int bin = u == 0 ? -1
                        : u <= wave_ub ? 0
                        : u <= block_ub ? 1
                        : dense_ok ? 2 : 3;
        if (r < m && bin < 0) row_nnz[r] = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint64_t mask = __ballot(bin == b);
            if (!mask) continue;
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if (dev::lane_id() == leader) base = dev::atomic_add_dev(ctl + b, (uint32_t)__popcll(mask));
            base = __shfl(base, leader, 64);
            if (bin == b) lists[(int64_t)b * m + base + __popcll(mask & dev::lanemask_lt())] = (int32_t)r;
        }
        if (bin == 3) atomicMax(ctl + 4, (uint32_t)min<int64_t>(u, 0x7fffffff));
    }
}

// insert key; returns true if this call created the entry
__device__ __forceinline__ bool tab_insert(int32_t *keys, uint32_t mask, int32_t key,
                                           uint32_t *slot_out) {
    uint32_t h = dev::fib_slot((uint32_t)key, mask);
    for (;;) {
        const int32_t prev = atomicCAS(keys + h, kEmpty, key);
        if (prev == kEmpty || prev == key) {
            *slot_out = h;
            return prev == kEmpty;
        }
        h = (h + 1) & mask;
    }
}

// bitonic sort of (keys, vals) of power-of-two size `size` by `nthreads`
// cooperating threads (thread index tid); `barrier` orders the stages
template <class Barrier>
__device__ __forceinline__ void bitonic(int32_t *keys, float *vals, uint32_t size, int tid,
                                        int nthreads, Barrier barrier) {
    for (uint32_t k = 2; k <= size; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < size; i += nthreads) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const bool up = (i & k) == 0;
                    const int32_t a = keys[i], b = keys[p];
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[p] = a;
                        if (vals) {
                            const float t = vals[i];
                            vals[i] = vals[p];
                            vals[p] = t;
                        }
                    }
                }
            }
            barrier();
        }
    }
}

// Small rows: one wave per row, products walked entry by entry of A_i (the
// lanes stride over B_k), 512-slot LDS table sorted whole.
template <bool NUMERIC>
__device__ __forceinline__ void wave_row(int32_t r, const int32_t *arp, const int32_t *acol,
                                         const float *aval, const int32_t *brp,
                                         const int32_t *bcol, const float *bval, int32_t *keys,
                                         float *vals, int32_t *row_nnz, const int32_t *crp,
                                         int32_t *ccol, float *cval, uint32_t *s_count) {
    const int tid = dev::lane_id();
    constexpr uint32_t tsize = kWaveTable, mask = kWaveTable - 1;
    for (uint32_t s = tid; s < tsize; s += 64) {
        keys[s] = kEmpty;
        if (NUMERIC) vals[s] = 0.0f;
    }
    if (tid == 0) *s_count = 0;
    dev::wave_sync();
    uint32_t created = 0;
    for (int32_t e = arp[r]; e < arp[r + 1]; ++e) {
        const int32_t k = acol[e];
        const float av = aval ? aval[e] : 1.0f;
        for (int32_t f = brp[k] + tid; f < brp[k + 1]; f += 64) {
            uint32_t slot;
            created += tab_insert(keys, mask, bcol[f], &slot) ? 1u : 0u;
            if (NUMERIC) atomicAdd(vals + slot, av * (bval ? bval[f] : 1.0f));
        }
    }
    if (!NUMERIC) {
        if (created) atomicAdd(s_count, created);
        dev::wave_sync();
        if (tid == 0) row_nnz[r] = (int32_t)*s_count;
        dev::wave_sync();
        return;
    }
    dev::wave_sync();
    bitonic(keys, vals, tsize, tid, 64, [] { dev::wave_sync(); });
    const int32_t c0 = crp[r], nz = crp[r + 1] - c0;
    for (int32_t s = tid; s < nz; s += 64) {
        ccol[c0 + s] = keys[s];
        cval[c0 + s] = vals[s];
    }
    dev::wave_sync();
}

template <bool NUMERIC>
__global__ __launch_bounds__(kBlock) void k_rows_wave(const int32_t *rows, const uint32_t *cnt,
                                                      const int32_t *arp, const int32_t *acol,
                                                      const float *aval, const int32_t *brp,
                                                      const int32_t *bcol, const float *bval,
                                                      int32_t *row_nnz, const int32_t *crp,
                                                      int32_t *ccol, float *cval) {
    constexpr int NW = kBlock / 64;
    __shared__ int32_t s_keys[NW][kWaveTable];
    __shared__ float s_vals[NUMERIC ? NW : 1][NUMERIC ? kWaveTable : 1];
    __shared__ uint32_t s_count[NW];
    const int w = threadIdx.x / 64;
    const int32_t n = (int32_t)cnt[0];
    for (int32_t i = blockIdx.x * NW + w; i < n; i += gridDim.x * NW)
        wave_row<NUMERIC>(rows[i], arp, acol, aval, brp, bcol, bval, s_keys[w],
                          NUMERIC ? s_vals[w] : nullptr, row_nnz, crp, ccol, cval, &s_count[w]);
}

// The products of row r, load-balanced over the block: f(col, value) per
// product.  pre / rs / av are LDS staging of kChunk entries of A_r.
template <int BLOCK, class F>
__device__ __forceinline__ void for_products(int32_t r, const int32_t *arp, const int32_t *acol,
                                             const float *aval, const int32_t *brp,
                                             const int32_t *bcol, const float *bval,
                                             int32_t *pre, int32_t *rs, float *av,
                                             uint32_t *s_scan, F &&f) {
    const int32_t ab = arp[r], an = arp[r + 1] - ab;
    for (int32_t c0 = 0; c0 < an; c0 += kChunk) {
        const int32_t cn = min(kChunk, an - c0);
        uint32_t carry = 0;
        for (int32_t x0 = 0; x0 < cn; x0 += BLOCK) {
            const int32_t x = x0 + threadIdx.x;
            uint32_t len = 0;
            if (x < cn) {
                const int32_t k = acol[ab + c0 + x];
                const int32_t b0 = brp[k];
                len = (uint32_t)(brp[k + 1] - b0);
                rs[x] = b0;
                av[x] = aval ? aval[ab + c0 + x] : 1.0f;
            }
            uint32_t tot;
            const uint32_t ex = dev::block_exclusive_add32<BLOCK>(len, s_scan, tot);
            if (x < cn) pre[x] = (int32_t)(carry + ex);
            carry += tot;
        }
        if (threadIdx.x == 0) pre[cn] = (int32_t)carry;
        __syncthreads();
        const int64_t W = carry;
        // kRun consecutive products per lane: the loads of B go out together,
        // then the products are applied
        for (int64_t x0 = (int64_t)threadIdx.x * kRun; x0 < W; x0 += (int64_t)BLOCK * kRun) {
            int j = dev::upper_bound_m1(pre, 0, cn, (int32_t)x0);
            int32_t key[kRun];
            float val[kRun];
#pragma unroll
            for (int q = 0; q < kRun; ++q) {
                const int64_t x = x0 + q;
                key[q] = -1;
                val[q] = 0.0f;
                if (x < W) {
                    while (pre[j + 1] <= x) ++j;
                    const int64_t e = (int64_t)rs[j] + (x - pre[j]);
                    key[q] = bcol[e];
                    val[q] = av[j] * (bval ? bval[e] : 1.0f);
                }
            }
#pragma unroll
            for (int q = 0; q < kRun; ++q)
                if (key[q] >= 0) f(key[q], val[q]);
        }
        __syncthreads();  // pre / rs / av are refilled by the next chunk
    }
}

// Moves the used slots of a table to its front, in slot order, 256 slots at
// a time (a slot's destination never lies beyond the slots already read), and
// returns their number; slots [count, next power of two) are set empty.
template <int BLOCK>
__device__ __forceinline__ uint32_t compact_table(int32_t *keys, float *vals, uint32_t tsize,
                                                  uint32_t *s_scan, uint32_t *s_n) {
    if (threadIdx.x == 0) *s_n = 0;
    __syncthreads();
    for (uint32_t s0 = 0; s0 < tsize; s0 += BLOCK) {
        const uint32_t s = s0 + threadIdx.x;
        int32_t k = kEmpty;
        float v = 0.0f;
        if (s < tsize) {
            k = keys[s];
            if (vals) v = vals[s];
        }
        const uint32_t used = k != kEmpty ? 1u : 0u;
        uint32_t tot;
        const uint32_t ex = dev::block_exclusive_add32<BLOCK>(used, s_scan, tot);  // syncs
        const uint32_t base = *s_n;
        if (used) {
            keys[base + ex] = k;
            if (vals) vals[base + ex] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) *s_n = base + tot;
        __syncthreads();
    }
    const uint32_t n = *s_n;
    uint32_t p = 1;
    while (p < n) p <<= 1;
    for (uint32_t s = n + threadIdx.x; s < p; s += BLOCK) keys[s] = kEmpty;
    __syncthreads();
    return n;
}

// Longer rows: one block per row with an open-addressing table in LDS
// (TABLE > 0, 2u <= TABLE) or in the block's region of global memory.
template <bool NUMERIC, int BLOCK, int TABLE>
__global__ __launch_bounds__(BLOCK) void k_rows_table(const int32_t *rows, const uint32_t *cnt,
                                                      const int32_t *arp, const int32_t *acol,
                                                      const float *aval, const int32_t *brp,
                                                      const int32_t *bcol, const float *bval,
                                                      int32_t *row_nnz, const int32_t *crp,
                                                      int32_t *ccol, float *cval,
                                                      const int64_t *ub, int32_t *g_keys,
                                                      float *g_vals, int64_t g_cap,
                                                      int32_t ncols) {
    constexpr bool LDS = TABLE > 0;
    __shared__ int32_t s_keys[LDS ? TABLE : 1];
    __shared__ float s_vals[(LDS && NUMERIC) ? TABLE : 1];
    __shared__ int32_t pre[kChunk + 1], rs[kChunk];
    __shared__ float av[kChunk];
    __shared__ uint32_t s_scan[BLOCK / 64 + 1];
    __shared__ uint32_t s_count;
    int32_t *keys = LDS ? s_keys : g_keys + (int64_t)blockIdx.x * g_cap;
    float *vals = NUMERIC ? (LDS ? s_vals : g_vals + (int64_t)blockIdx.x * g_cap) : nullptr;
    const int32_t nrows = (int32_t)cnt[0];
    for (int32_t i = blockIdx.x; i < nrows; i += gridDim.x) {
        const int32_t r = rows[i];
        uint32_t tsize = TABLE;
        if (!LDS) {
            tsize = 1;
            while (tsize < 2 * (uint64_t)ub[r]) tsize <<= 1;
        }
        const uint32_t mask = tsize - 1;
        for (uint32_t s = threadIdx.x; s < tsize; s += BLOCK) {
            keys[s] = kEmpty;
            if (NUMERIC) vals[s] = 0.0f;
        }
        if (threadIdx.x == 0) s_count = 0;
        __syncthreads();
        uint32_t created = 0;
        for_products<BLOCK>(r, arp, acol, aval, brp, bcol, bval, pre, rs, av, s_scan,
                            [&](int32_t col, float v) {
                                uint32_t slot;
                                created += tab_insert(keys, mask, col, &slot) ? 1u : 0u;
                                if (NUMERIC) atomicAdd(vals + slot, v);
                            });
        if (!NUMERIC) {
            if (created) atomicAdd(&s_count, created);
            __syncthreads();
            if (threadIdx.x == 0) row_nnz[r] = (int32_t)s_count;
            __syncthreads();
            continue;
        }
        if (LDS && ncols <= 32 * kChunk) {
            // order by rank instead of sorting: a bit per column (the chunk
            // staging arrays, free once the products are done), the words'
            // popcount prefix, and each entry's rank from its own word
            uint32_t *bits = (uint32_t *)pre, *wpre = (uint32_t *)rs;
            const int32_t nw = (ncols + 31) / 32;
            __syncthreads();  // every thread is past the last chunk
            for (int32_t q = threadIdx.x; q < nw; q += BLOCK) bits[q] = 0;
            __syncthreads();
            for (uint32_t t = threadIdx.x; t < tsize; t += BLOCK) {
                const int32_t k = keys[t];
                if (k != kEmpty) atomicOr(bits + (k >> 5), 1u << (k & 31));
            }
            __syncthreads();
            uint32_t carry = 0;
            for (int32_t q0 = 0; q0 < nw; q0 += BLOCK) {
                const int32_t q = q0 + threadIdx.x;
                uint32_t tot;
                const uint32_t ex = dev::block_exclusive_add32<BLOCK>(
                    q < nw ? (uint32_t)__popc(bits[q]) : 0u, s_scan, tot);
                if (q < nw) wpre[q] = carry + ex;
                carry += tot;
            }
            __syncthreads();
            const int32_t c0 = crp[r];
            for (uint32_t t = threadIdx.x; t < tsize; t += BLOCK) {
                const int32_t k = keys[t];
                if (k == kEmpty) continue;
                const uint32_t below = bits[k >> 5] & ((1u << (k & 31)) - 1u);
                const int32_t pos = (int32_t)(wpre[k >> 5] + (uint32_t)__popc(below));
                ccol[c0 + pos] = k;
                cval[c0 + pos] = vals[t];
            }
            __syncthreads();
            continue;
        }
        const uint32_t nz = compact_table<BLOCK>(keys, vals, tsize, s_scan, &s_count);
        uint32_t p = 1;
        while (p < nz) p <<= 1;
        bitonic(keys, vals, p, threadIdx.x, BLOCK, [] { __syncthreads(); });
        const int32_t c0 = crp[r];
        for (uint32_t s = threadIdx.x; s < nz; s += BLOCK) {
            ccol[c0 + s] = keys[s];
            cval[c0 + s] = vals[s];
        }
        __syncthreads();
    }
}

// Rows of matrices with at most kDenseN columns: a dense LDS accumulator, one
// value and one bit per column; the row comes out in column order.
template <bool NUMERIC>
__global__ __launch_bounds__(kBig) void k_rows_dense(const int32_t *rows, const uint32_t *cnt,
                                                     int32_t ncols, const int32_t *arp,
                                                     const int32_t *acol, const float *aval,
                                                     const int32_t *brp, const int32_t *bcol,
                                                     const float *bval, int32_t *row_nnz,
                                                     const int32_t *crp, int32_t *ccol,
                                                     float *cval) {
    constexpr int NWORD = kDenseN / 32;
    __shared__ float s_val[NUMERIC ? kDenseN : 1];
    __shared__ uint32_t s_bits[NWORD];
    __shared__ int32_t pre[kChunk + 1], rs[kChunk];
    __shared__ float av[kChunk];
    __shared__ uint32_t s_scan[kBig / 64 + 1];
    const int32_t nw = (ncols + 31) / 32;
    const int32_t nrows = (int32_t)cnt[0];
    for (int32_t i = blockIdx.x; i < nrows; i += gridDim.x) {
        const int32_t r = rows[i];
        for (int32_t q = threadIdx.x; q < nw; q += kBig) s_bits[q] = 0;
        if (NUMERIC)
            for (int32_t q = threadIdx.x; q < ncols; q += kBig) s_val[q] = 0.0f;
        __syncthreads();
        for_products<kBig>(r, arp, acol, aval, brp, bcol, bval, pre, rs, av, s_scan,
                           [&](int32_t col, float v) {
                               atomicOr(s_bits + (col >> 5), 1u << (col & 31));
                               if (NUMERIC) atomicAdd(s_val + col, v);
                           });
        // prefix of the words' popcounts -> each word's first output slot
        uint32_t carry = 0;
        const int32_t c0 = NUMERIC ? crp[r] : 0;
        for (int32_t q0 = 0; q0 < nw; q0 += kBig) {
            const int32_t q = q0 + threadIdx.x;
            const uint32_t bits = q < nw ? s_bits[q] : 0u;
            uint32_t tot;
            const uint32_t ex = dev::block_exclusive_add32<kBig>((uint32_t)__popc(bits), s_scan, tot);
            if (NUMERIC && bits) {
                uint32_t b = bits, o = carry + ex;
                while (b) {
                    const int t = __ffs(b) - 1;
                    b &= b - 1;
                    const int32_t col = q * 32 + t;
                    ccol[c0 + o] = col;
                    cval[c0 + o] = s_val[col];
                    ++o;
                }
            }
            carry += tot;
        }
        if (!NUMERIC && threadIdx.x == 0) row_nnz[r] = (int32_t)carry;
        __syncthreads();
    }
}

}  // namespace

Graph *spgemm_run(Ctx *c, const Graph *a, const Graph *b) {
    const int32_t m = a->m;
    const int grid = fill_grid(c, kBlock, 4);
    apspie_graph *C = new apspie_graph();
    C->owner = c;
    C->device = c->device;
    C->m = m;
    C->n = b->n;
    C->owns = true;
    try {
        C->rowptr = (int32_t *)dev_alloc(c, sizeof(int32_t) * (m + 1));
        int64_t *ub = (int64_t *)scratch(c, "sg_ub", sizeof(int64_t) * std::max(m, 1));
        int32_t *row_nnz = (int32_t *)scratch(c, "sg_rownnz", sizeof(int32_t) * (m + 1));
        int32_t *lists = (int32_t *)scratch(c, "sg_lists", sizeof(int32_t) * 4 * (size_t)std::max(m, 1));
        uint32_t *ctl = (uint32_t *)scratch(c, "sg_ctl", 64);
        fill_i32(c, (int32_t *)ctl, 0, 16);
        // tier bounds (tunable below the tables' capacity): the wave tier's
        // table holds rows of up to kWaveUb products, the block tier's up to
        // kBlockUb; longer rows use the dense accumulator (B with at most
        // kDenseN columns, unless switched off) or global-memory tables
        const bool dense_ok = b->n <= kDenseN && opt(c, "spgemm_dense", 1) != 0;
        const int32_t wave_ub =
            (int32_t)std::min<int64_t>(std::max<int64_t>(opt(c, "spgemm_wave_ub", kWaveUb), 1), kWaveUb);
        const int32_t block_ub = (int32_t)std::min<int64_t>(
            std::max<int64_t>(opt(c, "spgemm_block_ub", kBlockUb), wave_ub), kBlockUb);
        if (m > 0) {
            k_row_bound<<<grid, kBlock, 0, c->stream>>>(a->rowptr, a->col, b->rowptr, m, b->n, ub);
            k_bin_rows<<<grid, kBlock, 0, c->stream>>>(ub, m, wave_ub, block_ub, dense_ok, lists,
                                                       ctl, row_nnz);
            APSPIE_LAUNCH_CHECK();
        }
        uint32_t h[5] = {0, 0, 0, 0, 0};
        APSPIE_HIP(hipMemcpyAsync(h, ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        fill_i32(c, row_nnz + m, 0, 1);
        int64_t gcap = 1;
        while (gcap < 2 * (int64_t)h[4]) gcap <<= 1;
        // global tables: at most ~2 GiB of keys + values in flight
        const int64_t by_mem = std::max<int64_t>(1, ((int64_t)1 << 31) / (8 * gcap));
        const int gblocks = (int)std::min<int64_t>(
            std::min<int64_t>(std::max<uint32_t>(h[3], 1), c->cus), by_mem);
        int32_t *gkeys = h[3] ? (int32_t *)scratch(c, "sg_gkeys", sizeof(int32_t) * gcap * gblocks) : nullptr;
        float *gvals = h[3] ? (float *)scratch(c, "sg_gvals", sizeof(float) * gcap * gblocks) : nullptr;
        auto pass = [&](auto numeric_tag, int32_t *crp, int32_t *ccol, float *cval) {
            constexpr bool NUM = decltype(numeric_tag)::value;
            if (h[3])  // longest rows first
                k_rows_table<NUM, kBig, 0><<<gblocks, kBig, 0, c->stream>>>(
                    lists + 3 * (size_t)m, ctl + 3, a->rowptr, a->col, a->val, b->rowptr, b->col,
                    b->val, row_nnz, crp, ccol, cval, ub, gkeys, gvals, gcap, b->n);
            if (h[2])
                k_rows_dense<NUM><<<(int)std::min<int64_t>(h[2], c->cus), kBig, 0, c->stream>>>(
                    lists + 2 * (size_t)m, ctl + 2, b->n, a->rowptr, a->col, a->val, b->rowptr,
                    b->col, b->val, row_nnz, crp, ccol, cval);
            if (h[1])
                k_rows_table<NUM, kBlock, kBlockTable>
                    <<<(int)std::min<int64_t>(h[1], 4 * c->cus), kBlock, 0, c->stream>>>(
                        lists + m, ctl + 1, a->rowptr, a->col, a->val, b->rowptr, b->col, b->val,
                        row_nnz, crp, ccol, cval, ub, nullptr, nullptr, 0, b->n);
            if (h[0])
                k_rows_wave<NUM><<<(int)std::min<int64_t>(ceil_div(h[0], kBlock / 64), 8 * c->cus),
                                   kBlock, 0, c->stream>>>(lists, ctl + 0, a->rowptr, a->col,
                                                           a->val, b->rowptr, b->col, b->val,
                                                           row_nnz, crp, ccol, cval);
            APSPIE_LAUNCH_CHECK();
        };
        pass(std::false_type{}, nullptr, nullptr, nullptr);
        const int64_t nnz = scan_exclusive_i32(c, row_nnz, C->rowptr, m + 1);
        APSPIE_CHECK(nnz < INT32_MAX, APSPIE_ERR_OVERFLOW, "product has more than 2^31 entries");
        C->nnz = nnz;
        C->col = (int32_t *)dev_alloc(c, sizeof(int32_t) * std::max<int64_t>(nnz, 1));
        C->val = (float *)dev_alloc(c, sizeof(float) * std::max<int64_t>(nnz, 1));
        C->bytes = (int64_t)(sizeof(int32_t) * (m + 1) + 8 * std::max<int64_t>(nnz, 1));
        if (nnz > 0) pass(std::true_type{}, C->rowptr, C->col, C->val);
        sync(c);
        c->stats.edges_traversed = nnz;
        c->stats.iterations = 1;
    } catch (...) {
        if (C->rowptr) dev_free(c, C->rowptr, sizeof(int32_t) * (m + 1));
        if (C->col) dev_free(c, C->col, sizeof(int32_t) * std::max<int64_t>(C->nnz, 1));
        if (C->val) dev_free(c, C->val, sizeof(float) * std::max<int64_t>(C->nnz, 1));
        delete C;
        throw;
    }
    return C;
}

}  // namespace apspie
