// Masked sparse product C = (A * B) o M over (+, x), evaluated only on the
// pattern of M: c_val[e] = sum_k A[i,k] * B[k,j] for every entry e = (i, j) of
// M, in M's CSR order; *c_nnz = entries of M with at least one product (the
// structural nonzeros of the masked product).
//
// Reference: the TC composition materialises D = L * U with cuSPARSE and then
// keeps D where U is nonzero (spgemm() tc.cuh:62-165, ewiseMultTc()
// tc.cuh:168-199).  Here the product is never formed: one block per row i of
// M keeps M_i in an LDS hash table (column -> position) with a value and a
// hit counter per position, walks the products of A_i x B (row entries of A_i
// in chunks, each chunk's products load-balanced over the block by a scan of
// |B_k| and a binary search) and accumulates only those whose column is in
// M_i.  Rows of M longer than kLdsRow use a per-block region of global memory
// for the same table.  Values are exact for integer-valued inputs; otherwise
// the (+) order follows the LDS atomics.
#include <algorithm>

#include "common.hpp"
#include "device.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;
constexpr int kLdsRow = 2048;           // mask rows up to this length: LDS table
constexpr int kLdsTable = 2 * kLdsRow;  // open addressing, load <= 1/2
constexpr int kChunk = 1024;            // entries of A_i per enumeration chunk

__device__ __forceinline__ void mtable_insert(int32_t *keys, int32_t *pos, uint32_t mask,
                                              int32_t key, int32_t p) {
    uint32_t h = dev::fib_slot((uint32_t)key, mask);
    for (;;) {
        const int32_t prev = atomicCAS(keys + h, -1, key);
        if (prev == -1) {
            pos[h] = p;
            return;
        }
        if (prev == key) return;  // a repeated column keeps its first position
        h = (h + 1) & mask;
    }
}

// position of key in the table, -1 if absent (called after all inserts)
__device__ __forceinline__ int32_t mtable_find(const int32_t *keys, const int32_t *pos,
                                               uint32_t mask, int32_t key) {
    uint32_t h = dev::fib_slot((uint32_t)key, mask);
    for (;;) {
        const int32_t t = keys[h];
        if (t == key) return pos[h];
        if (t == -1) return -1;
        h = (h + 1) & mask;
    }
}

struct MaskedArgs {
    const int32_t *arp, *acol;
    const float *aval;  // null: all ones
    const int32_t *brp, *bcol;
    const float *bval;
    const int32_t *mrp, *mcol;
    int32_t m;
    int64_t bnnz;
    float *cval;
    uint64_t *nnz_out;
    // spill regions (rows of M longer than kLdsRow), per block
    int32_t *g_keys, *g_pos, *g_hits;
    float *g_acc;
    int64_t g_table, g_row;
};

__global__ __launch_bounds__(kBlock) void k_masked_rows(MaskedArgs p) {
    __shared__ int32_t s_keys[kLdsTable], s_pos[kLdsTable];
    __shared__ float s_acc[kLdsRow];
    __shared__ int32_t s_hits[kLdsRow];
    __shared__ int32_t s_pre[kChunk + 1], s_rs[kChunk];
    __shared__ float s_av[kChunk];
    __shared__ uint32_t s_scan[kBlock / 64 + 1];
    __shared__ uint64_t s_red[kBlock / 64];
    uint64_t my_nnz = 0;
    for (int32_t i = blockIdx.x; i < p.m; i += gridDim.x) {
        const int32_t mb = p.mrp[i], n = p.mrp[i + 1] - mb;
        if (n == 0) continue;  // uniform across the block
        const bool lds = n <= kLdsRow;
        uint32_t tsize = kLdsTable;
        if (!lds) {
            tsize = 1;
            while (tsize < 2u * (uint32_t)n) tsize <<= 1;
        }
        const uint32_t tmask = tsize - 1;
        int32_t *keys = lds ? s_keys : p.g_keys + (int64_t)blockIdx.x * p.g_table;
        int32_t *pos = lds ? s_pos : p.g_pos + (int64_t)blockIdx.x * p.g_table;
        float *acc = lds ? s_acc : p.g_acc + (int64_t)blockIdx.x * p.g_row;
        int32_t *hits = lds ? s_hits : p.g_hits + (int64_t)blockIdx.x * p.g_row;
        for (uint32_t s = threadIdx.x; s < tsize; s += kBlock) keys[s] = -1;
        for (int32_t q = threadIdx.x; q < n; q += kBlock) {
            acc[q] = 0.0f;
            hits[q] = 0;
        }
        __syncthreads();
        for (int32_t q = threadIdx.x; q < n; q += kBlock)
            mtable_insert(keys, pos, tmask, p.mcol[mb + q], q);
        __syncthreads();
        const int32_t ab = p.arp[i], an = p.arp[i + 1] - ab;
        for (int32_t c0 = 0; c0 < an; c0 += kChunk) {
            const int32_t cn = min(kChunk, an - c0);
            // products of this chunk: prefix of |B_k| over its entries
            uint32_t carry = 0;
            for (int32_t x0 = 0; x0 < cn; x0 += kBlock) {
                const int32_t x = x0 + threadIdx.x;
                uint32_t len = 0;
                if (x < cn) {
                    const int32_t k = p.acol[ab + c0 + x];
                    const int32_t rs = p.brp[k];
                    len = (uint32_t)(p.brp[k + 1] - rs);
                    s_rs[x] = rs;
                    s_av[x] = p.aval ? p.aval[ab + c0 + x] : 1.0f;
                }
                uint32_t tot;
                const uint32_t ex = dev::block_exclusive_add32<kBlock>(len, s_scan, tot);
                if (x < cn) s_pre[x] = (int32_t)(carry + ex);
                carry += tot;
            }
            if (threadIdx.x == 0) s_pre[cn] = (int32_t)carry;
            __syncthreads();
            const int32_t W = (int32_t)carry;
            for (int32_t x = threadIdx.x; x < W; x += kBlock) {
                const int j = dev::upper_bound_m1(s_pre, 0, cn, x);
                const int64_t e = (int64_t)s_rs[j] + (x - s_pre[j]);
                if ((uint64_t)e >= (uint64_t)p.bnnz) continue;  // guard only
                const int32_t q = mtable_find(keys, pos, tmask, p.bcol[e]);
                if (q < 0) continue;
                atomicAdd(acc + q, s_av[j] * (p.bval ? p.bval[e] : 1.0f));
                atomicAdd(hits + q, 1);
            }
            __syncthreads();  // s_pre / s_rs / s_av are refilled by the next chunk
        }
        uint32_t c = 0;
        for (int32_t q = threadIdx.x; q < n; q += kBlock) {
            p.cval[mb + q] = acc[q];
            c += hits[q] > 0 ? 1u : 0u;
        }
        my_nnz += c;
        __syncthreads();  // the table is reset for the next row
    }
    const uint64_t ws = dev::wave_sum_u64(my_nnz);
    if (dev::lane_id() == 0) s_red[threadIdx.x / 64] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int q = 0; q < kBlock / 64; ++q) t += s_red[q];
        if (t) dev::atomic_add_dev(p.nnz_out, t);
    }
}

__global__ void k_max_row(const int32_t *rp, int32_t m, int32_t *out) {
    int32_t best = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        best = max(best, rp[i + 1] - rp[i]);
    for (int off = 32; off > 0; off >>= 1) best = max(best, __shfl_xor(best, off, 64));
    if (dev::lane_id() == 0 && best > 0) atomicMax(out, best);
}

}  // namespace

int64_t spgemm_masked_run(Ctx *c, const Graph *a, const Graph *b, const Graph *mk, float *cval) {
    const int grid = fill_grid(c, kBlock, 4);
    uint64_t *ctl = (uint64_t *)scratch(c, "msk_ctl", 64);
    fill_i32(c, (int32_t *)ctl, 0, 16);
    if (mk->nnz == 0 || mk->m == 0) return 0;
    k_max_row<<<std::min(grid, (int)ceil_div(mk->m, kBlock)), kBlock, 0, c->stream>>>(
        mk->rowptr, mk->m, (int32_t *)(ctl + 1));
    APSPIE_LAUNCH_CHECK();
    int32_t maxn = 0;
    APSPIE_HIP(hipMemcpyAsync(&maxn, ctl + 1, 4, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    MaskedArgs p{};
    p.arp = a->rowptr;
    p.acol = a->col;
    p.aval = a->val;
    p.brp = b->rowptr;
    p.bcol = b->col;
    p.bval = b->val;
    p.mrp = mk->rowptr;
    p.mcol = mk->col;
    p.m = mk->m;
    p.bnnz = b->nnz;
    p.cval = cval;
    p.nnz_out = ctl;
    if (maxn > kLdsRow) {  // spill regions for the long mask rows
        int64_t t = 1;
        while (t < 2 * (int64_t)maxn) t <<= 1;
        p.g_table = t;
        p.g_row = maxn;
        p.g_keys = (int32_t *)scratch(c, "msk_gk", sizeof(int32_t) * t * grid);
        p.g_pos = (int32_t *)scratch(c, "msk_gp", sizeof(int32_t) * t * grid);
        p.g_acc = (float *)scratch(c, "msk_ga", sizeof(float) * maxn * (int64_t)grid);
        p.g_hits = (int32_t *)scratch(c, "msk_gh", sizeof(int32_t) * maxn * (int64_t)grid);
    }
    k_masked_rows<<<grid, kBlock, 0, c->stream>>>(p);
    APSPIE_LAUNCH_CHECK();
    uint64_t nnz = 0;
    APSPIE_HIP(hipMemcpyAsync(&nnz, ctl, 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return (int64_t)nnz;
}

}  // namespace apspie
