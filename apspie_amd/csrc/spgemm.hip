// General sparse product C = A * B over (+, x) by row-wise hash accumulation.
//
// Reference: the hash SpGEMM spgemm() (spgemm.cuh:131-342) with its insert
// kernels (spgemmKernel.cuh:179-343): open-addressing tables keyed by output
// coordinates, atomic value accumulation, a sort to order the output.  The
// reference tiles C in 2D over global-memory tables and drops products after
// 25 probes; here each output row gets its own table sized from the row's
// symbolic bound, so nothing is ever dropped:
//   0. symbolic bound ub[i] = sum_{k in A_i} |B_k| (products of row i)
//   1. rows binned by ub: wave + 512-slot LDS table (ub <= 256), block +
//      8192-slot LDS table (ub <= 4096), block + global-memory table (spill)
//   2. symbolic pass: distinct columns per row (key-only inserts)
//   3. exclusive scan -> C row offsets
//   4. numeric pass: key insert + atomic value add, bitonic sort of the
//      table in place (empty slots sort last), the first nnz(row) entries
//      are the sorted row.
// Values are exact for integer-valued inputs; float (+) order otherwise
// follows the atomic order.
#include <algorithm>

#include "common.hpp"
#include "device.hpp"
#include "primitives.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;
constexpr int kWaveUb = 256, kWaveTable = 512;
constexpr int kBlockUb = 4096, kBlockTable = 8192;
constexpr int32_t kEmpty = 0x7fffffff;  // sorts after every column id

__global__ void k_row_bound(const int32_t *arp, const int32_t *acol, const int32_t *brp,
                            int32_t m, int64_t *ub) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t s = 0;
        for (int32_t e = arp[i]; e < arp[i + 1]; ++e) {
            const int32_t k = acol[e];
            s += brp[k + 1] - brp[k];
        }
        ub[i] = s;
    }
}

// bins: 0 wave, 1 block-LDS, 2 block-global; ctl[0..2] counts; ctl[3] max ub (global bin)
__global__ void k_bin_rows(const int64_t *ub, int32_t m, int32_t *lists, uint32_t *ctl,
                           int32_t *row_nnz) {
    for (int64_t r0 = blockIdx.x * (int64_t)blockDim.x; r0 < m; r0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = r0 + threadIdx.x;
        const int64_t u = r < m ? ub[r] : 0;
        const int bin = u == 0 ? -1 : (u <= kWaveUb ? 0 : (u <= kBlockUb ? 1 : 2));
        if (r < m && bin < 0) row_nnz[r] = 0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const uint64_t mask = __ballot(bin == b);
            if (!mask) continue;
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if (dev::lane_id() == leader) base = dev::atomic_add_dev(ctl + b, (uint32_t)__popcll(mask));
            base = __shfl(base, leader, 64);
            if (bin == b) lists[(int64_t)b * m + base + __popcll(mask & dev::lanemask_lt())] = (int32_t)r;
        }
        if (bin == 2) atomicMax(ctl + 3, (uint32_t)min<int64_t>(u, 0x7fffffff));
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

// The row worker, shared by the three tiers.  NUMERIC=false counts distinct
// columns into row_nnz; NUMERIC=true writes the sorted row into C.
template <bool NUMERIC, bool WAVE>
__device__ __forceinline__ void row_product(int32_t r, const int32_t *arp, const int32_t *acol,
                                            const float *aval, const int32_t *brp,
                                            const int32_t *bcol, const float *bval,
                                            int32_t *keys, float *vals, uint32_t tsize,
                                            int tid, int nthreads, int32_t *row_nnz,
                                            const int32_t *crp, int32_t *ccol, float *cval,
                                            uint32_t *s_count) {
    auto barrier = [] {
        if (WAVE) dev::wave_sync(); else __syncthreads();
    };
    const uint32_t mask = tsize - 1;
    for (uint32_t s = tid; s < tsize; s += nthreads) {
        keys[s] = kEmpty;
        if (NUMERIC) vals[s] = 0.0f;
    }
    if (tid == 0) *s_count = 0;
    barrier();
    uint32_t created = 0;
    for (int32_t e = arp[r]; e < arp[r + 1]; ++e) {
        const int32_t k = acol[e];
        const float av = aval ? aval[e] : 1.0f;
        for (int32_t f = brp[k] + tid; f < brp[k + 1]; f += nthreads) {
            uint32_t slot;
            created += tab_insert(keys, mask, bcol[f], &slot) ? 1u : 0u;
            if (NUMERIC) atomicAdd(vals + slot, av * (bval ? bval[f] : 1.0f));
        }
    }
    if (!NUMERIC) {
        if (created) atomicAdd(s_count, created);
        barrier();
        if (tid == 0) row_nnz[r] = (int32_t)*s_count;
        barrier();
        return;
    }
    barrier();
    bitonic(keys, vals, tsize, tid, nthreads, barrier);
    const int32_t c0 = crp[r], nz = crp[r + 1] - c0;
    for (int32_t s = tid; s < nz; s += nthreads) {
        ccol[c0 + s] = keys[s];
        cval[c0 + s] = vals[s];
    }
    barrier();
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
        row_product<NUMERIC, true>(rows[i], arp, acol, aval, brp, bcol, bval, s_keys[w],
                                   NUMERIC ? s_vals[w] : nullptr, kWaveTable, dev::lane_id(), 64,
                                   row_nnz, crp, ccol, cval, &s_count[w]);
}

template <bool NUMERIC, bool GLOBAL>
__global__ __launch_bounds__(kBlock) void k_rows_block(const int32_t *rows, const uint32_t *cnt,
                                                       const int32_t *arp, const int32_t *acol,
                                                       const float *aval, const int32_t *brp,
                                                       const int32_t *bcol, const float *bval,
                                                       int32_t *row_nnz, const int32_t *crp,
                                                       int32_t *ccol, float *cval,
                                                       const int64_t *ub, int32_t *g_keys,
                                                       float *g_vals, int64_t g_cap) {
    __shared__ int32_t s_keys[GLOBAL ? 1 : kBlockTable];
    __shared__ float s_vals[(GLOBAL || !NUMERIC) ? 1 : kBlockTable];
    __shared__ uint32_t s_count;
    const int32_t n = (int32_t)cnt[0];
    int32_t *keys = GLOBAL ? g_keys + (int64_t)blockIdx.x * g_cap : s_keys;
    float *vals = GLOBAL ? (NUMERIC ? g_vals + (int64_t)blockIdx.x * g_cap : nullptr)
                         : (NUMERIC ? s_vals : nullptr);
    for (int32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int32_t r = rows[i];
        uint32_t tsize = kBlockTable;
        if (GLOBAL) {
            tsize = 1;
            while (tsize < 2 * (uint64_t)ub[r]) tsize <<= 1;
        }
        row_product<NUMERIC, false>(r, arp, acol, aval, brp, bcol, bval, keys, vals, tsize,
                                    threadIdx.x, kBlock, row_nnz, crp, ccol, cval, &s_count);
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
        int32_t *lists = (int32_t *)scratch(c, "sg_lists", sizeof(int32_t) * 3 * (size_t)std::max(m, 1));
        uint32_t *ctl = (uint32_t *)scratch(c, "sg_ctl", 64);
        fill_i32(c, (int32_t *)ctl, 0, 16);
        if (m > 0) {
            k_row_bound<<<grid, kBlock, 0, c->stream>>>(a->rowptr, a->col, b->rowptr, m, ub);
            k_bin_rows<<<grid, kBlock, 0, c->stream>>>(ub, m, lists, ctl, row_nnz);
            APSPIE_LAUNCH_CHECK();
        }
        uint32_t h[4] = {0, 0, 0, 0};
        APSPIE_HIP(hipMemcpyAsync(h, ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        fill_i32(c, row_nnz + m, 0, 1);
        int64_t gcap = 1;
        while (gcap < 2 * (int64_t)h[3]) gcap <<= 1;
        // spill tables: at most ~2 GiB of keys + values in flight
        const int64_t by_mem = std::max<int64_t>(1, ((int64_t)1 << 31) / (8 * gcap));
        const int gblocks = (int)std::min<int64_t>(
            std::min<int64_t>(std::max<uint32_t>(h[2], 1), 2 * c->cus), by_mem);
        int32_t *gkeys = h[2] ? (int32_t *)scratch(c, "sg_gkeys", sizeof(int32_t) * gcap * gblocks) : nullptr;
        float *gvals = h[2] ? (float *)scratch(c, "sg_gvals", sizeof(float) * gcap * gblocks) : nullptr;
        auto pass = [&](auto numeric_tag, int32_t *crp, int32_t *ccol, float *cval) {
            constexpr bool NUM = decltype(numeric_tag)::value;
            if (h[0])
                k_rows_wave<NUM><<<(int)std::min<int64_t>(ceil_div(h[0], kBlock / 64), 8 * c->cus),
                                   kBlock, 0, c->stream>>>(lists, ctl + 0, a->rowptr, a->col,
                                                           a->val, b->rowptr, b->col, b->val,
                                                           row_nnz, crp, ccol, cval);
            if (h[1])
                k_rows_block<NUM, false><<<(int)std::min<int64_t>(h[1], 4 * c->cus), kBlock, 0,
                                           c->stream>>>(lists + m, ctl + 1, a->rowptr, a->col,
                                                        a->val, b->rowptr, b->col, b->val, row_nnz,
                                                        crp, ccol, cval, ub, nullptr, nullptr, 0);
            if (h[2])
                k_rows_block<NUM, true><<<gblocks, kBlock, 0, c->stream>>>(
                    lists + 2 * (size_t)m, ctl + 2, a->rowptr, a->col, a->val, b->rowptr, b->col,
                    b->val, row_nnz, crp, ccol, cval, ub, gkeys, gvals, gcap);
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
