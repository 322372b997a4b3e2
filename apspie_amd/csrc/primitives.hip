// Scan, radix sort and compaction for gfx950 (see primitives.hpp).
#include "common.hpp"
#include "device.hpp"
#include "primitives.hpp"

namespace apspie {

namespace {

constexpr int kB = 256;          // threads per block
constexpr int kItems = 8;        // items per thread
constexpr int kTileN = kB * kItems;

// ------------------------------------------------------------------- scan
// Single-pass scan with decoupled look-back.  Tiles take ids from a ticket
// counter in launch order, so every tile a block waits on belongs to a block
// that started earlier and never waits on a later one: every wait ends.
// status word per tile: bits 62-63 = 0 nothing, 1 aggregate, 2 inclusive
// prefix; bits 0-61 = value.
constexpr uint64_t kFlagAgg = uint64_t(1) << 62;
constexpr uint64_t kFlagPre = uint64_t(2) << 62;
constexpr uint64_t kValMask = (uint64_t(1) << 62) - 1;

__global__ __launch_bounds__(kB) void k_scan(const int32_t *in, int32_t *out, int64_t n,
                                             uint64_t *status, uint32_t *ticket,
                                             int64_t *total_dev) {
    __shared__ uint64_t s_scan[kB / 64 + 1];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_tile;
    if (threadIdx.x == 0) s_tile = dev::atomic_add_dev(ticket, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t base = tile * kTileN + (int64_t)threadIdx.x * kItems;
    int32_t v[kItems];
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        sum += (uint64_t)(int64_t)v[k];
    }
    uint64_t agg;
    const uint64_t excl = dev::block_exclusive_add64<kB>(sum, s_scan, agg);
    if (threadIdx.x == 0) {
        uint64_t prefix = 0;
        if (tile == 0) {
            __hip_atomic_store(status, kFlagPre | (agg & kValMask), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(status + tile, kFlagAgg | (agg & kValMask), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            for (int64_t j = tile - 1; j >= 0;) {
                const uint64_t s = dev::load_relaxed(status + j);
                const uint64_t f = s & ~kValMask;
                if (f == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += s & kValMask;
                if (f == kFlagPre) break;
                --j;
            }
            __hip_atomic_store(status + tile, kFlagPre | ((prefix + agg) & kValMask),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_prefix = prefix;
        const int64_t ntiles = (n + kTileN - 1) / kTileN;
        if (tile == ntiles - 1 && total_dev) *total_dev = (int64_t)(prefix + agg);
    }
    __syncthreads();
    int64_t run = (int64_t)(s_prefix + excl);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (base + k < n) out[base + k] = (int32_t)run;
        run += v[k];
    }
}

__global__ void k_zero_u64(uint64_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// ------------------------------------------------------------- radix sort
// LSD passes over 8-bit digits.  Pass = per-tile digit histogram ->
// exclusive scan of the digit-major count table -> stable scatter.  Within a
// tile, elements are visited round by round (element = round*256 + thread),
// i.e. in input order; a lane's rank among equal digits in its wave comes from
// eight ballots (one per digit bit), waves and rounds are chained through LDS.
__global__ __launch_bounds__(kB) void k_rs_hist(const uint32_t *keys, int64_t n, int shift,
                                                int32_t *counts, int64_t ntiles) {
    __shared__ int32_t s_h[256];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTileN;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t i = base + k * kB + threadIdx.x;
        if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = s_h[threadIdx.x];
}

template <bool HAS_VAL>
__global__ __launch_bounds__(kB) void k_rs_scatter(const uint32_t *keys, const uint32_t *vals,
                                                   uint32_t *okeys, uint32_t *ovals, int64_t n,
                                                   int shift, const int32_t *offsets,
                                                   int64_t ntiles) {
    constexpr int NW = kB / 64;
    __shared__ int32_t s_base[256];
    __shared__ int32_t s_run[256];
    __shared__ int32_t s_wc[NW][256];
    const int w = threadIdx.x / 64;
    s_base[threadIdx.x] = offsets[(int64_t)threadIdx.x * ntiles + blockIdx.x];
    s_run[threadIdx.x] = 0;
    const int64_t base = (int64_t)blockIdx.x * kTileN;
    for (int k = 0; k < kItems; ++k) {
#pragma unroll
        for (int q = 0; q < NW; ++q) s_wc[q][threadIdx.x] = 0;
        __syncthreads();
        const int64_t i = base + k * kB + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bm = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bm : ~bm;
        }
        const int rank = __popcll(peers & dev::lanemask_lt());
        if (valid && rank == 0) s_wc[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = s_base[d] + s_run[d] + rank;
            for (int q = 0; q < w; ++q) pos += s_wc[q][d];
            if (pos < n) {  // always true for a consistent count table
                okeys[pos] = key;
                if (HAS_VAL) ovals[pos] = vals[i];
            }
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) add += s_wc[q][threadIdx.x];
        s_run[threadIdx.x] += add;
        __syncthreads();
    }
}

// ------------------------------------------------------------- compaction
__global__ __launch_bounds__(kB) void k_flag_count(const int32_t *flag, int64_t n,
                                                   int32_t *block_counts) {
    __shared__ int32_t s[kB / 64];
    const int64_t base = (int64_t)blockIdx.x * kTileN;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t i = base + k * kB + threadIdx.x;
        c += (i < n && flag[i] != 0) ? 1 : 0;
    }
    const uint32_t wc = (uint32_t)dev::wave_sum_u64((uint64_t)c);
    if (dev::lane_id() == 0) s[threadIdx.x / 64] = (int32_t)wc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t t = 0;
        for (int q = 0; q < kB / 64; ++q) t += s[q];
        block_counts[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kB) void k_flag_write(const int32_t *flag, int64_t n,
                                                   const int32_t *block_offsets, int32_t *out) {
    constexpr int NW = kB / 64;
    __shared__ int32_t s_wc[NW];
    __shared__ int32_t s_run;
    if (threadIdx.x == 0) s_run = block_offsets[blockIdx.x];
    const int w = threadIdx.x / 64;
    const int64_t base = (int64_t)blockIdx.x * kTileN;
    for (int k = 0; k < kItems; ++k) {
        const int64_t i = base + k * kB + threadIdx.x;
        const bool f = i < n && flag[i] != 0;
        const uint64_t bm = __ballot(f);
        if (dev::lane_id() == 0) s_wc[w] = __popcll(bm);
        __syncthreads();
        if (f) {
            int pos = s_run + __popcll(bm & dev::lanemask_lt());
            for (int q = 0; q < w; ++q) pos += s_wc[q];
            if (pos < n) out[pos] = (int32_t)i;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int q = 0; q < NW; ++q) t += s_wc[q];
            s_run += t;
        }
        __syncthreads();
    }
}

}  // namespace

void scan_exclusive_i32_async(Ctx *c, const int32_t *in, int32_t *out, int64_t n,
                              int64_t *total_dev) {
    if (n <= 0) {
        k_zero_u64<<<1, 64, 0, c->stream>>>((uint64_t *)total_dev, 1);
        APSPIE_LAUNCH_CHECK();
        return;
    }
    const int64_t ntiles = ceil_div(n, kTileN);
    uint64_t *status = (uint64_t *)scratch(c, "scan_status", sizeof(uint64_t) * (ntiles + 1));
    uint32_t *ticket = (uint32_t *)(status + ntiles);
    k_zero_u64<<<(int)std::min<int64_t>(ceil_div(ntiles + 1, 256), 1024), 256, 0, c->stream>>>(
        status, ntiles + 1);
    k_scan<<<(int)ntiles, kB, 0, c->stream>>>(in, out, n, status, ticket, total_dev);
    APSPIE_LAUNCH_CHECK();
}

int64_t scan_exclusive_i32(Ctx *c, const int32_t *in, int32_t *out, int64_t n) {
    int64_t *tot = (int64_t *)scratch(c, "scan_total", sizeof(int64_t));
    scan_exclusive_i32_async(c, in, out, n, tot);
    int64_t h = 0;
    APSPIE_HIP(hipMemcpyAsync(&h, tot, sizeof h, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return h;
}

int64_t scan_exclusive_dev(Ctx *c, const int32_t *in, int32_t *out, int64_t n) {
    return scan_exclusive_i32(c, in, out, n);
}

void radix_sort_pairs(Ctx *c, uint32_t *keys, uint32_t *vals, int64_t n, int bits) {
    if (n <= 1) return;
    APSPIE_CHECK(n < INT32_MAX, APSPIE_ERR_UNSUPPORTED, "sort size must fit int32");
    const int64_t ntiles = ceil_div(n, kTileN);
    uint32_t *k2 = (uint32_t *)scratch(c, "rs_k2", sizeof(uint32_t) * n);
    uint32_t *v2 = vals ? (uint32_t *)scratch(c, "rs_v2", sizeof(uint32_t) * n) : nullptr;
    int32_t *counts = (int32_t *)scratch(c, "rs_counts", sizeof(int32_t) * 256 * ntiles);
    int32_t *offs = (int32_t *)scratch(c, "rs_offs", sizeof(int32_t) * 256 * ntiles);
    int64_t *tot = (int64_t *)scratch(c, "rs_total", sizeof(int64_t));
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    const int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        k_rs_hist<<<(int)ntiles, kB, 0, c->stream>>>(ka, n, shift, counts, ntiles);
        scan_exclusive_i32_async(c, counts, offs, 256 * ntiles, tot);
        if (vals)
            k_rs_scatter<true><<<(int)ntiles, kB, 0, c->stream>>>(ka, va, kb, vb, n, shift, offs,
                                                                   ntiles);
        else
            k_rs_scatter<false><<<(int)ntiles, kB, 0, c->stream>>>(ka, nullptr, kb, nullptr, n,
                                                                    shift, offs, ntiles);
        APSPIE_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        APSPIE_HIP(hipMemcpyAsync(keys, ka, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, c->stream));
        if (vals)
            APSPIE_HIP(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice,
                                      c->stream));
    }
}

void sort_pairs_dev(Ctx *c, uint32_t *keys, uint32_t *vals, int64_t n, int bits) {
    radix_sort_pairs(c, keys, vals, n, bits);
}

int64_t compact_flags(Ctx *c, const int32_t *flag, int64_t n, int32_t *idx_out) {
    if (n <= 0) return 0;
    const int64_t nb = ceil_div(n, kTileN);
    int32_t *bc = (int32_t *)scratch(c, "cf_counts", sizeof(int32_t) * nb);
    int32_t *bo = (int32_t *)scratch(c, "cf_offs", sizeof(int32_t) * nb);
    k_flag_count<<<(int)nb, kB, 0, c->stream>>>(flag, n, bc);
    const int64_t total = scan_exclusive_i32(c, bc, bo, nb);
    k_flag_write<<<(int)nb, kB, 0, c->stream>>>(flag, n, bo, idx_out);
    APSPIE_LAUNCH_CHECK();
    return total;
}

}  // namespace apspie
