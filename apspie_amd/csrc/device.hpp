// Device-side building blocks for gfx950 (wave64): lane masks, wave and block
// scans, and atomics with explicit scope.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace apspie {
namespace dev {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// bit mask of lanes below this one
__device__ __forceinline__ uint64_t lanemask_lt() {
    return (uint64_t(1) << lane_id()) - 1;  // lane 0 -> 0
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// Orders LDS/global accesses of the lanes of one wave (a wave-level barrier
// that is also a memory fence, so the compiler waits for outstanding LDS ops).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive wave-wide scan (+) of 32-bit values.  Uses the DPP row shifts
// within each 16-lane row and row broadcasts across rows, the CDNA idiom.
__device__ __forceinline__ uint32_t wave_inclusive_add(uint32_t x) {
    // row_shr:1,2,4,8 (bound_ctrl: lanes without a source read 0)
    uint32_t t;
    t = __builtin_amdgcn_update_dpp(0u, x, 0x111, 0xf, 0xf, false); x += t;
    t = __builtin_amdgcn_update_dpp(0u, x, 0x112, 0xf, 0xf, false); x += t;
    t = __builtin_amdgcn_update_dpp(0u, x, 0x114, 0xf, 0xf, false); x += t;
    t = __builtin_amdgcn_update_dpp(0u, x, 0x118, 0xf, 0xf, false); x += t;
    // row_bcast:15 -> rows 1 and 3 get the last lane of the previous row
    t = __builtin_amdgcn_update_dpp(0u, x, 0x142, 0xa, 0xf, false); x += t;
    // row_bcast:31 -> rows 2 and 3 get lane 31
    t = __builtin_amdgcn_update_dpp(0u, x, 0x143, 0xc, 0xf, false); x += t;
    return x;
}

// Reference implementation with cross-lane shuffles (used for 64-bit values).
__device__ __forceinline__ uint64_t wave_inclusive_add64(uint64_t x) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, d, 64);
        if (lane_id() >= d) x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_read_lane(uint32_t x, int lane) {
    return __builtin_amdgcn_readlane(x, lane);
}

__device__ __forceinline__ uint64_t wave_read_lane64(uint64_t x, int lane) {
    uint32_t lo = __builtin_amdgcn_readlane(uint32_t(x), lane);
    uint32_t hi = __builtin_amdgcn_readlane(uint32_t(x >> 32), lane);
    return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t y = __shfl_xor(x, d, 64);
        x = x > y ? x : y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_or_u32(uint32_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x |= __shfl_xor(x, d, 64);
    return x;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t y = __shfl_xor(x, d, 64);
        x = x < y ? x : y;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Block-wide exclusive scan of uint64 values; returns the exclusive prefix and
// writes the block total.  `lds` needs (blockDim/64 + 1) uint64 slots.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_exclusive_add64(uint64_t x, uint64_t *lds,
                                                          uint64_t &total) {
    constexpr int NW = BLOCK / 64;
    const int w = threadIdx.x / 64;
    uint64_t inc = wave_inclusive_add64(x);
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            uint64_t t = lds[i];
            lds[i] = run;
            run += t;
        }
        lds[NW] = run;
    }
    __syncthreads();
    total = lds[NW];
    uint64_t r = lds[w] + inc - x;
    __syncthreads();
    return r;
}

template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_add32(uint32_t x, uint32_t *lds,
                                                          uint32_t &total) {
    constexpr int NW = BLOCK / 64;
    const int w = threadIdx.x / 64;
    uint32_t inc = wave_inclusive_add(x);
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            uint32_t t = lds[i];
            lds[i] = run;
            run += t;
        }
        lds[NW] = run;
    }
    __syncthreads();
    total = lds[NW];
    uint32_t r = lds[w] + inc - x;
    __syncthreads();
    return r;
}

// Slot of `key` in an open-addressing table of mask + 1 (a power of two)
// slots: Fibonacci hashing, the HIGH bits of key * 2^32 / phi.  (The low bits
// of the product depend only on the key's low bits, and R-MAT hub ids share
// long runs of low zero bits: masking the low bits puts them in one probe
// chain.)
__device__ __forceinline__ uint32_t fib_slot(uint32_t key, uint32_t mask) {
    return mask ? (key * 2654435769u) >> __clz(mask) : 0u;
}

// ---------------------------------------------------------------- atomics
// Device-scope relaxed atomics: every CU on every XCD sees one order.
__device__ __forceinline__ uint32_t atomic_or_dev(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_add_dev(uint64_t *p, uint64_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_add_dev(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t atomic_min_dev(int32_t *p, int32_t v) {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_max_dev(uint64_t *p, uint64_t v) {
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t atomic_exch_dev(int32_t *p, int32_t v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t atomic_cas_dev(int32_t *p, int32_t cmp, int32_t v) {
    __hip_atomic_compare_exchange_strong(p, &cmp, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return cmp;
}
__device__ __forceinline__ uint32_t load_relaxed(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t load_relaxed(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t load_relaxed(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// float <-> order-preserving int for non-negative floats (distances >= 0)
__device__ __forceinline__ int32_t fbits(float f) { return __float_as_int(f); }
__device__ __forceinline__ float bitsf(int32_t i) { return __int_as_float(i); }

// largest j in [lo, hi) with a[j] <= key (a sorted, a[lo] <= key)
template <class T>
__device__ __forceinline__ int upper_bound_m1(const T *a, int lo, int hi, T key) {
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid; else hi = mid;
    }
    return lo;
}

}  // namespace dev
}  // namespace apspie
