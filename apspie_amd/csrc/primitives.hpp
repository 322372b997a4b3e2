// Device-wide primitives of the SpMSpV pipeline: exclusive scan, stable
// key-value radix sort, stream compaction.  The reference takes these from
// moderngpu (mgpu::Scan, mXv.cuh:142,160) and CUB
// (cub::DeviceRadixSort::SortPairs / SortKeys, mXv.cuh:191-193); here they are
// written for wave64 and LDS directly.
#pragma once

#include "common.hpp"

namespace apspie {

// out[i] = sum_{j<i} in[j] (int32 in/out, int64 total returned, host sync)
int64_t scan_exclusive_i32(Ctx *c, const int32_t *in, int32_t *out, int64_t n);
// same without a host sync: total written to *total_dev
void scan_exclusive_i32_async(Ctx *c, const int32_t *in, int32_t *out, int64_t n,
                              int64_t *total_dev);

// stable LSD radix sort of (key, val) by the low `bits` bits of key; vals may
// be null.  Sorts in place (uses scratch ping-pong buffers).
void radix_sort_pairs(Ctx *c, uint32_t *keys, uint32_t *vals, int64_t n, int bits);

// idx_out = the i in [0, n) with flag[i] != 0, in increasing order; returns count
int64_t compact_flags(Ctx *c, const int32_t *flag, int64_t n, int32_t *idx_out);

inline int bits_for(int64_t maxval) {
    int b = 1;
    while (b < 32 && ((int64_t)1 << b) <= maxval) ++b;
    return b;
}

}  // namespace apspie
