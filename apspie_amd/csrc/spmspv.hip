// The SpMSpV primitive y = A^T x under a semiring, as explicit stages.
//
// Reference: sparse mXv (mXv.cuh:249-389) and its per-algorithm copies
// (spmspvBfs.cuh:174-222, spmspvSssp.cuh:172-241, spmspvMis.cuh:245-293):
//   1. frontier column gather: start/degree of every x entry (IntervalGather)
//   2. exclusive scan of column lengths (mgpu::Scan)
//   3. load-balanced edge expansion: per expanded edge (key = column,
//      value = x_u (x) a_uv) (IntervalGather / IntervalExpand / elementMult)
//      with the mask filter applied at expansion (bfs.cuh:16-30)
//   4a. SORT path: stable key-value radix sort on the low ceil(log2 n) bits
//       (cub::DeviceRadixSort::SortPairs) then segmented reduce by key
//       (mgpu::ReduceByKey).  Each run is folded by the thread that owns its
//       head, in sorted (= expansion) order, so results are bit-identical to
//       a sequential fold in (x order, CSR order) even for non-integer data.
//   4b. ATOMIC path: atomic semiring-add into a dense row, then an
//       order-preserving compaction of the touched entries.  Exact for
//       integer-valued data and for the idempotent semirings (or/min/max);
//       (+) on non-integer values depends on the atomic order.
// Output: sorted unique indices and their values.
#include <cmath>

#include "common.hpp"
#include "device.hpp"
#include "primitives.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 2048;  // expanded edges per partition

__device__ __forceinline__ float sr_mul(int sr, float x, float a) {
    switch (sr) {
    case APSPIE_SR_PLUS_TIMES: return x * a;
    case APSPIE_SR_BOOL: return 1.0f;
    case APSPIE_SR_MIN_PLUS: return x + a;
    default: return x;  // max-priority: the source's value
    }
}

__device__ __forceinline__ float sr_add(int sr, float a, float b) {
    switch (sr) {
    case APSPIE_SR_PLUS_TIMES: return a + b;
    case APSPIE_SR_BOOL: return 1.0f;
    case APSPIE_SR_MIN_PLUS: return b < a ? b : a;
    default: return b > a ? b : a;
    }
}

// stage 1: gather row start and length of every frontier entry
__global__ void k_gather(const int32_t *rowptr, const int32_t *x_ind, int64_t nx,
                         int32_t *start, int32_t *deg) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nx;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = x_ind[i];
        const int32_t s = rowptr[u];
        start[i] = s;
        deg[i] = rowptr[u + 1] - s;
    }
}

// stage 3a: tile partition -- tfirst[t] = last entry with off <= t*kTile
__global__ void k_partition(const int32_t *off, int64_t nx, int64_t ntiles, int32_t *tfirst) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t * kTile;
        int64_t lo = 0, hi = nx;  // largest i with off[i] <= e
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (off[mid] <= e) lo = mid; else hi = mid;
        }
        tfirst[t] = (int32_t)lo;
    }
}

// stage 3b: expansion.  Slot e belongs to the last entry i with off[i] <= e
// (zero-length entries share an offset with their successor and lose).
template <bool SORT_OUT>
__global__ __launch_bounds__(kBlock) void k_expand_pairs(
    int sr, const int32_t *col, int32_t nnz, const float *aval, const float *x_val,
    const int32_t *start,
    const int32_t *off, int64_t nx, const int32_t *tfirst, int64_t total, const int32_t *mask,
    int32_t n_out, uint32_t *keys, float *vals, float *ydense, uint32_t *touched) {
    const int64_t ntiles = (total + kTile - 1) / kTile;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int32_t i0 = tfirst[t];
        const int32_t i1 = (t + 1 < ntiles) ? tfirst[t + 1] : (int32_t)(nx - 1);
        for (int k = 0; k < kTile / kBlock; ++k) {
            const int64_t e = t * kTile + k * kBlock + threadIdx.x;
            if (e >= total) break;
            int32_t lo = i0, hi = i1 + 1;  // largest i in [i0, i1] with off[i] <= e
            while (hi - lo > 1) {
                const int32_t mid = (lo + hi) >> 1;
                if (off[mid] <= e) lo = mid; else hi = mid;
            }
            const int32_t idx = start[lo] + (int32_t)(e - off[lo]);
            if ((uint32_t)idx >= (uint32_t)nnz) continue;  // guard only
            const int32_t v = col[idx];
            const float xv = x_val ? x_val[lo] : 1.0f;
            const float prod = sr_mul(sr, xv, aval ? aval[idx] : 1.0f);
            const bool drop = mask && mask[v] != 0;
            if (SORT_OUT) {
                keys[e] = drop ? (uint32_t)n_out : (uint32_t)v;  // masked -> past the end
                vals[e] = prod;
            } else if (!drop) {
                float *p = ydense + v;
                switch (sr) {
                case APSPIE_SR_PLUS_TIMES: atomicAdd(p, prod); break;
                case APSPIE_SR_BOOL: *p = 1.0f; break;
                case APSPIE_SR_MIN_PLUS: {
                    float cur = *p;
                    while (prod < cur) {
                        const float seen = __int_as_float(atomicCAS(
                            (int *)p, __float_as_int(cur), __float_as_int(prod)));
                        if (seen == cur) break;
                        cur = seen;
                    }
                    break;
                }
                default: {
                    float cur = *p;
                    while (prod > cur) {
                        const float seen = __int_as_float(atomicCAS(
                            (int *)p, __float_as_int(cur), __float_as_int(prod)));
                        if (seen == cur) break;
                        cur = seen;
                    }
                    break;
                }
                }
                atomicOr(touched + (v >> 5), 1u << (v & 31));
            }
        }
    }
}

// stage 4a: run heads of the sorted keys (keys < n_out only)
__global__ void k_heads(const uint32_t *keys, int64_t total, uint32_t n_out, int32_t *head) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x)
        head[e] = keys[e] < n_out && (e == 0 || keys[e] != keys[e - 1]) ? 1 : 0;
}

// fold each run sequentially from its head
__global__ void k_fold_runs(int sr, const uint32_t *keys, const float *vals, int64_t total,
                            const int32_t *heads_idx, int64_t nruns, int32_t *y_ind,
                            float *y_val) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nruns;
         r += (int64_t)gridDim.x * blockDim.x) {
        int64_t e = heads_idx[r];
        const uint32_t key = keys[e];
        float acc = vals[e];
        for (++e; e < total && keys[e] == key; ++e) acc = sr_add(sr, acc, vals[e]);
        y_ind[r] = (int32_t)key;
        y_val[r] = acc;
    }
}

__global__ void k_fill_f32(float *p, float v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_touched_flags(const uint32_t *touched, int64_t n, int32_t *flag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = (touched[i >> 5] >> (i & 31)) & 1u;
}

__global__ void k_gather_dense(const int32_t *idx, int64_t k, const float *ydense, float *y_val) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k;
         i += (int64_t)gridDim.x * blockDim.x)
        y_val[i] = ydense[idx[i]];
}

// the dense mXv's frontier: entries above 0.5 (the reference's bitify)
__global__ void k_above_half(const float *x, int64_t n, int32_t *flag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = x[i] > 0.5f ? 1 : 0;
}

__global__ void k_scatter_dense(const int32_t *idx, const float *val, int64_t k, float *y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k;
         i += (int64_t)gridDim.x * blockDim.x)
        y[idx[i]] = val[i];
}

float host_identity(int sr) {
    switch (sr) {
    case APSPIE_SR_MIN_PLUS: return INFINITY;
    case APSPIE_SR_MAX_FIRST: return -INFINITY;
    default: return 0.0f;
    }
}

}  // namespace

int64_t spmspv_run(Ctx *c, const Graph *g, int sr, int64_t nx, const int32_t *x_ind,
                   const float *x_val, const int32_t *mask, int32_t *y_ind, float *y_val,
                   int path) {
    if (nx == 0 || g->nnz == 0) {
        c->stats.iterations = 1;
        return 0;
    }
    const int grid = fill_grid(c, kBlock, 4);
    // 1. gather
    int32_t *start = (int32_t *)scratch(c, "spv_start", sizeof(int32_t) * nx);
    int32_t *deg = (int32_t *)scratch(c, "spv_deg", sizeof(int32_t) * nx);
    int32_t *off = (int32_t *)scratch(c, "spv_off", sizeof(int32_t) * nx);
    k_gather<<<grid, kBlock, 0, c->stream>>>(g->rowptr, x_ind, nx, start, deg);
    APSPIE_LAUNCH_CHECK();
    // 2. scan of column lengths
    const int64_t total = scan_exclusive_i32(c, deg, off, nx);
    c->stats.host_syncs++;
    c->stats.edges_traversed = total;
    c->stats.vertices_reached = nx;
    c->stats.iterations = 1;
    if (total == 0) return 0;
    // 3. partition
    const int64_t ntiles = ceil_div(total, kTile);
    int32_t *tfirst = (int32_t *)scratch(c, "spv_tfirst", sizeof(int32_t) * ntiles);
    k_partition<<<(int)std::min<int64_t>(ceil_div(ntiles, kBlock), 1024), kBlock, 0, c->stream>>>(
        off, nx, ntiles, tfirst);
    APSPIE_LAUNCH_CHECK();
    if (path == APSPIE_PATH_AUTO) {
        const int64_t p = opt(c, "spmspv_path", 0);
        path = p ? (int)p
                 : (total <= opt(c, "spmspv_sort_max", 1 << 16) && total * 8 < g->n
                        ? APSPIE_PATH_SORT
                        : APSPIE_PATH_ATOMIC);
    }
    const int egrid = (int)std::min<int64_t>(ntiles, 4 * c->cus);
    if (path == APSPIE_PATH_SORT) {
        uint32_t *keys = (uint32_t *)scratch(c, "spv_keys", sizeof(uint32_t) * total);
        float *vals = (float *)scratch(c, "spv_vals", sizeof(float) * total);
        k_expand_pairs<true><<<egrid, kBlock, 0, c->stream>>>(
            sr, g->col, (int32_t)g->nnz, g->val, x_val, start, off, nx, tfirst, total, mask, g->n, keys, vals,
            nullptr, nullptr);
        APSPIE_LAUNCH_CHECK();
        // 4a. sort on the bits that can be set (keys <= n)
        radix_sort_pairs(c, keys, (uint32_t *)vals, total, bits_for(g->n));
        int32_t *head = (int32_t *)scratch(c, "spv_head", sizeof(int32_t) * total);
        int32_t *hidx = (int32_t *)scratch(c, "spv_hidx", sizeof(int32_t) * total);
        k_heads<<<grid, kBlock, 0, c->stream>>>(keys, total, (uint32_t)g->n, head);
        const int64_t nruns = compact_flags(c, head, total, hidx);
        k_fold_runs<<<grid, kBlock, 0, c->stream>>>(sr, keys, vals, total, hidx, nruns, y_ind,
                                                    y_val);
        APSPIE_LAUNCH_CHECK();
        return nruns;
    }
    // 4b. atomic accumulate into a dense row, then compact
    float *yd = (float *)scratch(c, "spv_dense", sizeof(float) * g->n);
    uint32_t *touched = (uint32_t *)scratch(c, "spv_touched", sizeof(uint32_t) * (g->n / 32 + 1));
    k_fill_f32<<<grid, kBlock, 0, c->stream>>>(yd, host_identity(sr), g->n);
    fill_i32(c, (int32_t *)touched, 0, g->n / 32 + 1);
    k_expand_pairs<false><<<egrid, kBlock, 0, c->stream>>>(sr, g->col, (int32_t)g->nnz, g->val, x_val, start, off,
                                                           nx, tfirst, total, mask, g->n, nullptr,
                                                           nullptr, yd, touched);
    int32_t *flag = (int32_t *)scratch(c, "spv_flag", sizeof(int32_t) * g->n);
    k_touched_flags<<<grid, kBlock, 0, c->stream>>>(touched, g->n, flag);
    APSPIE_LAUNCH_CHECK();
    const int64_t k = compact_flags(c, flag, g->n, y_ind);
    k_gather_dense<<<grid, kBlock, 0, c->stream>>>(y_ind, k, yd, y_val);
    APSPIE_LAUNCH_CHECK();
    return k;
}

// Dense mXv (mXv.cuh:110-247): the entries of x above 0.5 (bitify,
// mXv.cuh:50-55) compacted in index order (Scan + streamCompact, :142-147)
// with their values, expanded and reduced by the SpMSpV stages above, y
// zeroed (preprocessFlag, :185) and the sums scattered (scatterFloat, :205).
// Returns the sum of the frontier rows' lengths (`total`, :229); with no
// frontier y is left as it was and 0 is returned (:143-145).
int64_t mxv_dense_run(Ctx *c, const Graph *g, const float *x, float *y, int path) {
    const int grid = fill_grid(c, kBlock, 4);
    int32_t *flag = (int32_t *)scratch(c, "mxd_flag", sizeof(int32_t) * std::max(g->m, 1));
    int32_t *xi = (int32_t *)scratch(c, "mxd_xi", sizeof(int32_t) * std::max(g->m, 1));
    k_above_half<<<grid, kBlock, 0, c->stream>>>(x, g->m, flag);
    APSPIE_LAUNCH_CHECK();
    const int64_t nx = compact_flags(c, flag, g->m, xi);
    if (nx == 0) {
        c->stats.iterations = 1;
        return 0;
    }
    float *xv = (float *)scratch(c, "mxd_xv", sizeof(float) * nx);
    k_gather_dense<<<grid, kBlock, 0, c->stream>>>(xi, nx, x, xv);
    int32_t *yi = (int32_t *)scratch(c, "mxd_yi", sizeof(int32_t) * std::max(g->n, 1));
    float *yv = (float *)scratch(c, "mxd_yv", sizeof(float) * std::max(g->n, 1));
    APSPIE_LAUNCH_CHECK();
    const int64_t k = spmspv_run(c, g, APSPIE_SR_PLUS_TIMES, nx, xi, xv, nullptr, yi, yv, path);
    const int64_t total = c->stats.edges_traversed;
    c->stats.vertices_reached = nx;
    k_fill_f32<<<grid, kBlock, 0, c->stream>>>(y, 0.0f, g->n);
    if (k > 0) k_scatter_dense<<<grid, kBlock, 0, c->stream>>>(yi, yv, k, y);
    APSPIE_LAUNCH_CHECK();
    return total;
}

}  // namespace apspie
