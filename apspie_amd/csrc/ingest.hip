// Graph ingestion on the device: COO -> CSR (optionally symmetrized,
// deduplicated and loop-free) and CSR transpose.
//
// Reference: makeSymmetric (util.cuh:571-643: mirror, qsort by (row,col),
// dedup), buildMatrix (matrix.cpp:55-111: counting sort by row that keeps the
// input order inside a row), cusparseXcoo2csr (util.cuh:456-471) and
// cusparseScsr2csc (util.cuh:473-511).  Ordering is done with the stable
// radix sort: sorting by column and then by row orders entries by
// (row, column) with ties in input order, so a duplicate keeps the value of
// its first occurrence; the reference leaves self-loops in and moves values
// out of step with indices during dedup (util.cuh:606-641) -- here loops are
// dropped and values travel with their entry.
#include "common.hpp"
#include "device.hpp"
#include "primitives.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;

__global__ void k_mirror(const int32_t *row, const int32_t *col, int64_t nnz, int32_t *r2,
                         int32_t *c2, int32_t m, int32_t n, int32_t *bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = row[i], c = col[i];
        if (r < 0 || r >= m || c < 0 || c >= n) bad[0] = 1;
        r2[i] = r;
        c2[i] = c;
        r2[nnz + i] = c;
        c2[nnz + i] = r;
    }
}

__global__ void k_copy_check(const int32_t *row, const int32_t *col, int64_t nnz, int32_t *r2,
                             int32_t *c2, int32_t m, int32_t n, int32_t *bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = row[i], c = col[i];
        if (r < 0 || r >= m || c < 0 || c >= n) bad[0] = 1;
        r2[i] = r;
        c2[i] = c;
    }
}

__global__ void k_iota_key(const int32_t *src, int64_t n, uint32_t *keys, uint32_t *idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        idx[i] = (uint32_t)i;
        keys[i] = (uint32_t)src[i];
    }
}

__global__ void k_gather_key(const int32_t *src, const uint32_t *idx, int64_t n, uint32_t *keys) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        keys[i] = (uint32_t)src[idx[i]];
}

// keep sorted entry e unless it is a self-loop or repeats entry e-1
__global__ void k_keep(const int32_t *row, const int32_t *col, const uint32_t *idx, int64_t n,
                       int dedup, int32_t *keep) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t a = idx[e];
        int k = 1;
        if (dedup) {
            if (row[a] == col[a]) k = 0;
            if (e > 0) {
                const uint32_t b = idx[e - 1];
                if (row[b] == row[a] && col[b] == col[a]) k = 0;
            }
        }
        keep[e] = k;
    }
}

// output entry i = input entry idx[kept[i]] (the entries arrive sorted by
// output row): column, value and row id
__global__ void k_emit(const int32_t *row, const int32_t *col, const float *val,
                       const uint32_t *idx, const int32_t *kept, int64_t k, int64_t nnz_in,
                       int32_t *ocol, float *oval, int32_t *orow) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t a = idx[kept ? kept[i] : i];
        ocol[i] = col[a];
        if (oval) oval[i] = val ? val[a < nnz_in ? a : a - nnz_in] : 1.0f;
        orow[i] = row[a];
    }
}

// rowptr[r] = number of entries with row < r, by binary search over the
// sorted row ids (no per-entry atomics: a hub row would serialise them)
__global__ void k_rowptr_search(const int32_t *orow, int64_t k, int32_t m, int32_t *rowptr) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= m;
         r += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = k;  // first i with orow[i] >= r
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (orow[mid] < r) lo = mid + 1; else hi = mid;
        }
        rowptr[r] = (int32_t)lo;
    }
}

// row id of every CSR entry (binary search over rowptr)
__global__ void k_entry_rows(const int32_t *rowptr, int32_t m, int64_t nnz, int32_t *rows) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = m;  // largest r with rowptr[r] <= e
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (rowptr[mid] <= e) lo = mid; else hi = mid;
        }
        rows[e] = lo;
    }
}

void finish_rowptr(Ctx *c, Graph *g, const int32_t *orow) {
    const int grid = fill_grid(c, kBlock, 4);
    k_rowptr_search<<<grid, kBlock, 0, c->stream>>>(orow, g->nnz, g->m, g->rowptr);
    APSPIE_LAUNCH_CHECK();
}


// ---------------------------------------------------- row-block DCSC
// Part of CSR row r when the rows are cut into blocks of ps, the last of the
// pn parts taking every row from (pn - 1) * ps on.
__device__ __forceinline__ uint32_t dcsc_part(int32_t r, int32_t ps, int32_t pn) {
    return (uint32_t)min(r / ps, pn - 1);
}

__global__ void k_dcsc_part_keys(const int32_t *rows, const uint32_t *idx, int64_t n, int32_t ps,
                                 int32_t pn, uint32_t *keys) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        keys[i] = dcsc_part(rows[idx[i]], ps, pn);
}

// Entry j of the (part, column, row) order: its row and value, and whether
// it opens a (part, column) segment.
__global__ void k_dcsc_entries(const int32_t *rows, const int32_t *col, const float *val,
                               const uint32_t *part, const uint32_t *idx, int64_t n,
                               int32_t *row_ind, float *val_out, int32_t *opens) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t e = idx[j];
        row_ind[j] = rows[e];
        if (val_out) val_out[j] = val[e];
        opens[j] = j == 0 || part[j] != part[j - 1] || col[e] != col[idx[j - 1]];
    }
}

// Segment k starts at off[k]: its column and part; off[L] = n closes the last.
__global__ void k_dcsc_segments(const int32_t *col, const uint32_t *idx, const uint32_t *part,
                                int32_t *off, int64_t L, int64_t n, int32_t *col_ind,
                                uint32_t *seg_part) {
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (int64_t k = gid; k < L; k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t j = off[k];
        col_ind[k] = col[idx[j]];
        seg_part[k] = part[j];
    }
    if (gid == 0) off[L] = (int32_t)n;
}

// part_ptr[p] = segments in parts before p (first k with seg_part[k] >= p)
__global__ void k_dcsc_part_ptr(const uint32_t *seg_part, int64_t L, int32_t pn,
                                int32_t *part_ptr) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p <= pn;
         p += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = L;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((int64_t)seg_part[mid] < p) lo = mid + 1; else hi = mid;
        }
        part_ptr[p] = (int32_t)lo;
    }
}

}  // namespace

void graph_from_coo_dev(Ctx *c, Graph *g, const int32_t *row, const int32_t *col,
                        const float *val, int64_t nnz, int symmetrize) {
    const int grid = fill_grid(c, kBlock, 4);
    const int64_t E = symmetrize ? 2 * nnz : nnz;
    const int64_t Ea = std::max<int64_t>(E, 1);
    int32_t *r2 = (int32_t *)scratch(c, "ing_r", sizeof(int32_t) * Ea);
    int32_t *c2 = (int32_t *)scratch(c, "ing_c", sizeof(int32_t) * Ea);
    uint32_t *keys = (uint32_t *)scratch(c, "ing_k", sizeof(uint32_t) * Ea);
    uint32_t *idx = (uint32_t *)scratch(c, "ing_i", sizeof(uint32_t) * Ea);
    int32_t *flag = (int32_t *)scratch(c, "ing_flag", sizeof(int32_t) * Ea);
    int32_t *kept = (int32_t *)scratch(c, "ing_kept", sizeof(int32_t) * Ea);
    int32_t *bad = (int32_t *)scratch(c, "ing_bad", 256);
    fill_i32(c, bad, 0, 1);
    if (nnz > 0) {
        if (symmetrize)
            k_mirror<<<grid, kBlock, 0, c->stream>>>(row, col, nnz, r2, c2, g->m, g->n, bad);
        else
            k_copy_check<<<grid, kBlock, 0, c->stream>>>(row, col, nnz, r2, c2, g->m, g->n, bad);
        APSPIE_LAUNCH_CHECK();
    }
    int32_t hbad = 0;
    APSPIE_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    APSPIE_CHECK(hbad == 0, APSPIE_ERR_INVALID, "COO index out of range");
    int64_t k = E;
    if (E > 0) {
        if (symmetrize) {  // order by (row, col): sort by col, then stably by row
            k_iota_key<<<grid, kBlock, 0, c->stream>>>(c2, E, keys, idx);
            radix_sort_pairs(c, keys, idx, E, bits_for(g->n));
            k_gather_key<<<grid, kBlock, 0, c->stream>>>(r2, idx, E, keys);
        } else {           // buildMatrix: by row only, input order inside a row
            k_iota_key<<<grid, kBlock, 0, c->stream>>>(r2, E, keys, idx);
        }
        radix_sort_pairs(c, keys, idx, E, bits_for(g->m));
        k_keep<<<grid, kBlock, 0, c->stream>>>(r2, c2, idx, E, symmetrize, flag);
        APSPIE_LAUNCH_CHECK();
        k = symmetrize ? compact_flags(c, flag, E, kept) : E;
    }
    g->nnz = k;
    g->rowptr = (int32_t *)dev_alloc(c, sizeof(int32_t) * (g->m + 1));
    g->col = (int32_t *)dev_alloc(c, sizeof(int32_t) * std::max<int64_t>(k, 1));
    g->val = (float *)dev_alloc(c, sizeof(float) * std::max<int64_t>(k, 1));
    g->bytes = (int64_t)(sizeof(int32_t) * (g->m + 1) + 8 * std::max<int64_t>(k, 1));
    int32_t *orow = (int32_t *)scratch(c, "ing_orow", sizeof(int32_t) * std::max<int64_t>(k, 1));
    if (k > 0) {
        k_emit<<<grid, kBlock, 0, c->stream>>>(r2, c2, val, idx, symmetrize ? kept : nullptr, k,
                                               nnz, g->col, g->val, orow);
        APSPIE_LAUNCH_CHECK();
    }
    finish_rowptr(c, g, orow);
}

__global__ void k_neq(const int32_t *a, const int32_t *b, int64_t n, int32_t *flag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (a[i] != b[i]) flag[0] = 1;
}

// Structure (and, for a weighted graph, values) compared with the transpose,
// once per graph.  The transpose's rows come out sorted, so with equal
// structure the values line up position by position.
static void check_symmetry(Ctx *c, Graph *g) {
    if (g->symmetric >= 0) return;
    if (g->m != g->n) {
        g->symmetric = 0;
        g->sym_weights = 0;
        return;
    }
    Graph t;  // a temporary in the context's scratch
    graph_transpose_dev(c, g, &t, "sym_t");
    int32_t *flag = (int32_t *)scratch(c, "sym_flag", 256);
    fill_i32(c, flag, 0, 2);
    const int grid = fill_grid(c, kBlock, 4);
    k_neq<<<grid, kBlock, 0, c->stream>>>(g->rowptr, t.rowptr, g->m + 1, flag);
    k_neq<<<grid, kBlock, 0, c->stream>>>(g->col, t.col, g->nnz, flag);
    if (g->val && t.val)
        k_neq<<<grid, kBlock, 0, c->stream>>>((const int32_t *)g->val, (const int32_t *)t.val,
                                               g->nnz, flag + 1);
    APSPIE_LAUNCH_CHECK();
    int32_t h[2] = {1, 1};
    APSPIE_HIP(hipMemcpyAsync(h, flag, 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    g->symmetric = h[0] == 0 ? 1 : 0;
    g->sym_weights = (g->symmetric == 1 && g->val && h[1] == 0) ? 1 : 0;
}

__global__ void k_iso_bits(const int32_t *rowptr, int32_t m, uint32_t *bits) {
    const int64_t words = ((int64_t)m + 31) / 32;
    for (int64_t wi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; wi < words;
         wi += (int64_t)gridDim.x * blockDim.x) {
        uint32_t x = 0;
        int32_t prev = rowptr[wi * 32];
        for (int b = 0; b < 32 && wi * 32 + b < m; ++b) {
            const int32_t next = rowptr[wi * 32 + b + 1];
            if (next == prev) x |= 1u << b;
            prev = next;
        }
        bits[wi] = x;
    }
}

__global__ void k_max_degree(const int32_t *rowptr, int32_t m, int32_t *out) {
    int32_t best = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        best = max(best, rowptr[i + 1] - rowptr[i]);
    for (int off = 32; off > 0; off >>= 1) best = max(best, __shfl_xor(best, off, 64));
    if (dev::lane_id() == 0 && best > 0) atomicMax(out, best);
}

// Floats as unsigned keys that order like the floats (negative ones reversed).
__device__ __forceinline__ uint32_t f_order(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void k_weight_range(const float *w, int64_t n, uint32_t *out) {
    uint32_t lo = 0xffffffffu, hi = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t k = f_order(w[i]);
        lo = min(lo, k);
        hi = max(hi, k);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, off, 64));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, off, 64));
    }
    if (dev::lane_id() == 0) {
        atomicMin(out, lo);
        atomicMax(out + 1, hi);
    }
}

bool graph_weight_range(Ctx *c, Graph *g, float *lo, float *hi) {
    if (!g->val || g->nnz == 0) return false;
    std::lock_guard<std::mutex> lock(g->derived_mu);
    if (g->w_range < 0) {
        uint32_t *d = (uint32_t *)scratch(c, "w_range", 256);
        fill_i32(c, (int32_t *)d, -1, 1);
        fill_i32(c, (int32_t *)d + 1, 0, 1);
        k_weight_range<<<(int)std::min<int64_t>(ceil_div(g->nnz, kBlock), 4 * c->cus), kBlock, 0,
                         c->stream>>>(g->val, g->nnz, d);
        APSPIE_LAUNCH_CHECK();
        uint32_t h[2];
        APSPIE_HIP(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        auto back = [](uint32_t k) {
            const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
            return __builtin_bit_cast(float, b);
        };
        g->w_lo = back(h[0]);
        g->w_hi = back(h[1]);
        g->w_range = 1;
    }
    *lo = g->w_lo;
    *hi = g->w_hi;
    return true;
}

// Longest row, computed once per graph.
int32_t graph_max_degree(Ctx *c, Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    if (g->max_deg >= 0) return g->max_deg;
    int32_t *d = (int32_t *)scratch(c, "max_deg", 256);
    fill_i32(c, d, 0, 1);
    if (g->m > 0) {
        k_max_degree<<<(int)std::min<int64_t>(ceil_div(g->m, kBlock), 4 * c->cus), kBlock, 0,
                       c->stream>>>(g->rowptr, g->m, d);
        APSPIE_LAUNCH_CHECK();
    }
    int32_t h = 0;
    APSPIE_HIP(hipMemcpyAsync(&h, d, 4, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    g->max_deg = h;
    return h;
}

// Bitmap of the vertices without entries, built once per graph.
const uint32_t *graph_iso_bits(Ctx *c, Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    if (g->iso_bits) return g->iso_bits;
    const int64_t words = ((int64_t)g->m + 31) / 32;
    uint32_t *bits = nullptr;
    // owned by the graph (freed with it), like its other derived arrays
    APSPIE_HIP(hipMalloc(&bits, sizeof(uint32_t) * std::max<int64_t>(words, 1)));
    if (words > 0) {
        k_iso_bits<<<(int)std::min<int64_t>(ceil_div(words, kBlock), 4 * c->cus), kBlock, 0,
                     c->stream>>>(g->rowptr, g->m, bits);
        APSPIE_LAUNCH_CHECK();
    }
    sync(c);
    g->iso_bits = bits;
    return bits;
}

// ------------------------------------------------- connected components
// Lock-free union-find over the edges of a symmetric CSR: every root hooks
// under a smaller root with one compare-and-swap on its parent pointer, so a
// tree's root is always the smallest id in it and the final label of a
// vertex is the smallest vertex id of its component.  Parents only ever move
// to smaller ids of the same set, so any value a parent pointer ever held is
// an ancestor: a find may halve its path, and a stale read only costs a
// retry.  Every write to the parent array while unions run is an atomic
// read-modify-write (the hook, the halving swap, the final label exchange):
// the workgroups of different XCDs write neighbouring words of one line.
__device__ __forceinline__ int32_t cc_parent(const int32_t *p, int32_t x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t cc_find(int32_t *p, int32_t x) {
    int32_t px = cc_parent(p, x);
    while (px != x) {
        const int32_t gp = cc_parent(p, px);
        if (gp != px) {  // halve: x's parent becomes its grandparent, if still px
            int32_t expected = px;
            __hip_atomic_compare_exchange_strong(p + x, &expected, gp, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        x = gp;
        px = cc_parent(p, x);
    }
    return x;
}

__device__ __forceinline__ void cc_unite(int32_t *p, int32_t u, int32_t v) {
    u = cc_find(p, u);
    v = cc_find(p, v);
    while (u != v) {
        if (u < v) {
            const int32_t t = u;
            u = v;
            v = t;
        }
        // hook the larger root u under v; a failed swap means u got a parent
        int32_t expected = u;
        if (__hip_atomic_compare_exchange_strong(p + u, &expected, v, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return;
        u = cc_find(p, expected);
        v = cc_find(p, v);
    }
}

__global__ void k_cc_init(int32_t m, int32_t *p) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int32_t)i;
}

constexpr int32_t kCcShort = 32;

// rows of at most kCcShort entries: one thread per row; each undirected edge
// is united from its larger end only
__global__ void k_cc_short(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                           int32_t m, int32_t *p) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < m;
         u += (int64_t)gridDim.x * blockDim.x) {
        const int32_t rs = rowptr[u], re = rowptr[u + 1];
        if (re - rs > kCcShort) continue;
        for (int32_t e = rs; e < re; ++e) {
            const int32_t v = col[e];
            if (v < u) cc_unite(p, (int32_t)u, v);
        }
    }
}

// longer rows: one wave per row, lanes striding it
__global__ void k_cc_long(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                          int32_t m, int32_t *p) {
    const int lane = dev::lane_id();
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
    const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / 64;
    for (int64_t u = wave; u < m; u += nwaves) {
        const int32_t rs = rowptr[u], re = rowptr[u + 1];
        if (re - rs <= kCcShort) continue;
        for (int32_t e = rs + lane; e < re; e += 64) {
            const int32_t v = col[e];
            if (v < u) cc_unite(p, (int32_t)u, v);
        }
    }
}

__global__ void k_cc_label(int32_t m, int32_t *p) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = cc_find(p, (int32_t)i);
        __hip_atomic_exchange(p + i, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

void components_into(Ctx *c, const Graph *g, int32_t *label) {
    if (g->m == 0) return;
    const int grid = fill_grid(c, kBlock, 8);
    k_cc_init<<<grid, kBlock, 0, c->stream>>>(g->m, label);
    k_cc_short<<<grid, kBlock, 0, c->stream>>>(g->rowptr, g->col, g->m, label);
    k_cc_long<<<grid, kBlock, 0, c->stream>>>(g->rowptr, g->col, g->m, label);
    k_cc_label<<<grid, kBlock, 0, c->stream>>>(g->m, label);
    APSPIE_LAUNCH_CHECK();
}

const int32_t *graph_components(Ctx *c, Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    if (g->comp) return g->comp;
    int32_t *label = nullptr;
    // owned by the graph (freed with it), like its other derived arrays
    APSPIE_HIP(hipMalloc(&label, sizeof(int32_t) * std::max<int64_t>(g->m, 1)));
    components_into(c, g, label);
    sync(c);
    g->comp = label;
    return label;
}

int graph_symmetry_known(Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    return g->symmetric;
}

const uint32_t *graph_iso_bits_cached(Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    return g->iso_bits;
}

void iso_bits_into(Ctx *c, const Graph *g, uint32_t *bits) {
    const int64_t words = ((int64_t)g->m + 31) / 32;
    if (words == 0) return;
    k_iso_bits<<<(int)std::min<int64_t>(ceil_div(words, kBlock), 4 * c->cus), kBlock, 0,
                 c->stream>>>(g->rowptr, g->m, bits);
    APSPIE_LAUNCH_CHECK();
}

bool graph_is_symmetric(Ctx *c, Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    check_symmetry(c, g);
    return g->symmetric == 1;
}

bool graph_is_symmetric_weighted(Ctx *c, Graph *g) {
    std::lock_guard<std::mutex> lock(g->derived_mu);
    check_symmetry(c, g);
    return g->sym_weights == 1;
}

void graph_transpose_dev(Ctx *c, const Graph *g, Graph *t, const char *in_scratch) {
    const int grid = fill_grid(c, kBlock, 4);
    const int64_t E = g->nnz, Ea = std::max<int64_t>(E, 1);
    t->owner = c;
    t->device = c->device;
    t->m = g->n;
    t->n = g->m;
    t->nnz = E;
    t->owns = !in_scratch;
    int32_t *rows = (int32_t *)scratch(c, "tr_rows", sizeof(int32_t) * Ea);
    uint32_t *keys = (uint32_t *)scratch(c, "tr_k", sizeof(uint32_t) * Ea);
    uint32_t *idx = (uint32_t *)scratch(c, "tr_i", sizeof(uint32_t) * Ea);
    if (E > 0) {
        k_entry_rows<<<grid, kBlock, 0, c->stream>>>(g->rowptr, g->m, E, rows);
        k_iota_key<<<grid, kBlock, 0, c->stream>>>(g->col, E, keys, idx);
        APSPIE_LAUNCH_CHECK();
        radix_sort_pairs(c, keys, idx, E, bits_for(g->n));  // stable: rows stay ascending
    }
    if (in_scratch) {
        const std::string p(in_scratch);
        t->rowptr = (int32_t *)scratch(c, p + "_rowptr", sizeof(int32_t) * (t->m + 1));
        t->col = (int32_t *)scratch(c, p + "_col", sizeof(int32_t) * Ea);
        t->val = g->val ? (float *)scratch(c, p + "_val", sizeof(float) * Ea) : nullptr;
    } else {
        t->rowptr = (int32_t *)dev_alloc(c, sizeof(int32_t) * (t->m + 1));
        t->col = (int32_t *)dev_alloc(c, sizeof(int32_t) * Ea);
        t->val = g->val ? (float *)dev_alloc(c, sizeof(float) * Ea) : nullptr;
    }
    t->bytes = (int64_t)(sizeof(int32_t) * (t->m + 1) + (g->val ? 8 : 4) * Ea);
    int32_t *orow = (int32_t *)scratch(c, "tr_orow", sizeof(int32_t) * Ea);
    if (E > 0) {
        // new row = old column, new column = old row
        k_emit<<<grid, kBlock, 0, c->stream>>>(g->col, rows, g->val, idx, nullptr, E, E, t->col,
                                               t->val, orow);
        APSPIE_LAUNCH_CHECK();
    }
    finish_rowptr(c, t, orow);
}


// Row-block DCSC (csr_to_dcsc, matrix.cpp:286-468): two stable radix passes
// order the entries by (part, column) with rows ascending inside, the
// segment openings are compacted into the column offsets, and each part's
// first segment is found by binary search.  Outputs are device pointers;
// returns the number of (part, column) segments.
int64_t graph_to_dcsc_dev(Ctx *c, const Graph *g, int32_t ps, int32_t pn, int32_t *part_ptr,
                          int32_t *col_ind, int32_t *col_off, int32_t *row_ind, float *val) {
    const int grid = fill_grid(c, kBlock, 4);
    const int64_t E = g->nnz, Ea = std::max<int64_t>(E, 1);
    int64_t L = 0;
    uint32_t *seg_part = (uint32_t *)scratch(c, "dc_sp", sizeof(uint32_t) * Ea);
    if (E > 0) {
        int32_t *rows = (int32_t *)scratch(c, "dc_rows", sizeof(int32_t) * Ea);
        uint32_t *keys = (uint32_t *)scratch(c, "dc_k", sizeof(uint32_t) * Ea);
        uint32_t *idx = (uint32_t *)scratch(c, "dc_i", sizeof(uint32_t) * Ea);
        int32_t *opens = (int32_t *)scratch(c, "dc_open", sizeof(int32_t) * Ea);
        k_entry_rows<<<grid, kBlock, 0, c->stream>>>(g->rowptr, g->m, E, rows);
        k_iota_key<<<grid, kBlock, 0, c->stream>>>(g->col, E, keys, idx);
        APSPIE_LAUNCH_CHECK();
        radix_sort_pairs(c, keys, idx, E, bits_for(g->n));  // by column, rows kept ascending
        k_dcsc_part_keys<<<grid, kBlock, 0, c->stream>>>(rows, idx, E, ps, pn, keys);
        APSPIE_LAUNCH_CHECK();
        radix_sort_pairs(c, keys, idx, E, bits_for(pn - 1));  // by part, stable
        k_dcsc_entries<<<grid, kBlock, 0, c->stream>>>(rows, g->col, g->val, keys, idx, E,
                                                       row_ind, g->val ? val : nullptr, opens);
        APSPIE_LAUNCH_CHECK();
        L = compact_flags(c, opens, E, col_off);
        k_dcsc_segments<<<grid, kBlock, 0, c->stream>>>(g->col, idx, keys, col_off, L, E,
                                                        col_ind, seg_part);
    } else {
        k_dcsc_segments<<<1, 64, 0, c->stream>>>(g->col, nullptr, nullptr, col_off, 0, 0,
                                                 col_ind, seg_part);
    }
    APSPIE_LAUNCH_CHECK();
    k_dcsc_part_ptr<<<std::max(1, std::min(grid, (pn + kBlock) / kBlock)), kBlock, 0,
                      c->stream>>>(seg_part, L, pn, part_ptr);
    APSPIE_LAUNCH_CHECK();
    return L;
}

}  // namespace apspie
