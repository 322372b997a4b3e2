// Context, graph residency and the C ABI glue.
#include <algorithm>
#include <cstdlib>
#include <utility>

#include "common.hpp"
#include "device.hpp"

namespace apspie {

static const char *kOptionNames[] = {
    "bfs_tile",          // edges per expansion tile (power of two)
    "bfs_blocks_per_cu", // resident 256-thread groups per CU for the level kernels
    "bfs_levels_per_sync", // levels launched between host checks
    "bfs_direction",     // 0 push only, 1 direction-optimising
    "bfs_alpha",         // push->pull switch: frontier edges * alpha > unexplored edges
    "bfs_beta",          // pull->push switch: frontier vertices * beta < m
    "sssp_delta",        // near/far split width (0 = plain Bellman-Ford frontier,
                         // -1 = chosen per graph from its weights and mean row length)
    "sssp_pull_div",     // pull an SSSP round when frontier edges * this > nnz (0 = never;
                         // symmetric weights only)
    "sssp_light_heavy",  // near-far: heavy entries (value >= width) once per bucket, at
                         // the next split, instead of with every improvement (1 = on)
    "spmspv_path",       // 0 auto, 1 sort, 2 atomic
    "spmspv_sort_max",   // auto: use the sort path up to this many expanded edges
    "batch_blocks_per_cu",
    "batch_graph",       // 1 = replay batched level chunks from a captured hipGraph
    "batch_pull_div",    // batched BFS pulls a level when union-frontier edges * this > nnz
                         // (times 4 for batches of up to 128 sources on graphs whose rows
                         // are at most 64 long; 0 = never)
    "batch_sssp_delta",  // batched SSSP near-far threshold step (0 = Bellman-Ford frontier,
                         // -1 = chosen per graph from its weights and mean row length)
    "batch_snapshot_mb", // HBM for the bit rows of pulled batched BFS levels (0 = none)
    "batch_sssp_pull_div", // batched near-far SSSP pulls a round when union-frontier edges *
                           // this > nnz, on graphs with rows of at most 64 entries and
                           // symmetric weights (0 = never)
    "batch_long_row_group", // batched BFS on graphs with rows longer than 64: sources per
                            // pass (1..512)
    "tc_medium_row",     // TC: rows with |L_r| up to this go to the 256-thread LDS tier, longer
                         // ones to the 1024-thread tier (128..2048; 512 measured best or within
                         // 4 % on R-MAT-18/20/22: 1.2-1.6x faster than 2048)
    "spgemm_wave_ub",    // SpGEMM: rows with at most this many products to the wave tier (<= 256)
    "spgemm_block_ub",   // SpGEMM: up to this to the block LDS tier (<= 4096)
    "spgemm_dense",      // SpGEMM: longer rows to the dense accumulator when B has <= 32768
                         // columns (1) or always to global tables (0)
    "profiling",         // 1 = time the dominant kernel with hipEvents
    nullptr};

static bool option_default(const std::string &n, int64_t *v) {
    static const std::pair<const char *, int64_t> kDefaults[] = {
        {"bfs_tile", 2048},
        {"bfs_blocks_per_cu", 8},
        {"bfs_levels_per_sync", 8},
        {"bfs_direction", 1},
        {"bfs_alpha", 14},
        {"bfs_beta", 24},
        {"sssp_delta", -1},
        {"sssp_pull_div", 0},
        {"sssp_light_heavy", 1},
        {"spmspv_path", 0},
        {"spmspv_sort_max", 1 << 16},
        {"batch_blocks_per_cu", 4},
        {"batch_graph", 1},
        {"batch_pull_div", 4},
        {"batch_sssp_delta", -1},
        {"batch_snapshot_mb", 24576},
        {"batch_long_row_group", 256},
        {"batch_sssp_pull_div", 4},
        {"tc_medium_row", 512},
        {"spgemm_wave_ub", 256},
        {"spgemm_block_ub", 4096},
        {"spgemm_dense", 1},
        {"profiling", 0}};
    for (const auto &d : kDefaults)
        if (n == d.first) {
            *v = d.second;
            return true;
        }
    return false;
}

static bool option_known(const std::string &n) {
    for (int i = 0; kOptionNames[i]; ++i)
        if (n == kOptionNames[i]) return true;
    return false;
}

int64_t opt(Ctx *c, const char *name, int64_t dflt) {
    auto it = c->options.find(name);
    if (it != c->options.end()) return it->second;
    int64_t d = 0;
    return option_default(name, &d) ? d : dflt;
}

void set_device(Ctx *c) { APSPIE_HIP(hipSetDevice(c->device)); }

void sync(Ctx *c) { APSPIE_HIP(hipStreamSynchronize(c->stream)); }

void *dev_alloc(Ctx *c, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (c->budget > 0 && c->used + (int64_t)bytes > c->budget)
        throw Error{APSPIE_ERR_BUDGET, "HBM budget exceeded: need " + std::to_string(bytes) +
                                           " more bytes, " + std::to_string(c->budget - c->used) +
                                           " left"};
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{APSPIE_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + ") failed"};
    }
    c->used += (int64_t)bytes;
    return p;
}

void dev_free(Ctx *c, void *p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    if (bytes == 0) bytes = 256;
    c->used -= (int64_t)bytes;
}

namespace {

constexpr size_t kGranule = 256;

size_t granules(size_t bytes) { return (std::max<size_t>(bytes, 1) + kGranule - 1) / kGranule * kGranule; }

// first free range of the arena that holds `bytes` (a granule multiple)
void *arena_take(Ctx *c, size_t bytes) {
    for (auto it = c->arena_free.begin(); it != c->arena_free.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, len = it->second;
        c->arena_free.erase(it);
        if (len > bytes) c->arena_free.emplace(off + bytes, len - bytes);
        return c->arena + off;
    }
    return nullptr;
}

// return a range, merged with the free ranges on either side
void arena_give(Ctx *c, void *p, size_t bytes) {
    auto it = c->arena_free.emplace((size_t)((char *)p - c->arena), bytes).first;
    auto nx = std::next(it);
    if (nx != c->arena_free.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        c->arena_free.erase(nx);
    }
    if (it != c->arena_free.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            c->arena_free.erase(it);
        }
    }
}

// drop every scratch buffer; the caller has synchronised the stream
void drop_scratch(Ctx *c) {
    if (c->arena) {
        c->arena_free.clear();
        c->arena_free.emplace(0, c->arena_bytes);
    } else {
        for (auto &kv : c->scratch) dev_free(c, kv.second.ptr, kv.second.bytes);
    }
    c->scratch.clear();
}

}  // namespace

void *scratch(Ctx *c, const std::string &name, size_t bytes) {
    Buffer &b = c->scratch[name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (b.ptr) {
        // pending work may still read the old buffer
        sync(c);
        if (c->arena)
            arena_give(c, b.ptr, b.bytes);
        else
            dev_free(c, b.ptr, b.bytes);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    if (c->arena) {
        const size_t want = granules(bytes);
        b.ptr = arena_take(c, want);
        if (!b.ptr)
            throw Error{APSPIE_ERR_BUDGET, "scratch arena full: need " + std::to_string(want) +
                                               " bytes in one range, the largest free range is " +
                                               std::to_string(scratch_room(c)) +
                                               " (apspie_ctx_trim drops the cached buffers)"};
        b.bytes = want;
        return b.ptr;
    }
    size_t want = std::max<size_t>(bytes, 256);
    b.ptr = dev_alloc(c, want);
    b.bytes = want;
    return b.ptr;
}

int64_t scratch_room(Ctx *c) {
    if (c->arena) {
        size_t best = 0;
        for (const auto &kv : c->arena_free) best = std::max(best, kv.second);
        return (int64_t)best;
    }
    if (c->budget > 0) return std::max<int64_t>(0, c->budget - c->used);
    return INT64_MAX;
}

void copy_to_device(Ctx *c, void *dst, const void *src, size_t bytes, int mem) {
    if (!bytes) return;
    APSPIE_HIP(hipMemcpyAsync(dst, src, bytes,
                              mem == APSPIE_MEM_HOST ? hipMemcpyHostToDevice
                                                     : hipMemcpyDeviceToDevice,
                              c->stream));
}

void copy_from_device(Ctx *c, void *dst, const void *src, size_t bytes, int mem) {
    if (!bytes) return;
    APSPIE_HIP(hipMemcpyAsync(dst, src, bytes,
                              mem == APSPIE_MEM_HOST ? hipMemcpyDeviceToHost
                                                     : hipMemcpyDeviceToDevice,
                              c->stream));
}

int fill_grid(Ctx *c, int threads, int per_cu) {
    (void)threads;
    return c->cus * std::max(1, per_cu);
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_i32(Ctx *c, int32_t *p, int32_t v, int64_t n) {
    if (n <= 0) return;
    int blocks = (int)std::min<int64_t>(ceil_div(n, 256), 4 * c->cus);
    k_fill_i32<<<blocks, 256, 0, c->stream>>>(p, v, n);
    APSPIE_LAUNCH_CHECK();
}

static void validate_csr_host(const apspie_csr *h) {
    APSPIE_CHECK(h && h->rowptr && (h->col || h->nnz == 0), APSPIE_ERR_INVALID, "null CSR");
    APSPIE_CHECK(h->m >= 0 && h->n >= 0 && h->nnz >= 0, APSPIE_ERR_INVALID, "negative size");
    APSPIE_CHECK(h->nnz < INT32_MAX, APSPIE_ERR_UNSUPPORTED, "nnz must fit int32");
    APSPIE_CHECK(h->rowptr[0] == 0 && h->rowptr[h->m] == h->nnz, APSPIE_ERR_INVALID,
                 "rowptr does not span [0, nnz]");
    for (int32_t i = 0; i < h->m; ++i)
        APSPIE_CHECK(h->rowptr[i] <= h->rowptr[i + 1], APSPIE_ERR_INVALID, "rowptr not monotone");
    for (int64_t e = 0; e < h->nnz; ++e)
        APSPIE_CHECK(h->col[e] >= 0 && h->col[e] < h->n, APSPIE_ERR_INVALID,
                     "column index out of range");
}

static void check_graph(Ctx *c, const Graph *g) {
    APSPIE_CHECK(g, APSPIE_ERR_INVALID, "null graph");
    APSPIE_CHECK(g->device == c->device, APSPIE_ERR_INVALID, "graph lives on another device");
}

// SSSP orders distances by their bit patterns (an int32 atomic min), which
// orders finite non-negative floats only; the reference's CPU SSSP
// (testSssp.cpp:86-100) would accept negative edges, which this path cannot
// order, so they are refused instead of answered wrongly.  The range is
// reduced on the device once per graph.
static void check_sssp_weights(Ctx *c, const Graph *g) {
    APSPIE_CHECK(g->val, APSPIE_ERR_INVALID, "SSSP needs edge weights");
    float lo = 0.0f, hi = 0.0f;
    if (!graph_weight_range(c, const_cast<Graph *>(g), &lo, &hi)) return;  // no entries
    APSPIE_CHECK(lo >= 0.0f && hi <= 3.402823466e+38f, APSPIE_ERR_UNSUPPORTED,
                 "SSSP needs finite non-negative weights (smallest " + std::to_string(lo) +
                     ", largest " + std::to_string(hi) + ")");
}

}  // namespace apspie

using namespace apspie;


extern "C" {

int apspie_abi_version(void) { return APSPIE_ABI_VERSION; }

const char *apspie_status_str(int s) {
    switch (s) {
    case APSPIE_OK: return "ok";
    case APSPIE_ERR_INVALID: return "invalid argument";
    case APSPIE_ERR_OOM: return "out of device memory";
    case APSPIE_ERR_DEVICE: return "device error";
    case APSPIE_ERR_BUDGET: return "HBM budget exceeded";
    case APSPIE_ERR_OVERFLOW: return "internal capacity exceeded";
    case APSPIE_ERR_UNSUPPORTED: return "unsupported";
    case APSPIE_ERR_BUSY: return "busy";
    default: return "unknown status";
    }
}

int apspie_ctx_create(int device, apspie_ctx **out) {
    if (!out) return APSPIE_ERR_INVALID;
    *out = nullptr;
    apspie_ctx *c = new apspie_ctx();
    c->device = device;
    int st = guard(nullptr, [&] {
        int n = 0;
        APSPIE_HIP(hipGetDeviceCount(&n));
        APSPIE_CHECK(device >= 0 && device < n, APSPIE_ERR_INVALID, "no such device");
        set_device(c);
        APSPIE_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
        c->stream = c->own_stream;
        hipDeviceProp_t prop;
        APSPIE_HIP(hipGetDeviceProperties(&prop, device));
        c->cus = prop.multiProcessorCount;
        APSPIE_HIP(hipEventCreate(&c->ev0));
        APSPIE_HIP(hipEventCreate(&c->ev1));
        APSPIE_HIP(hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault));
        noalloc_init(device);  // the no-alloc BFS's static array, once per device
    });
    if (st != APSPIE_OK) {
        delete c;
        return st;
    }
    *out = c;
    return APSPIE_OK;
}

int apspie_ctx_destroy(apspie_ctx *c) {
    if (!c) return APSPIE_ERR_INVALID;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->arena)
        (void)hipFree(c->arena);
    else
        for (auto &kv : c->scratch)
            if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (auto &p : c->timer_events) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->mis_prio.ptr) (void)hipFree(c->mis_prio.ptr);
    // graphs outliving their context keep their memory but lose the accounting
    for (Graph *g : c->graphs) g->owner = nullptr;
    delete c;
    return APSPIE_OK;
}

int apspie_ctx_set_stream(apspie_ctx *c, void *s) {
    if (!c) return APSPIE_ERR_INVALID;
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return APSPIE_OK;
}

void *apspie_ctx_stream(apspie_ctx *c) { return c ? (void *)c->stream : nullptr; }

int apspie_ctx_set_hbm_budget(apspie_ctx *c, int64_t bytes) {
    if (!c || bytes < 0) return APSPIE_ERR_INVALID;
    c->budget = bytes;
    return APSPIE_OK;
}

int64_t apspie_ctx_hbm_used(apspie_ctx *c) { return c ? c->used : -1; }

int apspie_ctx_set_option(apspie_ctx *c, const char *name, int64_t v) {
    if (!c || !name) return APSPIE_ERR_INVALID;
    if (!option_known(name)) {
        c->last_error = std::string("unknown option ") + name;
        return APSPIE_ERR_INVALID;
    }
    std::string n(name);
    if (n == "bfs_tile" && v != 1024 && v != 2048 && v != 4096) return APSPIE_ERR_INVALID;
    if ((n == "bfs_blocks_per_cu" || n == "batch_blocks_per_cu") && (v < 1 || v > 32))
        return APSPIE_ERR_INVALID;
    if (n == "bfs_levels_per_sync" && (v < 1 || v > 4096)) return APSPIE_ERR_INVALID;
    if (n == "batch_pull_div" && (v < 0 || v > (1 << 20))) return APSPIE_ERR_INVALID;
    if (n == "batch_snapshot_mb" && (v < 0 || v > (1 << 22))) return APSPIE_ERR_INVALID;
    if (n == "batch_long_row_group" && (v < 1 || v > 512)) return APSPIE_ERR_INVALID;
    if (n == "batch_sssp_pull_div" && (v < 0 || v > (1 << 20))) return APSPIE_ERR_INVALID;
    if (n == "tc_medium_row" && (v < 128 || v > 2048)) return APSPIE_ERR_INVALID;
    if (n == "spgemm_wave_ub" && (v < 1 || v > 256)) return APSPIE_ERR_INVALID;
    if (n == "spgemm_block_ub" && (v < 1 || v > 4096)) return APSPIE_ERR_INVALID;
    if (n == "spgemm_dense" && (v < 0 || v > 1)) return APSPIE_ERR_INVALID;
    if (n == "batch_sssp_delta" && (v < -1 || v > (1 << 30))) return APSPIE_ERR_INVALID;
    if (n == "sssp_delta" && (v < -1 || v > (1 << 30))) return APSPIE_ERR_INVALID;
    if (n == "sssp_pull_div" && (v < 0 || v > (1 << 20))) return APSPIE_ERR_INVALID;
    if (n == "sssp_light_heavy" && (v < 0 || v > 1)) return APSPIE_ERR_INVALID;
    if (n == "profiling") c->profiling = v != 0;
    c->options[n] = v;
    return APSPIE_OK;
}

int apspie_ctx_get_option(apspie_ctx *c, const char *name, int64_t *v) {
    if (!c || !name || !v) return APSPIE_ERR_INVALID;
    if (!option_known(name)) return APSPIE_ERR_INVALID;
    *v = opt(c, name, 0);
    return APSPIE_OK;
}

int apspie_ctx_trim(apspie_ctx *c) {
    if (!c) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        sync(c);
        drop_scratch(c);
    });
}

int apspie_ctx_reserve_arena(apspie_ctx *c, int64_t bytes) {
    if (!c || bytes < 0) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        sync(c);
        drop_scratch(c);
        if (c->arena) {
            dev_free(c, c->arena, c->arena_bytes);
            c->arena = nullptr;
            c->arena_bytes = 0;
            c->arena_free.clear();
        }
        if (bytes == 0) return;
        // whole granules, rounded down: a budget-sized arena stays within the budget
        const size_t want = (size_t)bytes / kGranule * kGranule;
        APSPIE_CHECK(want > 0, APSPIE_ERR_INVALID, "arena smaller than 256 bytes");
        c->arena = (char *)dev_alloc(c, want);  // within the HBM budget
        c->arena_bytes = want;
        c->arena_free.emplace(0, want);
    });
}

int64_t apspie_ctx_arena_bytes(apspie_ctx *c) { return c ? (int64_t)c->arena_bytes : -1; }

const char *apspie_last_error(apspie_ctx *c) { return c ? c->last_error.c_str() : "no context"; }

int apspie_dev_alloc(apspie_ctx *c, int64_t bytes, void **ptr) {
    if (!c || !ptr || bytes < 0) return APSPIE_ERR_INVALID;
    *ptr = nullptr;
    return guard(c, [&] {
        set_device(c);
        *ptr = dev_alloc(c, (size_t)bytes);
    });
}

int apspie_dev_free(apspie_ctx *c, void *ptr, int64_t bytes) {
    if (!c || bytes < 0) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        sync(c);
        dev_free(c, ptr, (size_t)bytes);
    });
}

int apspie_memcpy(apspie_ctx *c, void *dst, const void *src, int64_t bytes, int kind) {
    if (!c || bytes < 0 || (bytes && (!dst || !src)) || kind < 0 || kind > 2)
        return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        const hipMemcpyKind k = kind == 0   ? hipMemcpyHostToDevice
                                : kind == 1 ? hipMemcpyDeviceToHost
                                            : hipMemcpyDeviceToDevice;
        if (bytes) APSPIE_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, c->stream));
        sync(c);
    });
}

int apspie_ctx_sync(apspie_ctx *c) {
    if (!c) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        sync(c);
    });
}

int apspie_last_stats(apspie_ctx *c, apspie_stats *out) {
    if (!c || !out) return APSPIE_ERR_INVALID;
    *out = c->stats;
    return APSPIE_OK;
}

// ------------------------------------------------------------------- graphs
int apspie_graph_upload(apspie_ctx *c, const apspie_csr *h, apspie_graph **out) {
    if (!c || !out) return APSPIE_ERR_INVALID;
    *out = nullptr;
    apspie_graph *g = new apspie_graph();
    int st = guard(c, [&] {
        validate_csr_host(h);
        set_device(c);
        g->owner = c;
        g->device = c->device;
        g->m = h->m;
        g->n = h->n;
        g->nnz = h->nnz;
        g->owns = true;
        const bool has_val = h->val != nullptr;  // a weighted graph, even if edgeless
        size_t rb = sizeof(int32_t) * (h->m + 1), cb = sizeof(int32_t) * h->nnz,
               vb = has_val ? sizeof(float) * h->nnz : 0;
        g->rowptr = (int32_t *)dev_alloc(c, rb);
        g->col = (int32_t *)dev_alloc(c, cb);
        if (has_val) g->val = (float *)dev_alloc(c, vb);
        g->bytes = (int64_t)(std::max<size_t>(rb, 256) + std::max<size_t>(cb, 256) +
                             (has_val ? std::max<size_t>(vb, 256) : 0));
        copy_to_device(c, g->rowptr, h->rowptr, rb, APSPIE_MEM_HOST);
        copy_to_device(c, g->col, h->col, cb, APSPIE_MEM_HOST);
        if (vb) copy_to_device(c, g->val, h->val, vb, APSPIE_MEM_HOST);
        sync(c);
    });
    if (st != APSPIE_OK) {
        if (g->rowptr) dev_free(c, g->rowptr, sizeof(int32_t) * (h->m + 1));
        if (g->col) dev_free(c, g->col, sizeof(int32_t) * h->nnz);
        if (g->val) dev_free(c, g->val, sizeof(float) * h->nnz);
        delete g;
        return st;
    }
    c->graphs.insert(g);
    *out = g;
    return APSPIE_OK;
}

int apspie_graph_wrap(apspie_ctx *c, const apspie_csr *d, apspie_graph **out) {
    if (!c || !d || !out || !d->rowptr || (!d->col && d->nnz)) return APSPIE_ERR_INVALID;
    if (d->m < 0 || d->nnz < 0 || d->nnz >= INT32_MAX) return APSPIE_ERR_INVALID;
    apspie_graph *g = new apspie_graph();
    g->owner = c;
    g->device = c->device;
    g->m = d->m;
    g->n = d->n;
    g->nnz = d->nnz;
    g->rowptr = const_cast<int32_t *>(d->rowptr);
    g->col = const_cast<int32_t *>(d->col);
    g->val = const_cast<float *>(d->val);
    g->owns = false;
    c->graphs.insert(g);
    *out = g;
    return APSPIE_OK;
}

int apspie_graph_from_coo(apspie_ctx *c, const apspie_coo *coo, int mem, int symmetrize,
                          apspie_graph **out) {
    if (!c || !coo || !out || coo->nnz < 0 || coo->m < 0) return APSPIE_ERR_INVALID;
    if ((!coo->row || !coo->col) && coo->nnz) return APSPIE_ERR_INVALID;
    *out = nullptr;
    apspie_graph *g = new apspie_graph();
    int st = guard(c, [&] {
        set_device(c);
        APSPIE_CHECK(!symmetrize || coo->m == coo->n, APSPIE_ERR_INVALID,
                     "symmetrize needs a square matrix");
        APSPIE_CHECK(2 * coo->nnz < INT32_MAX, APSPIE_ERR_UNSUPPORTED, "nnz must fit int32");
        g->owner = c;
        g->device = c->device;
        g->m = coo->m;
        g->n = coo->n;
        g->owns = true;
        const int32_t *r = coo->row, *cl = coo->col;
        const float *v = coo->val;
        if (mem == APSPIE_MEM_HOST && coo->nnz) {
            int32_t *dr = (int32_t *)scratch(c, "coo_in_r", sizeof(int32_t) * coo->nnz);
            int32_t *dc = (int32_t *)scratch(c, "coo_in_c", sizeof(int32_t) * coo->nnz);
            copy_to_device(c, dr, r, sizeof(int32_t) * coo->nnz, APSPIE_MEM_HOST);
            copy_to_device(c, dc, cl, sizeof(int32_t) * coo->nnz, APSPIE_MEM_HOST);
            float *dv = nullptr;
            if (v) {
                dv = (float *)scratch(c, "coo_in_v", sizeof(float) * coo->nnz);
                copy_to_device(c, dv, v, sizeof(float) * coo->nnz, APSPIE_MEM_HOST);
            }
            r = dr; cl = dc; v = dv;
        }
        graph_from_coo_dev(c, g, r, cl, v, coo->nnz, symmetrize);
        if (symmetrize) g->symmetric = 1;
        sync(c);
    });
    if (st != APSPIE_OK) {
        delete g;
        return st;
    }
    c->graphs.insert(g);
    *out = g;
    return APSPIE_OK;
}

int apspie_graph_destroy(apspie_graph *g) {
    if (!g) return APSPIE_ERR_INVALID;
    Ctx *c = g->owner;
    (void)hipSetDevice(g->device);
    if (c) {
        (void)hipStreamSynchronize(c->stream);
        c->graphs.erase(g);
    }
    if (g->owns) {
        if (c) {
            dev_free(c, g->rowptr, sizeof(int32_t) * (g->m + 1));
            dev_free(c, g->col, sizeof(int32_t) * g->nnz);
            if (g->val) dev_free(c, g->val, sizeof(float) * g->nnz);
        } else {
            (void)hipFree(g->rowptr);
            (void)hipFree(g->col);
            if (g->val) (void)hipFree(g->val);
        }
    }
    if (g->lowcnt) (void)hipFree(g->lowcnt);
    if (g->iso_bits) (void)hipFree(g->iso_bits);
    if (g->comp) (void)hipFree(g->comp);
    if (g->heavy_mem) (void)hipFree(g->heavy_mem);
    for (const auto &x : g->lf) (void)hipFree(x.mem);
    delete g;
    return APSPIE_OK;
}

int apspie_graph_info(const apspie_graph *g, int32_t *m, int32_t *n, int64_t *nnz,
                      int64_t *bytes) {
    if (!g) return APSPIE_ERR_INVALID;
    if (m) *m = g->m;
    if (n) *n = g->n;
    if (nnz) *nnz = g->nnz;
    if (bytes) *bytes = g->bytes;
    return APSPIE_OK;
}

int apspie_graph_view(const apspie_graph *g, apspie_csr *v) {
    if (!g || !v) return APSPIE_ERR_INVALID;
    v->m = g->m;
    v->n = g->n;
    v->nnz = g->nnz;
    v->rowptr = g->rowptr;
    v->col = g->col;
    v->val = g->val;
    return APSPIE_OK;
}

int apspie_graph_download(apspie_ctx *c, const apspie_graph *g, int32_t *rowptr, int32_t *col,
                          float *val) {
    if (!c || !g) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        if (rowptr)
            copy_from_device(c, rowptr, g->rowptr, sizeof(int32_t) * (g->m + 1), APSPIE_MEM_HOST);
        if (col) copy_from_device(c, col, g->col, sizeof(int32_t) * g->nnz, APSPIE_MEM_HOST);
        if (val) {
            APSPIE_CHECK(g->val, APSPIE_ERR_INVALID, "graph has no values");
            copy_from_device(c, val, g->val, sizeof(float) * g->nnz, APSPIE_MEM_HOST);
        }
        sync(c);
    });
}

int apspie_graph_to_dcsc(apspie_ctx *c, const apspie_graph *g, int32_t part_size,
                         int32_t part_num, int64_t *col_length, int32_t *part_ptr,
                         int32_t *col_ind, int32_t *col_off, int32_t *row_ind, float *val,
                         int mem) {
    if (!c || !g || !col_length || !part_ptr || !col_off) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(part_size >= 1 && part_num >= 1, APSPIE_ERR_INVALID,
                     "part_size and part_num must be positive");
        APSPIE_CHECK((int64_t)(part_num - 1) * part_size <= g->m, APSPIE_ERR_INVALID,
                     "a part starts past the last row");
        APSPIE_CHECK(g->nnz == 0 || (col_ind && row_ind), APSPIE_ERR_INVALID,
                     "col_ind / row_ind are NULL");
        APSPIE_CHECK(g->nnz == 0 || !val || g->val, APSPIE_ERR_INVALID, "graph has no values");
        set_device(c);
        const int64_t E = std::max<int64_t>(g->nnz, 1);
        const bool host = mem == APSPIE_MEM_HOST;
        const bool want_val = val && g->val;
        int32_t *pp = host ? (int32_t *)scratch(c, "dco_pp", sizeof(int32_t) * (part_num + 1))
                           : part_ptr;
        int32_t *ci = host ? (int32_t *)scratch(c, "dco_ci", sizeof(int32_t) * E) : col_ind;
        int32_t *co = host ? (int32_t *)scratch(c, "dco_co", sizeof(int32_t) * (E + 1)) : col_off;
        int32_t *ri = host ? (int32_t *)scratch(c, "dco_ri", sizeof(int32_t) * E) : row_ind;
        float *vv = !want_val ? nullptr
                    : host    ? (float *)scratch(c, "dco_v", sizeof(float) * E)
                              : val;
        int64_t L = 0;
        timed(c, [&] { L = graph_to_dcsc_dev(c, g, part_size, part_num, pp, ci, co, ri, vv); });
        if (host) {
            copy_from_device(c, part_ptr, pp, sizeof(int32_t) * (part_num + 1), mem);
            copy_from_device(c, col_off, co, sizeof(int32_t) * (L + 1), mem);
            if (L) copy_from_device(c, col_ind, ci, sizeof(int32_t) * L, mem);
            if (g->nnz) copy_from_device(c, row_ind, ri, sizeof(int32_t) * g->nnz, mem);
            if (want_val && g->nnz) copy_from_device(c, val, vv, sizeof(float) * g->nnz, mem);
        }
        sync(c);
        *col_length = L;
    });
}

int apspie_graph_transpose(apspie_ctx *c, const apspie_graph *g, apspie_graph **out) {
    if (!c || !g || !out) return APSPIE_ERR_INVALID;
    *out = nullptr;
    apspie_graph *t = new apspie_graph();
    int st = guard(c, [&] {
        check_graph(c, g);
        set_device(c);
        graph_transpose_dev(c, g, t);
        sync(c);
    });
    if (st != APSPIE_OK) {
        delete t;
        return st;
    }
    c->graphs.insert(t);
    *out = t;
    return APSPIE_OK;
}

// --------------------------------------------------------------- algorithms
int apspie_bfs(apspie_ctx *c, const apspie_graph *g, int32_t source, int32_t *depth, int mem) {
    if (!c || !depth) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(source >= 0 && source < g->m, APSPIE_ERR_INVALID, "source out of range");
        APSPIE_CHECK(g->m == g->n, APSPIE_ERR_INVALID, "BFS needs a square matrix");
        set_device(c);
        int32_t *out = mem == APSPIE_MEM_DEVICE
                           ? depth
                           : (int32_t *)scratch(c, "bfs_out", sizeof(int32_t) * g->m);
        timed(c, [&] { bfs_run(c, g, source, out); });
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, depth, out, sizeof(int32_t) * g->m, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_sssp(apspie_ctx *c, const apspie_graph *g, int32_t source, float *dist, int mem) {
    if (!c || !dist) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(source >= 0 && source < g->m, APSPIE_ERR_INVALID, "source out of range");
        APSPIE_CHECK(g->m == g->n, APSPIE_ERR_INVALID, "SSSP needs a square matrix");
        set_device(c);
        check_sssp_weights(c, g);
        float *out = mem == APSPIE_MEM_DEVICE ? dist
                                              : (float *)scratch(c, "sssp_out", sizeof(float) * g->m);
        timed(c, [&] { sssp_run(c, g, source, out); });
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, dist, out, sizeof(float) * g->m, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_mis(apspie_ctx *c, const apspie_graph *g, const uint32_t *prio, int prio_mem,
               int32_t *in_set, int mem) {
    if (!c || !in_set) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(g->m == g->n, APSPIE_ERR_INVALID, "MIS needs a square matrix");
        set_device(c);
        const uint32_t *p = prio;
        if (!prio) {
            if (c->mis_prio_m != g->m) {
                if (c->mis_prio.ptr) dev_free(c, c->mis_prio.ptr, c->mis_prio.bytes);
                c->mis_prio = Buffer{};
                c->mis_prio_m = -1;
                const size_t bytes = sizeof(uint32_t) * std::max<int32_t>(g->m, 1);
                c->mis_prio.ptr = dev_alloc(c, bytes);
                c->mis_prio.bytes = bytes;
                mt19937_priorities(c, 5489u, g->m, (uint32_t *)c->mis_prio.ptr);
                c->mis_prio_m = g->m;
            }
            p = (const uint32_t *)c->mis_prio.ptr;
        } else if (prio_mem == APSPIE_MEM_HOST) {
            uint32_t *d = (uint32_t *)scratch(c, "mis_prio", sizeof(uint32_t) * g->m);
            copy_to_device(c, d, prio, sizeof(uint32_t) * g->m, APSPIE_MEM_HOST);
            p = d;
        }
        int32_t *out = mem == APSPIE_MEM_DEVICE
                           ? in_set
                           : (int32_t *)scratch(c, "mis_out", sizeof(int32_t) * g->m);
        timed(c, [&] { mis_run(c, g, p, out); });
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, in_set, out, sizeof(int32_t) * g->m, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_tc(apspie_ctx *c, const apspie_graph *g, uint64_t *total, int32_t *per_row, int mem) {
    if (!c || !total) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(g->m == g->n, APSPIE_ERR_INVALID, "TC needs a square matrix");
        set_device(c);
        int32_t *out = nullptr;
        if (per_row)
            out = mem == APSPIE_MEM_DEVICE
                      ? per_row
                      : (int32_t *)scratch(c, "tc_out", sizeof(int32_t) * std::max(g->m, 1));
        timed(c, [&] { tc_run(c, g, total, out); });
        if (per_row && mem == APSPIE_MEM_HOST) {
            copy_from_device(c, per_row, out, sizeof(int32_t) * g->m, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_spmspv(apspie_ctx *c, const apspie_graph *g, int semiring, int64_t nx,
                  const int32_t *x_ind, const float *x_val, const int32_t *mask, int64_t *ny,
                  int32_t *y_ind, float *y_val, int mem, int path) {
    if (!c || !ny || !y_ind || !y_val || nx < 0 || (nx && !x_ind)) return APSPIE_ERR_INVALID;
    if (semiring < 0 || semiring > 3 || path < 0 || path > 2) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(nx <= g->m, APSPIE_ERR_INVALID, "more frontier entries than rows");
        set_device(c);
        const int32_t *xi = x_ind, *mk = mask;
        const float *xv = x_val;
        int32_t *yi = y_ind;
        float *yv = y_val;
        if (mem == APSPIE_MEM_HOST) {
            int32_t *d = (int32_t *)scratch(c, "spv_xi", sizeof(int32_t) * std::max<int64_t>(nx, 1));
            copy_to_device(c, d, x_ind, sizeof(int32_t) * nx, APSPIE_MEM_HOST);
            xi = d;
            if (x_val) {
                float *dv = (float *)scratch(c, "spv_xv", sizeof(float) * std::max<int64_t>(nx, 1));
                copy_to_device(c, dv, x_val, sizeof(float) * nx, APSPIE_MEM_HOST);
                xv = dv;
            }
            if (mask) {
                int32_t *dm = (int32_t *)scratch(c, "spv_mask", sizeof(int32_t) * g->n);
                copy_to_device(c, dm, mask, sizeof(int32_t) * g->n, APSPIE_MEM_HOST);
                mk = dm;
            }
            yi = (int32_t *)scratch(c, "spv_yi", sizeof(int32_t) * std::max(g->n, 1));
            yv = (float *)scratch(c, "spv_yv", sizeof(float) * std::max(g->n, 1));
        }
        int64_t k = 0;
        timed(c, [&] { k = spmspv_run(c, g, semiring, nx, xi, xv, mk, yi, yv, path); });
        *ny = k;
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, y_ind, yi, sizeof(int32_t) * k, APSPIE_MEM_HOST);
            copy_from_device(c, y_val, yv, sizeof(float) * k, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_components(apspie_ctx *c, const apspie_graph *g, int32_t *label, int mem) {
    if (!c || !label) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(g->m == g->n, APSPIE_ERR_INVALID, "components need a square matrix");
        APSPIE_CHECK(mem == APSPIE_MEM_HOST || mem == APSPIE_MEM_DEVICE, APSPIE_ERR_INVALID,
                     "mem must be APSPIE_MEM_HOST or APSPIE_MEM_DEVICE");
        set_device(c);
        Graph *gg = const_cast<Graph *>(static_cast<const Graph *>(g));
        APSPIE_CHECK(graph_is_symmetric(c, gg), APSPIE_ERR_UNSUPPORTED,
                     "components need a symmetric CSR (an undirected graph)");
        const int32_t *lab = nullptr;
        timed(c, [&] { lab = graph_components(c, gg); });
        copy_from_device(c, label, lab, sizeof(int32_t) * g->m, mem);
        sync(c);
    });
}

int apspie_mxv_dense(apspie_ctx *c, const apspie_graph *g, const float *x, float *y,
                     int64_t *frontier_edges, int mem) {
    if (!c || !x || !y) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(mem == APSPIE_MEM_HOST || mem == APSPIE_MEM_DEVICE, APSPIE_ERR_INVALID,
                     "mem must be APSPIE_MEM_HOST or APSPIE_MEM_DEVICE");
        set_device(c);
        const float *xd = x;
        float *yd = y;
        if (mem == APSPIE_MEM_HOST) {
            float *d = (float *)scratch(c, "mxd_x", sizeof(float) * std::max(g->m, 1));
            copy_to_device(c, d, x, sizeof(float) * g->m, APSPIE_MEM_HOST);
            xd = d;
            yd = (float *)scratch(c, "mxd_y", sizeof(float) * std::max(g->n, 1));
        }
        int64_t total = 0;
        timed(c, [&] { total = mxv_dense_run(c, g, xd, yd, APSPIE_PATH_AUTO); });
        if (frontier_edges) *frontier_edges = total;
        // an empty frontier leaves y as it was: nothing to copy back
        if (mem == APSPIE_MEM_HOST && c->stats.vertices_reached > 0) {
            copy_from_device(c, y, yd, sizeof(float) * g->n, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_spgemm(apspie_ctx *c, const apspie_graph *a, const apspie_graph *b,
                  apspie_graph **out) {
    if (!c || !a || !b || !out) return APSPIE_ERR_INVALID;
    *out = nullptr;
    return guard(c, [&] {
        check_graph(c, a);
        check_graph(c, b);
        APSPIE_CHECK(a->n == b->m, APSPIE_ERR_INVALID, "inner dimensions differ");
        set_device(c);
        Graph *r = nullptr;
        timed(c, [&] { r = spgemm_run(c, a, b); });
        c->graphs.insert(r);
        *out = static_cast<apspie_graph *>(r);
    });
}

// Column j of C = A * B is A * B[:,j], one sparse mXv per column of B
// (mXm tc.cuh:314-375).  All columns advance at once as one frontier matrix:
// a CSC array pair read as CSR is the transpose, so C^T = B^T * A^T is the
// row-wise hash product of the two graphs in swapped order, and its CSR is
// C's CSC with the rows of each column sorted.
int apspie_mxm(apspie_ctx *c, const apspie_graph *a_csc, const apspie_graph *b_csc,
               apspie_graph **c_csc) {
    if (!c || !a_csc || !b_csc || !c_csc) return APSPIE_ERR_INVALID;
    *c_csc = nullptr;
    return guard(c, [&] {
        check_graph(c, a_csc);
        check_graph(c, b_csc);
        // a_csc holds cols(A) columns over rows(A); b_csc cols(B) over rows(B)
        APSPIE_CHECK(b_csc->n == a_csc->m, APSPIE_ERR_INVALID,
                     "rows of B differ from columns of A");
        set_device(c);
        Graph *r = nullptr;
        timed(c, [&] { r = spgemm_run(c, b_csc, a_csc); });
        c->graphs.insert(r);
        *c_csc = static_cast<apspie_graph *>(r);
    });
}

int apspie_spgemm_masked(apspie_ctx *c, const apspie_graph *a, const apspie_graph *b,
                         const apspie_graph *mask, float *c_val, int64_t *c_nnz, int mem) {
    if (!c || !a || !b || !mask || !c_nnz) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, a);
        check_graph(c, b);
        check_graph(c, mask);
        APSPIE_CHECK(a->n == b->m, APSPIE_ERR_INVALID, "inner dimensions differ");
        APSPIE_CHECK(mask->m == a->m && mask->n == b->n, APSPIE_ERR_INVALID,
                     "mask shape differs from the product's");
        APSPIE_CHECK(mask->nnz == 0 || c_val, APSPIE_ERR_INVALID, "c_val is NULL");
        set_device(c);
        float *out = mem == APSPIE_MEM_DEVICE
                         ? c_val
                         : (float *)scratch(c, "msk_out",
                                            sizeof(float) * std::max<int64_t>(mask->nnz, 1));
        int64_t nnz = 0;
        timed(c, [&] { nnz = spgemm_masked_run(c, a, b, mask, out); });
        if (mem == APSPIE_MEM_HOST && mask->nnz > 0) {
            copy_from_device(c, c_val, out, sizeof(float) * mask->nnz, APSPIE_MEM_HOST);
            sync(c);
        }
        *c_nnz = nnz;
    });
}

int apspie_scan_exclusive(apspie_ctx *c, const int32_t *in, int32_t *out, int64_t n,
                          int64_t *total, int mem) {
    if (!c || n < 0 || (n && (!in || !out))) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        const int32_t *di = in;
        int32_t *dout = out;
        if (mem == APSPIE_MEM_HOST) {
            int32_t *t = (int32_t *)scratch(c, "scan_in", sizeof(int32_t) * std::max<int64_t>(n, 1));
            copy_to_device(c, t, in, sizeof(int32_t) * n, APSPIE_MEM_HOST);
            di = t;
            dout = (int32_t *)scratch(c, "scan_out", sizeof(int32_t) * std::max<int64_t>(n, 1));
        }
        int64_t tot = 0;
        timed(c, [&] { tot = scan_exclusive_dev(c, di, dout, n); });
        if (total) *total = tot;
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, out, dout, sizeof(int32_t) * n, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_sort_pairs(apspie_ctx *c, uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                      int mem) {
    if (!c || n < 0 || (n && !keys) || bits < 1 || bits > 32) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        set_device(c);
        uint32_t *k = keys, *v = vals;
        if (mem == APSPIE_MEM_HOST) {
            k = (uint32_t *)scratch(c, "sortp_k", sizeof(uint32_t) * std::max<int64_t>(n, 1));
            copy_to_device(c, k, keys, sizeof(uint32_t) * n, APSPIE_MEM_HOST);
            if (vals) {
                v = (uint32_t *)scratch(c, "sortp_v", sizeof(uint32_t) * std::max<int64_t>(n, 1));
                copy_to_device(c, v, vals, sizeof(uint32_t) * n, APSPIE_MEM_HOST);
            }
        }
        timed(c, [&] { sort_pairs_dev(c, k, v, n, bits); });
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, keys, k, sizeof(uint32_t) * n, APSPIE_MEM_HOST);
            if (vals) copy_from_device(c, vals, v, sizeof(uint32_t) * n, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_bfs_batch(apspie_ctx *c, const apspie_graph *g, const int32_t *sources, int32_t nsrc,
                     int32_t *depth, int mem) {
    if (!c || !sources || !depth || nsrc <= 0) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(g->m == g->n, APSPIE_ERR_INVALID, "BFS needs a square matrix");
        for (int32_t i = 0; i < nsrc; ++i)
            APSPIE_CHECK(sources[i] >= 0 && sources[i] < g->m, APSPIE_ERR_INVALID,
                         "source out of range");
        set_device(c);
        size_t bytes = sizeof(int32_t) * (size_t)nsrc * g->m;
        int32_t *out = mem == APSPIE_MEM_DEVICE ? depth : (int32_t *)scratch(c, "bfsb_out", bytes);
        timed(c, [&] { bfs_batch_run(c, g, sources, nsrc, out); });
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, depth, out, bytes, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

int apspie_sssp_batch(apspie_ctx *c, const apspie_graph *g, const int32_t *sources,
                      int32_t nsrc, float *dist, int mem) {
    if (!c || !sources || !dist || nsrc <= 0) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        check_graph(c, g);
        APSPIE_CHECK(g->m == g->n, APSPIE_ERR_INVALID, "SSSP needs a square matrix");
        for (int32_t i = 0; i < nsrc; ++i)
            APSPIE_CHECK(sources[i] >= 0 && sources[i] < g->m, APSPIE_ERR_INVALID,
                         "source out of range");
        set_device(c);
        check_sssp_weights(c, g);
        size_t bytes = sizeof(float) * (size_t)nsrc * g->m;
        float *out = mem == APSPIE_MEM_DEVICE ? dist : (float *)scratch(c, "ssspb_out", bytes);
        timed(c, [&] { sssp_batch_run(c, g, sources, nsrc, out); });
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, dist, out, bytes, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}

}  // extern "C"
