// Internal definitions shared by the apspie gfx950 kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/apspie.h"

namespace apspie {

constexpr int kWave = 64;  // CDNA wavefront width

// ------------------------------------------------------------------ errors
struct Error {
    int status;
    std::string msg;
};

#define APSPIE_HIP(call)                                                        \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            throw ::apspie::Error{APSPIE_ERR_DEVICE,                             \
                                  std::string(#call) + ": " + hipGetErrorString(e_)}; \
        }                                                                       \
    } while (0)

#define APSPIE_CHECK(cond, status, msg)                   \
    do {                                                  \
        if (!(cond)) throw ::apspie::Error{(status), (msg)}; \
    } while (0)

#define APSPIE_LAUNCH_CHECK() APSPIE_HIP(hipGetLastError())

// ------------------------------------------------------------------ context
struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct Ctx;

struct Graph {
    Ctx *owner = nullptr;
    int device = 0;
    int32_t m = 0, n = 0;
    int64_t nnz = 0;
    int32_t *rowptr = nullptr;
    int32_t *col = nullptr;
    float *val = nullptr;
    bool owns = false;
    int64_t bytes = 0;
    // derived, built lazily and cached with the graph
    int32_t *lowcnt = nullptr;  // per row: number of cols < row (TC)
    uint32_t *iso_bits = nullptr;  // bit per vertex without entries (BFS visited start)
    int32_t *comp = nullptr;       // connected-component label per vertex (symmetric graphs)
    int32_t max_deg = -1;          // longest row, once known
    int w_range = -1;              // 1 once w_lo / w_hi are known
    float w_lo = 0.0f, w_hi = 0.0f; // smallest and largest value
    int symmetric = -1;         // 1 / 0 once known (pull traversal needs A == A^T)
    int sym_weights = -1;       // 1 / 0 once known: values equal to the transpose's too
    // rows longer than the SSSP pull's lane limit, as a frontier over those
    // vertices (entry vertex, row start, edge offset, tile owners; one block
    // of device memory owned by the graph) and its (entries << 32 | edges)
    void *heavy_mem = nullptr;
    int32_t *heavy_v = nullptr, *heavy_start = nullptr, *heavy_off = nullptr,
            *heavy_tfirst = nullptr;
    uint64_t heavy_cnt = 0;
    // near-far SSSP with light / heavy entries, per width delta: a copy of
    // col / val whose rows hold their light entries (value below delta)
    // first, in CSR order, then the heavy ones, and the end of each row's
    // light part; one block of device memory each, owned by the graph and
    // kept until it is destroyed (a call on another context may still be
    // launching rounds over the copy of another width)
    struct LightFirst {
        float delta;
        void *mem;
        int32_t *col, *end;
        float *val;
    };
    std::vector<LightFirst> lf;
    std::mutex derived_mu;      // guards lazily built derived data
};

struct Ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int64_t budget = 0;  // 0 = unlimited
    int64_t used = 0;    // bytes held: uploaded graphs + scratch
    std::map<std::string, Buffer> scratch;
    // optional arena: one block of device memory the scratch buffers are
    // carved from (first fit, 256-byte granules, free ranges coalesced), so
    // that growing or dropping them never calls hipFree, which waits for every
    // stream on the device
    char *arena = nullptr;
    size_t arena_bytes = 0;
    std::map<size_t, size_t> arena_free;  // offset -> length
    std::map<std::string, int64_t> options;
    std::string last_error;
    apspie_stats stats{};
    bool profiling = false;
    int cus = 256;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ev1_recorded = false;  // the call recorded ev1 after its last work itself
    // pinned host word for counters read back by the host
    void *pinned = nullptr;
    // the default MIS priorities (mt19937 seeded 5489, one draw per vertex)
    // for mis_prio_m vertices, kept between calls: the stream is sequential
    // and generated on the host
    Buffer mis_prio;
    int32_t mis_prio_m = -1;
    // graphs whose memory is accounted to this context
    std::set<Graph *> graphs;
    // event pairs of the kernel timer (option "profiling"), reused by calls
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timer_events;
};

int64_t opt(Ctx *c, const char *name, int64_t dflt);

// Scratch buffer by name, grown on demand (contents not preserved on growth).
void *scratch(Ctx *c, const std::string &name, size_t bytes);
// Bytes a new scratch buffer could still take: the largest free range of the
// arena, else what the HBM budget leaves (INT64_MAX without either).
int64_t scratch_room(Ctx *c);
void *dev_alloc(Ctx *c, size_t bytes);  // counted against the budget
void dev_free(Ctx *c, void *p, size_t bytes);

// Run `f` translating exceptions into status codes and recording the message.
template <class F>
int guard(Ctx *c, F &&f) {
    try {
        f();
        return APSPIE_OK;
    } catch (const Error &e) {
        if (c) c->last_error = e.msg;
        return e.status;
    } catch (const std::bad_alloc &) {
        if (c) c->last_error = "host allocation failed";
        return APSPIE_ERR_OOM;
    }
}

void set_device(Ctx *c);
void sync(Ctx *c);

// device time of one algorithm call, bracketed by events on the context stream
template <class F>
void timed(Ctx *c, F &&f) {
    c->stats = apspie_stats{};
    c->ev1_recorded = false;
    APSPIE_HIP(hipEventRecord(c->ev0, c->stream));
    f();
    if (!c->ev1_recorded) APSPIE_HIP(hipEventRecord(c->ev1, c->stream));
    APSPIE_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    APSPIE_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.elapsed_ms = ms;
}
// copy-in / copy-out helpers honouring apspie_mem
void copy_to_device(Ctx *c, void *dst, const void *src, size_t bytes, int mem);
void copy_from_device(Ctx *c, void *dst, const void *src, size_t bytes, int mem);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// number of blocks to fill the chip for a kernel with `threads` per block
int fill_grid(Ctx *c, int threads, int per_cu);

// ----------------------------------------------------------- kernel entry
// (host-side launchers defined in the .hip files)
void bfs_run(Ctx *c, const Graph *g, int32_t source, int32_t *depth_dev);
// BFS over a caller-provided device block instead of the scratch pool (no
// allocation): the bytes a graph needs, and the run (APSPIE_ERR_OVERFLOW when
// the block is too small).  Pulls only on a graph already known symmetric.
int64_t bfs_block_bytes(Ctx *c, const Graph *g);
// device address of the no-alloc BFS's static array, looked up once per device
void noalloc_init(int device);
void bfs_run_in(Ctx *c, const Graph *g, int32_t source, int32_t *depth_dev, void *block,
                int64_t bytes);
void sssp_run(Ctx *c, const Graph *g, int32_t source, float *dist_dev);
void mis_run(Ctx *c, const Graph *g, const uint32_t *prio_dev, int32_t *out_dev);
void tc_run(Ctx *c, const Graph *g, uint64_t *total, int32_t *per_row_dev);
int64_t spmspv_run(Ctx *c, const Graph *g, int semiring, int64_t nx, const int32_t *x_ind,
                   const float *x_val, const int32_t *mask, int32_t *y_ind, float *y_val,
                   int path);
// dense x (m) -> dense y (n) under (+,x); returns the frontier's edge count
int64_t mxv_dense_run(Ctx *c, const Graph *g, const float *x, float *y, int path);
Graph *spgemm_run(Ctx *c, const Graph *a, const Graph *b);  // returns new apspie_graph
// (A * B) o M on M's pattern into cval (device, nnz(M)); returns the entries
// of M with at least one product
int64_t spgemm_masked_run(Ctx *c, const Graph *a, const Graph *b, const Graph *mask, float *cval);
int64_t scan_exclusive_dev(Ctx *c, const int32_t *in, int32_t *out, int64_t n);
void sort_pairs_dev(Ctx *c, uint32_t *keys, uint32_t *vals, int64_t n, int bits);
void bfs_batch_run(Ctx *c, const Graph *g, const int32_t *src_host, int32_t nsrc,
                   int32_t *depth_dev);
void sssp_batch_run(Ctx *c, const Graph *g, const int32_t *src_host, int32_t nsrc,
                    float *dist_dev);
void graph_from_coo_dev(Ctx *c, Graph *g, const int32_t *row, const int32_t *col,
                        const float *val, int64_t nnz, int symmetrize);
// Transposed copy of g into t.  With `in_scratch` its arrays are the
// context's scratch buffers of that name prefix (a temporary: t owns nothing
// and is valid until those buffers are reused), else t owns new device arrays.
void graph_transpose_dev(Ctx *c, const Graph *g, Graph *t, const char *in_scratch = nullptr);
int64_t graph_to_dcsc_dev(Ctx *c, const Graph *g, int32_t ps, int32_t pn, int32_t *part_ptr,
                          int32_t *col_ind, int32_t *col_off, int32_t *row_ind, float *val);
void mt19937_priorities(Ctx *c, uint32_t seed, int32_t m, uint32_t *out_dev);
void fill_i32(Ctx *c, int32_t *p, int32_t v, int64_t n);
void graph_prepare_tc(Ctx *c, Graph *g);
// Whether the CSR equals its transpose (cached with the graph).
bool graph_is_symmetric(Ctx *c, Graph *g);
// symmetric structure and values (pulling a weighted relaxation reads w(v, u)
// for w(u, v))
bool graph_is_symmetric_weighted(Ctx *c, Graph *g);
const uint32_t *graph_iso_bits(Ctx *c, Graph *g);
// Connected components of a symmetric graph, built once and kept with it:
// label[v] = the smallest vertex id of v's component.
const int32_t *graph_components(Ctx *c, Graph *g);
void components_into(Ctx *c, const Graph *g, int32_t *label_dev);
// What is known without computing anything: the symmetry flag (-1 unknown,
// 0, 1) and the isolated-vertex bits (null until built)
int graph_symmetry_known(Graph *g);
const uint32_t *graph_iso_bits_cached(Graph *g);
// the isolated-vertex bits of g into caller memory (ceil(m / 32) words)
void iso_bits_into(Ctx *c, const Graph *g, uint32_t *bits);
int32_t graph_max_degree(Ctx *c, Graph *g);
// Smallest and largest value of a weighted graph with entries, computed once
// (false: no values or no entries).
bool graph_weight_range(Ctx *c, Graph *g, float *lo, float *hi);
// Near-far split width for sssp_delta = -1 (0 = Bellman-Ford rounds), and the
// width the weights alone suggest (what the autotuner centres its candidates on).
double sssp_auto_delta(Ctx *c, Graph *g);
double sssp_weight_delta(Ctx *c, Graph *g);
// The batched SSSP width for batch_sssp_delta = -1 (0 = Bellman-Ford rounds).
double sssp_batch_auto_delta(Ctx *c, Graph *g);

}  // namespace apspie

// The opaque ABI handles are the internal structs.
struct apspie_ctx : apspie::Ctx {};
struct apspie_graph : apspie::Graph {};
