/*
 * apspie.h -- C ABI of the MI355X (gfx950) SpMSpV graph engine and hash
 * SpGEMM.  Plain C types only; every function returns an int status
 * (APSPIE_OK == 0, negative apspie_status on failure) and never exits.
 *
 * The reference (owensgroup/apspie) exposes this path as C++ templates in
 * headers, parameterised by raw device pointers and an mgpu::CudaContext&.
 * Each entry point below names the reference interface it replaces
 * (file:line in the apspie tree).  A header-only C++ layer with the
 * reference's own names and argument order sits on top (apspie.hpp).
 *
 * Threading: one apspie_ctx per host thread.  A context owns one HIP stream,
 * its scratch memory and its tuning table; calls on a context are
 * synchronous unless the name ends in _async.  Graphs are resident in HBM
 * and may be shared by contexts on the same device.
 */
#ifndef APSPIE_H
#define APSPIE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APSPIE_ABI_VERSION 1

typedef struct apspie_ctx apspie_ctx;
typedef struct apspie_graph apspie_graph;
typedef struct apspie_server apspie_server;

typedef enum {
    APSPIE_OK = 0,
    APSPIE_ERR_INVALID = -1,     /* bad argument (null pointer, id out of range) */
    APSPIE_ERR_OOM = -2,         /* device allocation failed */
    APSPIE_ERR_DEVICE = -3,      /* HIP runtime / kernel error */
    APSPIE_ERR_BUDGET = -4,      /* request exceeds the context's HBM budget */
    APSPIE_ERR_OVERFLOW = -5,    /* internal capacity exceeded */
    APSPIE_ERR_UNSUPPORTED = -6, /* combination not implemented */
    APSPIE_ERR_BUSY = -7         /* server queue full / not ready */
} apspie_status;

/* where a caller-owned buffer lives */
typedef enum { APSPIE_MEM_HOST = 0, APSPIE_MEM_DEVICE = 1 } apspie_mem;

/* semirings of the SpMSpV pipeline */
typedef enum {
    APSPIE_SR_PLUS_TIMES = 0, /* (+, x)      mXv, mXv.cuh:110-247 */
    APSPIE_SR_BOOL = 1,       /* (or, and)   spmspvBfs, spmspvBfs.cuh:96-243 */
    APSPIE_SR_MIN_PLUS = 2,   /* (min, +)    spmspvSssp, spmspvSssp.cuh:91-291 */
    APSPIE_SR_MAX_FIRST = 3   /* (max, x_u)  spmspvMis, spmspvMis.cuh:277-289 */
} apspie_semiring;

/* SpMSpV execution path (both give identical results on integer data) */
typedef enum {
    APSPIE_PATH_AUTO = 0,   /* tuner picks per call */
    APSPIE_PATH_SORT = 1,   /* expand -> key-value radix sort -> segmented reduce */
    APSPIE_PATH_ATOMIC = 2  /* expand -> atomic accumulate into a dense row -> compact */
} apspie_path;

/* Unreached SSSP distance, the reference's sentinel (sssp.cuh:48-49). */
#define APSPIE_SSSP_INF 1.70141e+38f

/* CSR graph / matrix.  32-bit row offsets and column indices, rows sorted.
 * val may be NULL (pattern only).  Mirrors d_matrix (matrix.hpp:6-37) in its
 * CSR use (h_/d_cscColPtr, RowInd, Val fed with (row, col), matrix.cpp:55-111). */
typedef struct {
    int32_t m, n;
    int64_t nnz;
    const int32_t *rowptr; /* m+1 */
    const int32_t *col;    /* nnz */
    const float *val;      /* nnz or NULL */
} apspie_csr;

/* COO triples, the three raw arrays the drivers pass around
 * (test/bfs/main.cu:70-86, coo.cuh:24-100). */
typedef struct {
    int32_t m, n;
    int64_t nnz;
    const int32_t *row;
    const int32_t *col;
    const float *val; /* nnz or NULL */
} apspie_coo;

/* ------------------------------------------------------------------ context */
/* Replaces mgpu::CreateCudaDevice / mgpu::CudaContext& (test/bfs/main.cu:27). */
int apspie_ctx_create(int device, apspie_ctx **out);
int apspie_ctx_destroy(apspie_ctx *ctx);
/* Run on a caller stream (hipStream_t) instead of the context's own; NULL
 * restores the context stream. */
int apspie_ctx_set_stream(apspie_ctx *ctx, void *hip_stream);
void *apspie_ctx_stream(apspie_ctx *ctx);
/* Cap on HBM the context may hold (graphs it uploads + scratch); 0 = none. */
int apspie_ctx_set_hbm_budget(apspie_ctx *ctx, int64_t bytes);
int64_t apspie_ctx_hbm_used(apspie_ctx *ctx);
/* Tuning knobs by name (tile sizes, bin thresholds, path choice); see
 * apspie_ctx_get_option for the list of names. */
int apspie_ctx_set_option(apspie_ctx *ctx, const char *name, int64_t value);
int apspie_ctx_get_option(apspie_ctx *ctx, const char *name, int64_t *value);
/* Release cached scratch memory (with an arena: mark it all free again). */
int apspie_ctx_trim(apspie_ctx *ctx);
/* Carve the context's scratch buffers out of one device block of `bytes`
 * (counted against the HBM budget) instead of one hipMalloc each: growing or
 * dropping a buffer then synchronises only the context's stream, where a
 * hipFree waits for every stream on the device -- what a server with one
 * context per worker needs.  A buffer that no longer fits the arena's largest
 * free range fails with APSPIE_ERR_BUDGET (apspie_ctx_trim, then call again).
 * The block is `bytes` rounded down to a multiple of 256 (at least 256).
 * 0 releases the arena; every call drops the cached scratch.  The reference's
 * callers allocate one workspace up front and hand it to every call
 * (d_scratch from allocScratch, scratch.hpp:5-32, util.cuh:646-672); the
 * arena is that workspace kept inside the context. */
int apspie_ctx_reserve_arena(apspie_ctx *ctx, int64_t bytes);
int64_t apspie_ctx_arena_bytes(apspie_ctx *ctx);
const char *apspie_last_error(apspie_ctx *ctx);
const char *apspie_status_str(int status);
int apspie_abi_version(void);

/* Device memory for callers without their own HIP plumbing: replaces the
 * drivers' cudaMalloc / cudaMemcpy / cudaFree of the CSR and result arrays
 * (test/bfs/main.cu:105-120, 147-152).  kind: 0 host->device,
 * 1 device->host, 2 device->device.  Memory from apspie_dev_alloc counts
 * against the context's HBM budget. */
int apspie_dev_alloc(apspie_ctx *ctx, int64_t bytes, void **ptr);
int apspie_dev_free(apspie_ctx *ctx, void *ptr, int64_t bytes);
int apspie_memcpy(apspie_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind);
/* Wait for all work queued on the context's stream. */
int apspie_ctx_sync(apspie_ctx *ctx);

/* Counters from the last algorithm call on this context. */
typedef struct {
    int64_t edges_traversed; /* sum of out-degrees of expanded frontier vertices */
    int64_t vertices_reached;
    int32_t iterations;      /* BFS/SSSP levels (near-far SSSP: rounds, splits included),
                                MIS rounds */
    int32_t host_syncs;      /* host round trips taken */
    double elapsed_ms;       /* device time of the call (hipEvent) */
    double kernel_ms;        /* summed time of the dominant kernel (profiling on) */
    int64_t kernel_launches; /* launches of the dominant kernel (profiling on) */
    int64_t edges_checked;   /* CSR entries actually read: push levels expand whole rows,
                                pull levels stop at the first frontier neighbour */
} apspie_stats;
int apspie_last_stats(apspie_ctx *ctx, apspie_stats *out);

/* ------------------------------------------------------------------- graphs */
/* Copy a host CSR into HBM (replaces the H2D copies of the drivers,
 * test/bfs/main.cu:118-120, and buildMatrix's device half, matrix.cpp:107-110). */
int apspie_graph_upload(apspie_ctx *ctx, const apspie_csr *host, apspie_graph **out);
/* Wrap caller-owned device CSR arrays without copying: the form the
 * reference's entry points take (bfs.cuh:32-33, sssp.cuh:27-28). */
int apspie_graph_wrap(apspie_ctx *ctx, const apspie_csr *device, apspie_graph **out);
/* Build a CSR on the device from COO triples (host or device memory):
 * optional symmetrize + dedup + self-loop removal (makeSymmetric,
 * util.cuh:571-643) then COO->CSR (buildMatrix matrix.cpp:55-111;
 * cusparseXcoo2csr util.cuh:456-471). */
int apspie_graph_from_coo(apspie_ctx *ctx, const apspie_coo *coo, int mem, int symmetrize,
                          apspie_graph **out);
int apspie_graph_destroy(apspie_graph *g);
int apspie_graph_info(const apspie_graph *g, int32_t *m, int32_t *n, int64_t *nnz,
                      int64_t *hbm_bytes);
/* Device addresses of a resident graph's arrays (the d_cscColPtr /
 * d_cscRowInd / d_cscVal members of d_matrix, matrix.hpp:25-28). */
int apspie_graph_view(const apspie_graph *g, apspie_csr *device_view);
/* Copy the CSR back to host arrays (rowptr m+1, col nnz, val nnz or NULL). */
int apspie_graph_download(apspie_ctx *ctx, const apspie_graph *g, int32_t *rowptr,
                          int32_t *col, float *val);
/* Transpose (csr2csc, util.cuh:473-511; cusparseScsr2csc). */
int apspie_graph_transpose(apspie_ctx *ctx, const apspie_graph *g, apspie_graph **out);
/* Row-block doubly compressed storage (csr_to_dcsc, matrix.cpp:286-468; the
 * d_dcscPartPtr / d_dcscColPtr_ind / d_dcscColPtr_off / d_dcscRowInd /
 * d_dcscVal members of d_matrix, matrix.hpp:7-14).  The rows are cut into
 * part_num parts of part_size rows, the last part taking every row from
 * (part_num-1)*part_size on (which must not exceed m).  Inside part p the
 * entries are grouped by column: segments k in [part_ptr[p], part_ptr[p+1])
 * are its non-empty columns col_ind[k] (ascending), and segment k's entries
 * [col_off[k], col_off[k+1]) hold their rows (row_ind, ascending) and
 * values (val; NULL skips them).  *col_length (host) = number of segments.
 * Capacities: part_ptr part_num+1, col_ind nnz, col_off nnz+1, row_ind and
 * val nnz; all arrays in `mem`. */
int apspie_graph_to_dcsc(apspie_ctx *ctx, const apspie_graph *g, int32_t part_size,
                         int32_t part_num, int64_t *col_length, int32_t *part_ptr,
                         int32_t *col_ind, int32_t *col_off, int32_t *row_ind, float *val,
                         int mem);

/* --------------------------------------------------------------- algorithms */
/* Level-synchronous BFS, depth[v] = hops from source, -1 unreached.
 * Replaces bfs<T>() (bfs.cuh:32-125) and spmspvBfs() (spmspvBfs.cuh:96-243);
 * runs to convergence (the reference's depth argument is not needed). */
int apspie_bfs(apspie_ctx *ctx, const apspie_graph *g, int32_t source, int32_t *depth,
               int mem);
/* apspie_bfs that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture elsewhere in the process, serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and depths as apspie_bfs; its frontier,
 * bitmaps and control block live in one static 96 MiB device array per GPU
 * (held until the process ends).  The graph needs 16(m+1) + 8(nnz/1024+2)
 * + 20(m/32+1) bytes plus 2 KiB (and 4m more for a host output): R-MAT
 * scale 22 fits; larger graphs return APSPIE_ERR_OVERFLOW.  Levels pull only
 * on a graph already known to be symmetric (built with symmetrize = 1, or
 * after an apspie_bfs on it); otherwise they push, with the same depths.
 * Calls on one GPU share the array and keep the last call's contents;
 * overlapping calls would race on it, so the library serialises this
 * process's calls per GPU (a host mutex; each call is synchronous). */
int apspie_bfs_noalloc(apspie_ctx *ctx, const apspie_graph *g, int32_t source, int32_t *depth,
                       int mem);
/* Single-source shortest paths under (min,+) over the graph's FP32 weights;
 * unreached = APSPIE_SSSP_INF.  Replaces sssp<V>() (sssp.cuh:27-115) +
 * spmspvSssp() (spmspvSssp.cuh:91-291), with real weights in every round.
 * Weights must be finite and non-negative: a graph with a negative, NaN or
 * infinite weight is refused with APSPIE_ERR_UNSUPPORTED (the same holds for
 * apspie_sssp_batch).  dist[v] is the smallest FP32 left-fold sum
 * (((w1 + w2) + w3) ...) over the paths to v, which does not depend on the
 * order of relaxation; with integer weights and sums below 2^24 it is exact. */
int apspie_sssp(apspie_ctx *ctx, const apspie_graph *g, int32_t source, float *dist,
                int mem);
/* Maximal independent set by Luby rounds under the max-priority semiring, on
 * an undirected graph (a CSR equal to its transpose, as the reference's
 * drivers build it; a directed CSR gives a set that is not specified);
 * in_set[v] = 1 in the set, 0 not.  prio (m uint32, host or device per
 * prio_mem) may be NULL: then prio[v] = v-th draw of mt19937 seeded 5489.
 * Ties break to the lower id.  Replaces spmspvMis() (spmspvMis.cuh:149-344). */
int apspie_mis(apspie_ctx *ctx, const apspie_graph *g, const uint32_t *prio, int prio_mem,
               int32_t *in_set, int mem);
/* Triangle count by masked SpGEMM L.L o L on the strictly-lower part:
 * per_row[r] = triangles whose largest vertex is r (may be NULL), *total on
 * the host.  Replaces spgemm() + ewiseMultTc() + Reduce (tc.cuh:62-199,
 * test/tc/main.cu:243-249), exact in uint64. */
int apspie_tc(apspie_ctx *ctx, const apspie_graph *g, uint64_t *total, int32_t *per_row,
              int mem);

/* --------------------------------------------------------------- primitives */
/* y = A^T x under a semiring (sparse mXv, mXv.cuh:249-389): x as sorted
 * indices + values (x_val NULL => all ones); mask (dense int32 m, NULL => none)
 * drops outputs with mask[v] != 0.  y_ind/y_val need capacity m; *ny is
 * written on the host.  All arrays live in `mem`. */
int apspie_spmspv(apspie_ctx *ctx, const apspie_graph *g, int semiring, int64_t nx,
                  const int32_t *x_ind, const float *x_val, const int32_t *mask,
                  int64_t *ny, int32_t *y_ind, float *y_val, int mem, int path);
/* Connected components of an undirected graph (a symmetric CSR; otherwise
 * APSPIE_ERR_UNSUPPORTED): label[v] = the smallest vertex id in v's
 * component (m entries in `mem`).  Built once per graph by lock-free
 * union-find on the device and kept with the graph; the batched BFS uses it
 * to mark pairs of a vertex and a source in different components as
 * settled from the start.  No reference counterpart. */
int apspie_components(apspie_ctx *ctx, const apspie_graph *g, int32_t *label, int mem);
/* Dense mXv (mXv.cuh:110-247): y = A^T x over (+,x) with x (m entries) and
 * y (n entries) dense.  The entries of x above 0.5 form the frontier (the
 * reference's bitify, mXv.cuh:50-55) with their values as x_v; y is zeroed,
 * then y[r] = sum over frontier v of x_v * a_vr (a_vr = 1 for a graph
 * without values).  *frontier_edges (host, may be NULL) = the sum of the
 * frontier rows' lengths, the reference's return value (mXv.cuh:229).  An
 * empty frontier leaves y untouched and gives 0 (mXv.cuh:143-145).  x and y
 * live in `mem`. */
int apspie_mxv_dense(apspie_ctx *ctx, const apspie_graph *g, const float *x, float *y,
                     int64_t *frontier_edges, int mem);
/* C = A * B over (+,x) by row-binned hash SpGEMM (spgemm(), spgemm.cuh:131-342
 * + LaunchKernel spgemmKernel.cuh:438-738).  C is a new device graph with
 * sorted rows. */
int apspie_spgemm(apspie_ctx *ctx, const apspie_graph *a, const apspie_graph *b,
                  apspie_graph **c);
/* C = A * B over (+,x) in CSC, column by column: column j of C is the sparse
 * mXv A * B[:,j] (mXm, tc.cuh:314-375, over mXv mXv.cuh:249-389).  a_csc and
 * b_csc hold CSC arrays (their rowptr/col/val are column offsets, row indices
 * and values; a graph's m is then the column count and n the row count);
 * *c_csc is a new device graph in the same form, rows sorted within each
 * column, entries kept where their products cancel to 0.  All columns of B
 * run together through the SpGEMM kernels instead of one launch each. */
int apspie_mxm(apspie_ctx *ctx, const apspie_graph *a_csc, const apspie_graph *b_csc,
               apspie_graph **c_csc);
/* Masked product C = (A * B) o M over (+,x), evaluated on M's pattern only:
 * c_val[e] = sum_k A[i,k] * B[k,j] for every entry e = (i,j) of M in M's CSR
 * order (0 where no product lands; a column repeated within a row of M gets
 * its value at the first occurrence), *c_nnz = entries of M with at least one
 * product.  Replaces the TC composition's materialised product + mask,
 * spgemm() + ewiseMultTc() (tc.cuh:62-199).  Unweighted graphs count as all
 * ones.  c_val (nnz(M) floats) lives in `mem`. */
int apspie_spgemm_masked(apspie_ctx *ctx, const apspie_graph *a, const apspie_graph *b,
                         const apspie_graph *mask, float *c_val, int64_t *c_nnz, int mem);
/* Exclusive prefix sum of n int32 (mgpu::Scan<MgpuScanTypeExc>); *total on host. */
int apspie_scan_exclusive(apspie_ctx *ctx, const int32_t *in, int32_t *out, int64_t n,
                          int64_t *total, int mem);
/* Stable key-value radix sort on the low `bits` bits of uint32 keys
 * (cub::DeviceRadixSort::SortPairs); vals may be NULL (SortKeys). */
int apspie_sort_pairs(apspie_ctx *ctx, uint32_t *keys, uint32_t *vals, int64_t n, int bits,
                      int mem);

/* ------------------------------------------------------------------ batched */
/* Many sources over one resident graph, advanced together as a sparse
 * frontier matrix (one bit / one distance column per source).  Output is
 * row-major [nsrc][m].  No reference counterpart: equal, per source, to
 * apspie_bfs / apspie_sssp. */
int apspie_bfs_batch(apspie_ctx *ctx, const apspie_graph *g, const int32_t *sources,
                     int32_t nsrc, int32_t *depth, int mem);
int apspie_sssp_batch(apspie_ctx *ctx, const apspie_graph *g, const int32_t *sources,
                      int32_t nsrc, float *dist, int mem);
/* Autotune the traversal launch choices for this graph (batch == 0: the
 * single-source BFS / SSSP ones; batch > 0: the batched ones at that many
 * sources) and keep the winners in the context's options.  No reference
 * counterpart: the reference hard-codes NTHREADS and its mgpu LaunchBox
 * tables (spgemm.cuh:181-185). */
int apspie_autotune(apspie_ctx *ctx, const apspie_graph *g, int32_t batch);
/* Autotune the hash machinery on this graph: TC's tier bound (tc_medium_row)
 * and SpGEMM's (spgemm_wave_ub, spgemm_block_ub, spgemm_dense) for A * A --
 * which table size and block shape serves a row; SpGEMM is left untuned when
 * A * A does not fit one call (more than 2^31 entries, or the budget).  Replaces the reference's
 * fixed TABLE_SIZE / MAX_PROBES / THREADS / BLOCKS (spgemmKernel.cuh:9-26). */
int apspie_autotune_hash(apspie_ctx *ctx, const apspie_graph *g);
/* The tuned options of this context as the entry for (this GPU, this graph's
 * class) in a text table at `path` (one line per pair; an existing entry for
 * the pair is replaced), and back: apspie_tune_load applies the entry that
 * matches, setting *found to 1 (0 and nothing applied when there is none or
 * the file does not exist).  The class is log2 of the vertex count, of the
 * mean and of the longest row, symmetric or not, weighted or not. */
int apspie_tune_save(apspie_ctx *ctx, const apspie_graph *g, const char *path);
int apspie_tune_load(apspie_ctx *ctx, const apspie_graph *g, const char *path, int32_t *found);

/* ---------------------------------------------------------------- synthetic */
/* Host-side graph construction (synthetic inputs; the reference has none).
 * Two-call protocol: pass NULL arrays to get sizes. */
int apspie_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                    uint32_t seed, int32_t *rowptr, int32_t *col, int64_t *nnz);
int apspie_gen_lattice(int32_t width, uint32_t seed, int32_t *rowptr, int32_t *col,
                       int64_t *nnz);
int apspie_gen_weights(int32_t m, const int32_t *rowptr, const int32_t *col, uint32_t seed,
                       float *val);
/* Read a MatrixMarket file into COO host arrays (readEdge/readMtx,
 * util.cuh:354-430; first index -> col, second -> row). */
int apspie_read_mtx(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int32_t *row,
                    int32_t *col, float *val);

/* Binary CSR cache: a graph's CSR arrays behind a 64-byte header (magic
 * "APSPCSR1", version, flags, m, n, nnz, checksum), loaded without parsing.
 * Replaces re-reading the text file on every run (readEdge / readMtx,
 * util.cuh:354-430) and the COO->CSR conversion after it (util.cuh:456-471).
 * Host-only write / read (two-call protocol for read: NULL arrays give the
 * sizes), and device-side save / load of a resident graph.  A file whose
 * size, magic or checksum does not match its header is refused with
 * APSPIE_ERR_INVALID. */
int apspie_csr_write(const char *path, const apspie_csr *host);
int apspie_csr_read(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int32_t *has_val,
                    int32_t *rowptr, int32_t *col, float *val);
int apspie_graph_save(apspie_ctx *ctx, const apspie_graph *g, const char *path);
int apspie_graph_load(apspie_ctx *ctx, const char *path, apspie_graph **out);

/* ------------------------------------------------------------------ serving */
typedef enum { APSPIE_Q_BFS = 0, APSPIE_Q_SSSP = 1, APSPIE_Q_TC = 2 } apspie_query_kind;

typedef struct {
    int32_t max_batch;     /* sources per batched launch */
    int32_t slots;         /* concurrent batches in flight (streams) */
    int64_t hbm_budget;    /* bytes for resident graphs + per-slot scratch */
    int32_t queue_capacity;
} apspie_server_config;

/* Continuous-batching query server over HBM-resident graphs. */
int apspie_server_create(int device, const apspie_server_config *cfg, apspie_server **out);
int apspie_server_destroy(apspie_server *s);
/* Admit a host CSR into the resident pool; fails with APSPIE_ERR_BUDGET. */
int apspie_server_add_graph(apspie_server *s, const apspie_csr *host, int32_t *graph_id);
/* Admit a graph built on the device from COO triples (host or device memory;
 * symmetrize as apspie_graph_from_coo), or loaded from a binary CSR cache
 * file.  Admissions may run while queries are in flight; the pool's budget
 * is checked against the built graph's size. */
int apspie_server_add_graph_coo(apspie_server *s, const apspie_coo *coo, int mem, int symmetrize,
                                int32_t *graph_id);
int apspie_server_add_graph_file(apspie_server *s, const char *path, int32_t *graph_id);
/* Message of the last failed admission. */
const char *apspie_server_last_error(apspie_server *s);
/* Enqueue a query; returns a ticket.  source ignored for TC. */
int apspie_server_submit(apspie_server *s, int32_t graph_id, int kind, int32_t source,
                         int64_t *ticket);
/* Enqueue a BFS / SSSP query whose result the server writes into the
 * caller's memory: `out` holds m int32 depths / m float distances, in host
 * or device memory (`mem`), and must stay valid until the wait returns --
 * the reference's entry points take their result arrays from the caller the
 * same way (d_bfsResult, bfs.cuh:32-33).  A TC query ignores `out`; its count
 * comes back from the wait.  The wait then only reports status and latency
 * (pass out = NULL there). */
int apspie_server_submit_into(apspie_server *s, int32_t graph_id, int kind, int32_t source,
                              void *out, int mem, int64_t *ticket);
/* Wait for a ticket; copies the result (m int32 depths / m float distances /
 * 1 uint64 count) into `out` (host) and reports arrival->completion latency. */
int apspie_server_wait(apspie_server *s, int64_t ticket, void *out, int64_t out_bytes,
                       double *latency_ms);
/* Zero-copy wait: *data points at the result in the server's pinned host
 * memory (*bytes long, layout as above) and stays valid until
 * apspie_server_release(ticket) or apspie_server_destroy.  No reference
 * counterpart (the reference has no server). */
int apspie_server_wait_view(apspie_server *s, int64_t ticket, const void **data, int64_t *bytes,
                            double *latency_ms);
int apspie_server_release(apspie_server *s, int64_t ticket);
int apspie_server_drain(apspie_server *s);
/* Work done on one graph for one query kind since the last reset: batches
 * launched, queries answered and the workers' wall time on them (launch to
 * results in host memory; concurrent workers add up).  reset != 0 zeroes
 * the counters after reading. */
int apspie_server_busy(apspie_server *s, int32_t graph_id, int kind, int64_t *batches,
                       int64_t *queries, double *busy_ms, int reset);
/* The same counters split by phase: the traversal / count calls, the copies
 * of result rows to host memory, and obtaining pinned result buffers (wall
 * time of each, summed over batches); busy_ms also holds the rest (retries
 * after a budget failure, bookkeeping).  No reference counterpart. */
typedef struct {
    int64_t batches, queries;
    double busy_ms, traverse_ms, copy_ms, alloc_ms;
    int64_t retries;  /* attempts that ran out of budget or memory */
    double retry_ms;  /* wall time of those attempts and of the trims after them */
} apspie_server_stats;
int apspie_server_busy_stats(apspie_server *s, int32_t graph_id, int kind,
                             apspie_server_stats *out, int reset);

#ifdef __cplusplus
}
#endif
#endif
