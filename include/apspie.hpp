// apspie.hpp -- header-only C++ layer over the C ABI (apspie.h) that keeps the
// reference's entry points: the same names, argument order and meaning as
// apspie's bfs() / spmspvBfs() / sssp() / spmspvMis() / spgemm() /
// buildMatrix() / d_matrix, with the mgpu::CudaContext& argument replaced by
// apspie::Context&.  Existing callers (the test/<alg>/main.cu drivers) switch
// by including this header instead of the .cuh files.
//
// Differences a caller can see:
//   * the level / round loops run to convergence; the reference's `depth`
//     argument (a CPU-computed level count, test/bfs/main.cu:129-130) is
//     accepted and ignored -- with the CPU depth the outputs are the same;
//   * sssp() uses the real weights in every round (the reference switches to
//     unit weights after its first relaxation round, sssp.cuh:83,91);
//   * spmspvMis() takes the priorities as given (any positive float order, or
//     uint32) with ties broken to the lower id; `delta` (the reference's
//     candidate threshold schedule) changes no result and is ignored;
//   * errors throw apspie::Error (status + message) instead of exit(); define
//     APSPIE_EXIT_ON_ERROR to print and exit(EXIT_FAILURE) like CUDA_SAFE_CALL
//     (mXv.cuh:12-27).
#ifndef APSPIE_HPP
#define APSPIE_HPP

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "apspie.h"

namespace apspie {

class Error : public std::runtime_error {
  public:
    Error(int status, const std::string &msg) : std::runtime_error(msg), status(status) {}
    int status;
};

inline void check(int st, apspie_ctx *ctx, const char *what) {
    if (st == APSPIE_OK) return;
    std::string msg = std::string(what) + ": " + apspie_status_str(st);
    if (ctx) {
        const char *d = apspie_last_error(ctx);
        if (d && *d) msg += std::string(" (") + d + ")";
    }
#ifdef APSPIE_EXIT_ON_ERROR
    std::fprintf(stderr, "apspie error in %s\n", msg.c_str());
    std::exit(EXIT_FAILURE);
#else
    throw Error(st, msg);
#endif
}

// Replaces mgpu::CudaContext (one device, one stream).  It also keeps the
// graph views of the caller's device arrays between calls, so what is derived
// once per graph (the symmetry check that enables pulling, TC's row counts)
// is not redone on every call with the same arrays -- the reference's drivers
// call bfs() / sssp() repeatedly on one CSR.  After changing the contents of
// arrays already passed in, call forget() (or use new arrays).
class Context {
  public:
    explicit Context(int device = 0) { check(apspie_ctx_create(device, &h_), nullptr, "context"); }
    ~Context() {
        forget();
        if (h_) apspie_ctx_destroy(h_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    apspie_ctx *get() const { return h_; }

    const apspie_graph *view(int m, int n, int64_t nnz, const int *d_rowptr, const int *d_col,
                             const float *d_val) {
        const Key k{m, n, nnz, d_rowptr, d_col, d_val};
        for (auto &e : views_)
            if (e.first == k) return e.second;
        apspie_csr c{m, n, nnz, d_rowptr, d_col, d_val};
        apspie_graph *g = nullptr;
        check(apspie_graph_wrap(h_, &c, &g), h_, "graph_wrap");
        views_.emplace_back(k, g);
        return g;
    }
    void forget() {
        for (auto &e : views_) apspie_graph_destroy(e.second);
        views_.clear();
    }

  private:
    struct Key {
        int m, n;
        int64_t nnz;
        const void *rowptr, *col, *val;
        bool operator==(const Key &o) const {
            return m == o.m && n == o.n && nnz == o.nnz && rowptr == o.rowptr && col == o.col &&
                   val == o.val;
        }
    };
    apspie_ctx *h_ = nullptr;
    std::vector<std::pair<Key, apspie_graph *>> views_;
};

// mgpu::CreateCudaDevice(0) (test/bfs/main.cu:27)
inline Context *CreateDevice(int device) { return new Context(device); }

// A non-owning graph view over caller-owned device CSR arrays (kept by the
// context for later calls with the same arrays).
class GraphView {
  public:
    GraphView(Context &ctx, int m, int n, int64_t nnz, const int *d_rowptr, const int *d_col,
              const float *d_val)
        : g_(ctx.view(m, n, nnz, d_rowptr, d_col, d_val)) {}
    const apspie_graph *get() const { return g_; }

  private:
    const apspie_graph *g_;
};

// d_matrix (matrix.hpp:6-37) in its CSR use: device arrays owned by a
// resident graph, plus the host row offsets as the reference keeps them.
struct d_matrix {
    int m = 0, n = 0, nnz = 0;
    int *h_cscColPtr = nullptr;  // host row offsets (m+1)
    int *d_cscColPtr = nullptr;  // device row offsets
    int *d_cscRowInd = nullptr;  // device column indices
    float *d_cscVal = nullptr;   // device values
    apspie_graph *g = nullptr;
    // row-block DCSC (matrix.hpp:7-14), filled by csr_to_dcsc
    int part = 0, col_length = 0;
    int *h_dcscPartPtr = nullptr;     // host copy of d_dcscPartPtr (copy_part)
    int *d_dcscPartPtr = nullptr;     // part+1
    int *d_dcscColPtr_ind = nullptr;  // non-empty columns per part (nnz capacity)
    int *d_dcscColPtr_off = nullptr;  // their entry offsets (nnz+1 capacity)
    int *d_dcscRowInd = nullptr;      // nnz
    float *d_dcscVal = nullptr;       // nnz
    apspie_ctx *dcsc_ctx = nullptr;   // the context the DCSC arrays came from
    int64_t dcsc_cap = 0;             // entries the DCSC arrays were sized for
};

inline void dcsc_free(d_matrix *A) {
    if (!A->dcsc_ctx) return;
    const int64_t e = A->dcsc_cap;
    apspie_dev_free(A->dcsc_ctx, A->d_dcscPartPtr, 4 * ((int64_t)A->part + 1));
    apspie_dev_free(A->dcsc_ctx, A->d_dcscColPtr_ind, 4 * e);
    apspie_dev_free(A->dcsc_ctx, A->d_dcscColPtr_off, 4 * (e + 1));
    apspie_dev_free(A->dcsc_ctx, A->d_dcscRowInd, 4 * e);
    apspie_dev_free(A->dcsc_ctx, A->d_dcscVal, 4 * e);
    std::free(A->h_dcscPartPtr);
    A->h_dcscPartPtr = nullptr;
    A->d_dcscPartPtr = A->d_dcscColPtr_ind = A->d_dcscColPtr_off = A->d_dcscRowInd = nullptr;
    A->d_dcscVal = nullptr;
    A->dcsc_ctx = nullptr;
    A->dcsc_cap = 0;
}

inline void matrix_new(d_matrix *A, int m, int n) {
    A->m = m;
    A->n = n;
    A->nnz = 0;
    A->h_cscColPtr = static_cast<int *>(std::calloc((size_t)m + 1, sizeof(int)));
}

inline void matrix_delete(d_matrix *A) {
    dcsc_free(A);
    if (A->g) apspie_graph_destroy(A->g);
    std::free(A->h_cscColPtr);
    *A = d_matrix{};
}

inline void attach(d_matrix *A, apspie_graph *g, Context &ctx) {
    if (A->g) apspie_graph_destroy(A->g);
    A->g = g;
    apspie_csr v;
    check(apspie_graph_view(g, &v), ctx.get(), "graph_view");
    A->m = v.m;
    A->n = v.n;
    A->nnz = (int)v.nnz;
    A->d_cscColPtr = const_cast<int *>(v.rowptr);
    A->d_cscRowInd = const_cast<int *>(v.col);
    A->d_cscVal = const_cast<float *>(v.val);
    std::free(A->h_cscColPtr);
    A->h_cscColPtr = static_cast<int *>(std::malloc(((size_t)A->m + 1) * sizeof(int)));
    check(apspie_graph_download(ctx.get(), g, A->h_cscColPtr, nullptr, nullptr), ctx.get(),
          "graph_download");
}

// buildMatrix (matrix.cpp:55-111): host COO -> device CSR by a counting sort
// on the row index that keeps the input order inside a row.
template <typename T>
void buildMatrix(d_matrix *A, int numEdge, const int *h_cooRowInd, const int *h_cooColInd,
                 const T *h_cooVal, Context &ctx) {
    static_assert(sizeof(T) == sizeof(float), "values are FP32");
    apspie_coo coo{A->m, A->n, numEdge, h_cooRowInd, h_cooColInd,
                   reinterpret_cast<const float *>(h_cooVal)};
    apspie_graph *g = nullptr;
    check(apspie_graph_from_coo(ctx.get(), &coo, APSPIE_MEM_HOST, 0, &g), ctx.get(), "buildMatrix");
    attach(A, g, ctx);
}

// makeSymmetric (util.cuh:571-643) + buildMatrix: mirror, sort, dedup, drop
// self-loops, CSR with sorted rows.
template <typename T>
void buildSymmetricMatrix(d_matrix *A, int numEdge, const int *h_cooRowInd,
                          const int *h_cooColInd, const T *h_cooVal, Context &ctx) {
    apspie_coo coo{A->m, A->n, numEdge, h_cooRowInd, h_cooColInd,
                   reinterpret_cast<const float *>(h_cooVal)};
    apspie_graph *g = nullptr;
    check(apspie_graph_from_coo(ctx.get(), &coo, APSPIE_MEM_HOST, 1, &g), ctx.get(),
          "makeSymmetric");
    attach(A, g, ctx);
}

// csr_to_dcsc (matrix.cpp:286-289): with alloc the DCSC arrays are sized
// for partNum parts and A's nnz; without it A's CSR is converted into them
// (allocating first if that was not done).  The rows are cut into partNum
// parts of partSize rows, the last taking the rest.
template <typename typeVal>
void csr_to_dcsc(d_matrix *A, int partSize, int partNum, Context &ctx, bool alloc = false) {
    static_assert(sizeof(typeVal) == sizeof(float), "values are FP32");
    const int64_t e = (int64_t)A->nnz > 0 ? A->nnz : 1;
    if (alloc || !A->dcsc_ctx || A->part != partNum || A->dcsc_cap < e) {
        dcsc_free(A);
        A->part = partNum;
        A->dcsc_ctx = ctx.get();
        A->dcsc_cap = e;
        A->h_dcscPartPtr = static_cast<int *>(std::calloc((size_t)partNum + 1, sizeof(int)));
        check(apspie_dev_alloc(ctx.get(), 4 * ((int64_t)partNum + 1), (void **)&A->d_dcscPartPtr),
              ctx.get(), "csr_to_dcsc");
        check(apspie_dev_alloc(ctx.get(), 4 * e, (void **)&A->d_dcscColPtr_ind), ctx.get(),
              "csr_to_dcsc");
        check(apspie_dev_alloc(ctx.get(), 4 * (e + 1), (void **)&A->d_dcscColPtr_off), ctx.get(),
              "csr_to_dcsc");
        check(apspie_dev_alloc(ctx.get(), 4 * e, (void **)&A->d_dcscRowInd), ctx.get(),
              "csr_to_dcsc");
        check(apspie_dev_alloc(ctx.get(), 4 * e, (void **)&A->d_dcscVal), ctx.get(),
              "csr_to_dcsc");
        if (alloc) return;
    }
    int64_t len = 0;
    check(apspie_graph_to_dcsc(ctx.get(), A->g, partSize, partNum, &len, A->d_dcscPartPtr,
                               A->d_dcscColPtr_ind, A->d_dcscColPtr_off, A->d_dcscRowInd,
                               A->d_cscVal ? A->d_dcscVal : nullptr, APSPIE_MEM_DEVICE),
          ctx.get(), "csr_to_dcsc");
    A->col_length = (int)len;
}

// copy_part (matrix.cpp:470-473): the part offsets to the host
inline void copy_part(d_matrix *A) {
    check(apspie_memcpy(A->dcsc_ctx, A->h_dcscPartPtr, A->d_dcscPartPtr,
                        4 * ((int64_t)A->part + 1), 1),
          A->dcsc_ctx, "copy_part");
}

// csr2csc (util.cuh:473-511)
inline void csr2csc(d_matrix *out, const d_matrix *A, Context &ctx) {
    apspie_graph *t = nullptr;
    check(apspie_graph_transpose(ctx.get(), A->g, &t), ctx.get(), "csr2csc");
    attach(out, t, ctx);
}

// bfs<T> (bfs.cuh:32-33): d_cscValA is unused by the reference too.
template <typename T>
void bfs(int vertex, int edge, int m, const T * /*d_cscValA*/, const int *d_cscColPtrA,
         const int *d_cscRowIndA, int *d_bfsResult, int /*depth*/, Context &ctx) {
    GraphView g(ctx, m, m, edge, d_cscColPtrA, d_cscRowIndA, nullptr);
    check(apspie_bfs(ctx.get(), g.get(), vertex, d_bfsResult, APSPIE_MEM_DEVICE), ctx.get(), "bfs");
}

// spmspvBfs (spmspvBfs.cuh:96): the keys-only variant; same result.
inline void spmspvBfs(int vertex, int edge, int m, const int * /*h_csrRowPtr*/,
                      const int *d_csrRowPtr, const int *d_csrColInd, int *d_bfsResult,
                      int /*depth*/, Context &ctx) {
    GraphView g(ctx, m, m, edge, d_csrRowPtr, d_csrColInd, nullptr);
    check(apspie_bfs(ctx.get(), g.get(), vertex, d_bfsResult, APSPIE_MEM_DEVICE), ctx.get(),
          "spmspvBfs");
}

// sssp<V> (sssp.cuh:27-28); unreached = 1.70141e+38 as in the reference.
template <typename V>
void sssp(int vertex, int edge, int m, const V *d_csrValA, const int *d_csrRowPtrA,
          const int *d_csrColIndA, V *d_ssspResult, int /*depth*/, Context &ctx) {
    static_assert(sizeof(V) == sizeof(float), "weights and distances are FP32");
    GraphView g(ctx, m, m, edge, d_csrRowPtrA, d_csrColIndA,
                reinterpret_cast<const float *>(d_csrValA));
    check(apspie_sssp(ctx.get(), g.get(), vertex, reinterpret_cast<float *>(d_ssspResult),
                      APSPIE_MEM_DEVICE),
          ctx.get(), "sssp");
}

// spmspvMis<V> (spmspvMis.cuh:149-150): d_randVec holds one priority per
// vertex on the device; positive floats order like their bit patterns, so
// they are used as uint32 keys directly.  d_misResult: 1 in the set, 0 not.
template <typename V>
void spmspvMis(int edge, int m, const int * /*h_csrRowPtr*/, const int *d_csrRowPtr,
               const int *d_csrColInd, const V *d_randVec, int *d_misResult, float /*delta*/,
               Context &ctx) {
    static_assert(sizeof(V) == sizeof(uint32_t), "priorities are 32-bit");
    GraphView g(ctx, m, m, edge, d_csrRowPtr, d_csrColInd, nullptr);
    check(apspie_mis(ctx.get(), g.get(), reinterpret_cast<const uint32_t *>(d_randVec),
                     APSPIE_MEM_DEVICE, d_misResult, APSPIE_MEM_DEVICE),
          ctx.get(), "spmspvMis");
}

// Triangle counting, the composition of the TC driver (test/tc/main.cu:
// 171-175 buildLower, 243 spgemm, 246 ewiseMultTc, 249 Reduce): per-row
// counts (triangles whose largest vertex is the row) into d_perRow (device,
// may be null) and the exact total as the return value.
inline uint64_t tc(int edge, int m, const int *d_csrRowPtr, const int *d_csrColInd,
                   int *d_perRow, Context &ctx) {
    GraphView g(ctx, m, m, edge, d_csrRowPtr, d_csrColInd, nullptr);
    uint64_t total = 0;
    check(apspie_tc(ctx.get(), g.get(), &total, d_perRow, APSPIE_MEM_DEVICE), ctx.get(), "tc");
    return total;
}

// spgemm (spgemm.cuh:132): C = A * B.  partSize / partNum (the reference's
// 2D tiling of C) are accepted and ignored.
inline void spgemm(d_matrix *C, d_matrix *A, d_matrix *B, int /*partSize*/, int /*partNum*/,
                   Context &ctx) {
    apspie_graph *c = nullptr;
    check(apspie_spgemm(ctx.get(), A->g, B->g, &c), ctx.get(), "spgemm");
    attach(C, c, ctx);
}

// The TC driver's product-then-mask step (spgemm tc.cuh:62-165 followed by
// ewiseMultTc tc.cuh:168-199) as one masked product: d_cVal[e] = (A*B)[i,j]
// for every entry e = (i,j) of M (device, nnz(M) floats); returns the number
// of entries of M that received at least one product.
inline int64_t spgemmMasked(float *d_cVal, d_matrix *A, d_matrix *B, d_matrix *M,
                            Context &ctx) {
    int64_t nnz = 0;
    check(apspie_spgemm_masked(ctx.get(), A->g, B->g, M->g, d_cVal, &nnz, APSPIE_MEM_DEVICE),
          ctx.get(), "spgemmMasked");
    return nnz;
}

// The context reference callers get when a reference entry point takes no
// context (allocScratch): one per process, on device 0, as the reference
// drivers create it (mgpu::CreateCudaDevice(0), test/bfs/main.cu:27).
inline Context &default_context() {
    static Context ctx(0);
    return ctx;
}

// d_scratch (scratch.hpp:5-32): the workspace a reference caller allocates
// once with allocScratch and hands to mXv.  The device arrays are the
// reference's, sized as it sizes them, taken from a context's memory
// (apspie_dev_alloc: counted against its HBM budget) so that callers that
// fill them (d_index, d_ones, d_cscColDiff, tc.cuh:236-254) keep working;
// mXv itself draws its working memory from the context's scratch pool.
struct d_scratch {
    int *d_cscVecInd = nullptr, *d_cscSwapInd = nullptr;
    float *d_cscVecVal = nullptr, *d_cscSwapVal = nullptr, *d_cscTempVal = nullptr;
    int *d_cscColGood = nullptr, *d_cscColBad = nullptr, *d_cscColDiff = nullptr;
    int *d_ones = nullptr, *d_index = nullptr;
    void *d_temp_storage = nullptr;
    int *d_randVecInd = nullptr;
    int *h_cscVecInd = nullptr;
    float *h_cscVecVal = nullptr;
    int *h_cscColDiff = nullptr, *h_ones = nullptr, *h_index = nullptr, *h_bfsResult = nullptr;
    float *h_spmvResult = nullptr, *h_bfsValA = nullptr;
    apspie_ctx *ctx = nullptr;  // the context the device arrays came from
    int edge = 0, m = 0;
};

// allocScratch (util.cuh:646-672): device arrays of edge or m entries as the
// reference allocates them, host arrays by malloc.
inline void allocScratch(d_scratch **d, const int edge, const int m, Context &ctx) {
    d_scratch *s = new d_scratch();
    s->ctx = ctx.get();
    s->edge = edge;
    s->m = m;
    const int64_t e = edge > 0 ? edge : 1, n = m > 0 ? m : 1;
    auto dev = [&](void **p, int64_t count) {
        check(apspie_dev_alloc(s->ctx, 4 * count, p), s->ctx, "allocScratch");
    };
    dev((void **)&s->d_cscVecInd, e);
    dev((void **)&s->d_cscSwapInd, e);
    dev((void **)&s->d_cscVecVal, e);
    dev((void **)&s->d_cscSwapVal, e);
    dev((void **)&s->d_cscTempVal, e);
    dev((void **)&s->d_cscColGood, e);
    dev((void **)&s->d_cscColBad, n);
    dev((void **)&s->d_cscColDiff, n);
    dev((void **)&s->d_ones, n);
    dev((void **)&s->d_index, n);
    dev((void **)&s->d_randVecInd, n);
    s->h_cscVecInd = static_cast<int *>(std::malloc(4 * (size_t)e));
    s->h_cscVecVal = static_cast<float *>(std::malloc(4 * (size_t)e));
    s->h_cscColDiff = static_cast<int *>(std::malloc(4 * (size_t)n));
    s->h_ones = static_cast<int *>(std::malloc(4 * (size_t)n));
    s->h_index = static_cast<int *>(std::malloc(4 * (size_t)n));
    s->h_bfsResult = static_cast<int *>(std::malloc(4 * (size_t)n));
    s->h_spmvResult = static_cast<float *>(std::malloc(4 * (size_t)n));
    s->h_bfsValA = static_cast<float *>(std::malloc(4 * (size_t)e));
    *d = s;
}

// The reference signature: the arrays come from default_context().
inline void allocScratch(d_scratch **d, const int edge, const int m) {
    allocScratch(d, edge, m, default_context());
}

// The reference never frees its scratch; this returns everything
// allocScratch took (device arrays to their context).
inline void freeScratch(d_scratch *d) {
    if (!d) return;
    const int64_t e = d->edge > 0 ? d->edge : 1, n = d->m > 0 ? d->m : 1;
    void *dev_e[] = {d->d_cscVecInd, d->d_cscSwapInd, d->d_cscVecVal,
                     d->d_cscSwapVal, d->d_cscTempVal, d->d_cscColGood};
    void *dev_n[] = {d->d_cscColBad, d->d_cscColDiff, d->d_ones, d->d_index, d->d_randVecInd};
    for (void *p : dev_e)
        if (p) apspie_dev_free(d->ctx, p, 4 * e);
    for (void *p : dev_n)
        if (p) apspie_dev_free(d->ctx, p, 4 * n);
    void *host[] = {d->h_cscVecInd, d->h_cscVecVal, d->h_cscColDiff, d->h_ones,
                    d->h_index,     d->h_bfsResult, d->h_spmvResult, d->h_bfsValA};
    for (void *p : host) std::free(p);
    delete d;
}

// Dense mXv (mXv.cuh:110-111): d_mmResult = A x over (+,x) with A in CSC
// (column v's rows d_cscRowInd[d_cscColPtr[v] ..]) and x = d_randVec, both
// vectors dense on the device.  The entries of x above 0.5 form the frontier;
// d_mmResult is zeroed and receives the sums; returns the sum of the
// frontier columns' lengths.  With no frontier d_mmResult is left as it was
// and 0 is returned.  `d` is accepted for the reference's signature.
template <typename T>
int mXv(const T *d_randVec, const int edge, const int m, const T *d_cscVal,
        const int *d_cscColPtr, const int *d_cscRowInd, T *d_mmResult, d_scratch * /*d*/,
        Context &context) {
    static_assert(sizeof(T) == sizeof(float), "values are FP32");
    GraphView g(context, m, m, edge, d_cscColPtr, d_cscRowInd,
                reinterpret_cast<const float *>(d_cscVal));
    int64_t total = 0;
    check(apspie_mxv_dense(context.get(), g.get(), reinterpret_cast<const float *>(d_randVec),
                           reinterpret_cast<float *>(d_mmResult), &total, APSPIE_MEM_DEVICE),
          context.get(), "mXv");
    return (int)total;
}

// Sparse mXv (mXv.cuh:249-250): y = A x over (+,x) with x as nnz sorted
// indices d_randVecInd and values d_randVecVal (device), A in CSC.  On return
// nnz is y's entry count and d_resultInd / d_resultVal (device, capacity m)
// hold y's indices in ascending order and their sums.  nnz == 0 returns at
// once (mXv.cuh:281-283).  `d` is accepted for the reference's signature.
template <typename T>
void mXv(const int *d_randVecInd, const T *d_randVecVal, const int edge, const int m, int &nnz,
         const T *d_cscVal, const int *d_cscColPtr, const int *d_cscRowInd, int *d_resultInd,
         T *d_resultVal, d_scratch * /*d*/, Context &context) {
    static_assert(sizeof(T) == sizeof(float), "values are FP32");
    if (nnz == 0) return;
    GraphView g(context, m, m, edge, d_cscColPtr, d_cscRowInd,
                reinterpret_cast<const float *>(d_cscVal));
    int64_t ny = 0;
    check(apspie_spmspv(context.get(), g.get(), APSPIE_SR_PLUS_TIMES, nnz, d_randVecInd,
                        reinterpret_cast<const float *>(d_randVecVal), nullptr, &ny, d_resultInd,
                        reinterpret_cast<float *>(d_resultVal), APSPIE_MEM_DEVICE,
                        APSPIE_PATH_AUTO),
          context.get(), "mXv");
    nnz = (int)ny;
}

// SpMSpV under any of the four semirings (the per-algorithm copies of mXv:
// spmspvBfs.cuh:174-222, spmspvSssp.cuh:172-241, spmspvMis.cuh:245-293);
// returns nnz(y).
template <typename T>
int spmspv(const int *d_xInd, const T *d_xVal, int edge, int m, int nnz_x, const T *d_cscVal,
           const int *d_cscColPtr, const int *d_cscRowInd, int *d_resultInd, T *d_resultVal,
           int semiring, Context &ctx) {
    GraphView g(ctx, m, m, edge, d_cscColPtr, d_cscRowInd,
                reinterpret_cast<const float *>(d_cscVal));
    int64_t ny = 0;
    check(apspie_spmspv(ctx.get(), g.get(), semiring, nnz_x, d_xInd,
                        reinterpret_cast<const float *>(d_xVal), nullptr, &ny, d_resultInd,
                        reinterpret_cast<float *>(d_resultVal), APSPIE_MEM_DEVICE,
                        APSPIE_PATH_AUTO),
          ctx.get(), "spmspv");
    return (int)ny;
}

// mXm (tc.cuh:314-316): C = A * B over (+,x) with m x m A and B in CSC on the
// device, column j of C being mXv of A with column j of B; edge = nnz(A),
// nnz(B) = h_cscColPtrB[m].  The reference allocates C's arrays inside and
// does not hand them back; here C comes back in a d_matrix whose
// d_cscColPtr / d_cscRowInd / d_cscVal are C's column offsets, row indices
// (sorted within each column) and values.  Returns nnz(C).
template <typename T>
int mXm(int edge, int m, const T *d_cscValA, const int *d_cscColPtrA, const int *d_cscRowIndA,
        const T *d_cscValB, const int *h_cscColPtrB, const int *d_cscColPtrB,
        const int *d_cscRowIndB, d_matrix *C, Context &ctx) {
    static_assert(sizeof(T) == sizeof(float), "values are FP32");
    GraphView a(ctx, m, m, edge, d_cscColPtrA, d_cscRowIndA,
                reinterpret_cast<const float *>(d_cscValA));
    GraphView b(ctx, m, m, h_cscColPtrB[m], d_cscColPtrB, d_cscRowIndB,
                reinterpret_cast<const float *>(d_cscValB));
    apspie_graph *c = nullptr;
    check(apspie_mxm(ctx.get(), a.get(), b.get(), &c), ctx.get(), "mXm");
    attach(C, c, ctx);
    return C->nnz;
}

// mXm with the reference's argument list (tc.cuh:314-316): C's column
// offsets (m+1), row indices and values go into caller-owned device arrays.
// The reference sizes C's entry arrays at 16 * edge (tc.cuh:328-329); a
// product with more entries than that returns APSPIE_ERR_OVERFLOW instead of
// writing past them.  Returns nnz(C).
template <typename T>
int mXm(const int edge, const int m, const T *d_cscValA, const int *d_cscColPtrA,
        const int *d_cscRowIndA, const T *d_cscValB, const int *h_cscColPtrB,
        const int *d_cscColPtrB, const int *d_cscRowIndB, T *d_cscValC, int *d_cscColPtrC,
        int *d_cscRowIndC, Context &ctx) {
    d_matrix C;
    const int nnz = mXm<T>(edge, m, d_cscValA, d_cscColPtrA, d_cscRowIndA, d_cscValB,
                           h_cscColPtrB, d_cscColPtrB, d_cscRowIndB, &C, ctx);
    if ((int64_t)nnz > 16 * (int64_t)edge) {
        matrix_delete(&C);
        check(APSPIE_ERR_OVERFLOW, nullptr, "mXm: C has more than 16 * edge entries");
    }
    check(apspie_memcpy(ctx.get(), d_cscColPtrC, C.d_cscColPtr, 4 * ((int64_t)m + 1), 2),
          ctx.get(), "mXm");
    check(apspie_memcpy(ctx.get(), d_cscRowIndC, C.d_cscRowInd, 4 * (int64_t)nnz, 2), ctx.get(),
          "mXm");
    check(apspie_memcpy(ctx.get(), d_cscValC, C.d_cscVal, 4 * (int64_t)nnz, 2), ctx.get(),
          "mXm");
    check(apspie_ctx_sync(ctx.get()), ctx.get(), "mXm");
    matrix_delete(&C);
    return nnz;
}

}  // namespace apspie

#endif
