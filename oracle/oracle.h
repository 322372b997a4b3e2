/*
 * oracle.h -- CPU restatement of the apspie SpMSpV / hash-SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  Only the test suite, the smoke check
 * and the CPU-baseline leg of the benchmark load it; the product library
 * (apspie_amd/libapspie.so) never links, loads or calls anything here.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the apspie source tree).  Each function is pinned by
 * tests/test_oracle.py against: the reference's own mt19937ar build
 * (oracle/_ref), the MatrixMarket fixtures of dataset/small with known
 * answers, independent library routines (scipy.sparse.csgraph, networkx,
 * dense numpy products) and closed forms (lattices, complete graphs, paths).
 *
 * Conventions: int32 vertex ids, int64 edge counts, CSR with sorted rows.
 * All return codes are 0 on success, negative on error.
 */
#ifndef APSPIE_ORACLE_H
#define APSPIE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MT19937 (restates src/mt19937ar.cpp:57-137, 167-171) ---------------- */
typedef struct {
    uint32_t mt[624];
    int mti;
} or_mt19937;

void     or_mt_seed(or_mt19937 *g, uint32_t seed);
uint32_t or_mt_next(or_mt19937 *g);
double   or_mt_res53(or_mt19937 *g);
/* n draws from a freshly seeded generator (convenience for tests). */
void     or_mt_stream(uint32_t seed, int64_t n, uint32_t *out);
void     or_mt_res53_stream(uint32_t seed, int64_t n, double *out);

/* ---- COO / CSR construction ----------------------------------------------- */
/* Parse a MatrixMarket coordinate file held in memory (util.cuh:354-430).
 * Like the reference, the FIRST index column lands in col[] and the second in
 * row[] (util.cuh:382-386), both shifted to 0-based.  A missing value column
 * gives 1.0.  *m, *n, *nnz_decl are the header sizes; entries are written to
 * row/col/val (capacity cap); returns the number read or <0 on error. */
int64_t or_parse_mtx(const char *text, int64_t cap, int32_t *m, int32_t *n,
                     int64_t *nnz_decl, int32_t *row, int32_t *col, float *val,
                     int32_t *weighted);

/* Symmetrize -> sort by (row,col) -> dedup -> drop self-loops
 * (fixed restatement of makeSymmetric, util.cuh:571-643).  Operates in place
 * on arrays of capacity >= 2*nnz; returns the new nnz.  Duplicates keep the
 * value of their first occurrence in the mirrored input order (stable). */
int64_t or_symmetrize(int64_t nnz, int32_t *row, int32_t *col, float *val);

/* COO -> CSR counting sort keeping input order within a row
 * (buildMatrix, matrix.cpp:55-111).  rowptr has m+1 entries. */
int or_coo_to_csr(int32_t m, int64_t nnz, const int32_t *row, const int32_t *col,
                  const float *val, int64_t *rowptr, int32_t *col_out,
                  float *val_out);

/* ---- synthetic inputs ------------------------------------------------------ */
/* R-MAT edge list: ef*2^scale edges, per level one res53 double compared with
 * a, a+b, a+b+c (quadrants (0,0),(0,1),(1,0),(1,1)).  No noise, no permutation. */
int or_rmat_edges(int scale, int ef, double a, double b, double c, uint32_t seed,
                  int32_t *src, int32_t *dst);

/* W x W 4-neighbour lattice with floor(0.10*E) edges removed by a seeded
 * Fisher-Yates shuffle and floor(0.001*E) random long-range edges added.
 * Writes the undirected edge list (u,v) into eu/ev (capacity 2*W*(W-1)+E/1000+1)
 * and returns its length. */
int64_t or_lattice_edges(int32_t W, uint32_t seed, int32_t *eu, int32_t *ev);

/* Symmetric integer weights in [1,255] per undirected pair u<v in CSR order,
 * drawn from mt19937(seed) as 1 + (draw % 255), mirrored to (v,u). */
int or_assign_weights(int32_t m, const int64_t *rowptr, const int32_t *col,
                      uint32_t seed, float *val);

/* ---- graph algorithms ------------------------------------------------------ */
/* Deque BFS (SimpleReferenceBfs, testBfs.cpp:6-72).  depth[] = -1 unreached.
 * Returns the search depth as the reference does (max depth + 1). */
int32_t or_bfs(int32_t m, const int64_t *rowptr, const int32_t *col,
               int32_t src, int32_t *depth);

/* Label-correcting heap SSSP (SimpleReferenceSssp, testSssp.cpp:33-115).
 * dist[] = -1 for unreached vertices, as in the reference CPU code. */
int32_t or_sssp(int32_t m, const int64_t *rowptr, const int32_t *col,
                const float *w, int32_t src, float *dist);

/* Greedy MIS over vertices ordered by (prio desc, id asc): the fixed point of
 * Luby rounds with fixed priorities (spmspvMis.cuh:149-344). in_set[v] in {0,1}. */
int or_mis_greedy(int32_t m, const int64_t *rowptr, const int32_t *col,
                  const uint32_t *prio, int32_t *in_set);

/* Luby rounds restated step by step from spmspvMis.cuh:232-322 with a
 * (prio, id) total order and the threshold schedule minimum -= delta
 * (minimum measured on prio / 2^32).  Returns the number of rounds. */
int or_mis_luby(int32_t m, const int64_t *rowptr, const int32_t *col,
                const uint32_t *prio, float delta, int32_t *in_set);

/* Triangle count with the semantics of ewiseMultTc (tc.cuh:168-199) on the
 * strictly-lower part L: per_row[r] = sum_{k in L_r} |L_r cap L_k| (triangles
 * whose largest vertex is r); returns the uint64 total. */
uint64_t or_tc(int32_t m, const int64_t *rowptr, const int32_t *col,
               int32_t *per_row);

/* Connected-component labels (smallest vertex id of the component). */
int or_components(int32_t m, const int64_t *rowptr, const int32_t *col,
                  int32_t *label);

/* ---- primitives ------------------------------------------------------------ */
enum {
    OR_SR_PLUS_TIMES = 0, /* mXv (+,x), mXv.cuh:110-247 */
    OR_SR_BOOL = 1,       /* spmspvBfs keys-only, spmspvBfs.cuh:96-243 */
    OR_SR_MIN_PLUS = 2,   /* spmspvSssp, spmspvSssp.cuh:86-89, 232 */
    OR_SR_MAX_FIRST = 3   /* spmspvMis max-priority, spmspvMis.cuh:277-289 */
};

/* Sparse y = A^T x under a semiring (sparse mXv, mXv.cuh:249-389): x is given
 * as (x_ind sorted ascending, x_val); y is returned as sorted unique indices
 * and values folded sequentially in (frontier order, CSR order).  If mask is
 * non-NULL, outputs with mask[v] != 0 are dropped (visited filter, bfs.cuh:16-30).
 * Returns ny. */
int64_t or_spmspv(int32_t m, const int64_t *rowptr, const int32_t *col,
                  const float *aval, int semiring, int64_t nx,
                  const int32_t *x_ind, const float *x_val, const int32_t *mask,
                  int32_t *y_ind, float *y_val);

/* Gustavson SpGEMM C = A*B with a dense accumulator (the product that
 * spgemm(), spgemm.cuh:131-342, forms).  Two-call protocol: call with
 * c_col == NULL to get c_rowptr and the nnz, then again to fill. */
int64_t or_spgemm(int32_t m, int32_t k, int32_t n,
                  const int64_t *a_rowptr, const int32_t *a_col, const float *a_val,
                  const int64_t *b_rowptr, const int32_t *b_col, const float *b_val,
                  int64_t *c_rowptr, int32_t *c_col, float *c_val);

/* Column-by-column C = A*B in CSC (mXm, tc.cuh:314-375, one sparse mXv per
 * column of B): A is p x q, B is q x r, both CSC; rows of each column of C
 * sorted, cancelled sums kept as 0.  Two-call protocol as or_spgemm
 * (c_row == NULL: c_colptr and the nnz only). */
int64_t or_mxm(int32_t p, int32_t q, int32_t r,
               const int64_t *a_colptr, const int32_t *a_row, const float *a_val,
               const int64_t *b_colptr, const int32_t *b_row, const float *b_val,
               int64_t *c_colptr, int32_t *c_row, float *c_val);

/* Row-block DCSC of a CSR (csr_to_dcsc, matrix.cpp:286-468): rows cut into pn
 * parts of ps rows (the last takes the rest); per part the non-empty columns
 * col_ind[part_ptr[p]..part_ptr[p+1]) ascending, segment k's rows and values
 * at [col_off[k], col_off[k+1]).  Capacities: part_ptr pn+1, col_ind nnz,
 * col_off nnz+1, row_ind / val_out nnz.  Returns the segment count. */
int64_t or_dcsc(int32_t m, const int64_t *rowptr, const int32_t *col, const float *val,
                int32_t ps, int32_t pn, int64_t *part_ptr, int32_t *col_ind,
                int64_t *col_off, int32_t *row_ind, float *val_out);

#ifdef __cplusplus
}
#endif
/* result invariants that fix BFS depths / SSSP distances exactly (0 = all hold) */
int32_t or_check_bfs(int32_t m, const int64_t *rowptr, const int32_t *col, const int32_t *d,
                     int32_t src);
int32_t or_check_sssp(int32_t m, const int64_t *rowptr, const int32_t *col, const float *w,
                      const float *d, int32_t src, float inf);

#endif
