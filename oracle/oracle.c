/*
 * oracle.c -- CPU restatement of the apspie SpMSpV / hash-SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Loaded
 * by tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py,
 * never by the product library.  Written from the reference's behaviour, not
 * its text; citations are file:line in the apspie source tree.
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* MT19937 -- the Matsumoto-Nishimura generator that src/mt19937ar.cpp       */
/* implements (init_genrand :57-71, genrand_int32 :102-137, genrand_res53    */
/* :167-171).  Written from the published algorithm (ACM TOMACS 8(1), 1998,  */
/* with the 2002 initialisation), state kept per instance instead of in      */
/* file-scope statics.                                                       */
/* ======================================================================== */
#define MT_N 624
#define MT_M 397

void or_mt_seed(or_mt19937 *g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < MT_N; ++i) {
        uint32_t prev = g->mt[i - 1];
        g->mt[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
    }
    g->mti = MT_N;
}

static void mt_twist(or_mt19937 *g) {
    for (int i = 0; i < MT_N; ++i) {
        uint32_t x = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % MT_N] & 0x7fffffffu);
        uint32_t xa = x >> 1;
        if (x & 1u) xa ^= 0x9908b0dfu;
        g->mt[i] = g->mt[(i + MT_M) % MT_N] ^ xa;
    }
    g->mti = 0;
}

uint32_t or_mt_next(or_mt19937 *g) {
    if (g->mti >= MT_N) mt_twist(g);
    uint32_t y = g->mt[g->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double or_mt_res53(or_mt19937 *g) {
    /* 27 high bits of the first draw, 26 of the second, over 2^53 */
    uint32_t hi = or_mt_next(g) >> 5;
    uint32_t lo = or_mt_next(g) >> 6;
    return ((double)hi * 67108864.0 + (double)lo) / 9007199254740992.0;
}

void or_mt_stream(uint32_t seed, int64_t n, uint32_t *out) {
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = or_mt_next(&g);
}

void or_mt_res53_stream(uint32_t seed, int64_t n, double *out) {
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = or_mt_res53(&g);
}

/* ======================================================================== */
/* MatrixMarket parsing (readEdge util.cuh:354-368, readMtx :371-430)       */
/* ======================================================================== */
static const char *skip_ws(const char *p) {
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    return p;
}

int64_t or_parse_mtx(const char *text, int64_t cap, int32_t *m, int32_t *n,
                     int64_t *nnz_decl, int32_t *row, int32_t *col, float *val,
                     int32_t *weighted) {
    const char *p = text;
    /* comment lines start with '%' (readEdge skips them, util.cuh:356-364) */
    while (*p == '%') {
        while (*p && *p != '\n') ++p;
        if (*p == '\n') ++p;
    }
    long long hm, hn, hz;
    int used = 0;
    if (sscanf(p, "%lld %lld %lld%n", &hm, &hn, &hz, &used) != 3) return -1;
    *m = (int32_t)hm;
    *n = (int32_t)hn;
    *nnz_decl = hz;
    p += used;
    *weighted = 0;
    int64_t k = 0;
    while (k < hz && k < cap) {
        while (*p == '\n' || *p == ' ' || *p == '\t' || *p == '\r') ++p;
        if (!*p) break;
        if (*p == '%') { /* stray comment */
            while (*p && *p != '\n') ++p;
            continue;
        }
        long long a, b;
        int u2 = 0;
        if (sscanf(p, "%lld %lld%n", &a, &b, &u2) != 2) return -2;
        p += u2;
        const char *q = skip_ws(p);
        float v = 1.0f;
        if (*q && *q != '\n') {
            double dv;
            int u3 = 0;
            if (sscanf(q, "%lf%n", &dv, &u3) == 1) {
                v = (float)dv;
                q += u3;
                if (k == 0) *weighted = 1;
            }
        }
        p = q;
        /* first column -> col, second -> row (util.cuh:382-386), 1-based */
        col[k] = (int32_t)(a - 1);
        row[k] = (int32_t)(b - 1);
        val[k] = v;
        ++k;
    }
    return k;
}

/* ======================================================================== */
/* Symmetrize / dedup / self-loop removal (makeSymmetric util.cuh:571-643,  */
/* with the self-loop removal that the reference leaves commented out and   */
/* the value array moved together with the indices).                       */
/* ======================================================================== */
typedef struct {
    int32_t r, c;
    int64_t ord;
    float v;
} or_trip;

static int trip_cmp(const void *x, const void *y) {
    const or_trip *a = (const or_trip *)x, *b = (const or_trip *)y;
    if (a->r != b->r) return a->r < b->r ? -1 : 1;
    if (a->c != b->c) return a->c < b->c ? -1 : 1;
    if (a->ord != b->ord) return a->ord < b->ord ? -1 : 1;
    return 0;
}

int64_t or_symmetrize(int64_t nnz, int32_t *row, int32_t *col, float *val) {
    or_trip *t = (or_trip *)malloc(sizeof(or_trip) * (size_t)(2 * nnz + 1));
    if (!t) return -1;
    int64_t k = 0;
    for (int64_t i = 0; i < nnz; ++i) {
        t[k].r = row[i]; t[k].c = col[i]; t[k].v = val[i]; t[k].ord = k; ++k;
    }
    for (int64_t i = 0; i < nnz; ++i) {
        if (row[i] == col[i]) continue; /* mirror only off-diagonal entries */
        t[k].r = col[i]; t[k].c = row[i]; t[k].v = val[i]; t[k].ord = k; ++k;
    }
    qsort(t, (size_t)k, sizeof(or_trip), trip_cmp);
    int64_t out = 0;
    for (int64_t i = 0; i < k; ++i) {
        if (t[i].r == t[i].c) continue;                       /* self-loop */
        if (out > 0 && row[out - 1] == t[i].r && col[out - 1] == t[i].c)
            continue;                                         /* duplicate */
        row[out] = t[i].r; col[out] = t[i].c; val[out] = t[i].v; ++out;
    }
    free(t);
    return out;
}

/* COO -> CSR counting sort (buildMatrix, matrix.cpp:83-105). */
int or_coo_to_csr(int32_t m, int64_t nnz, const int32_t *row, const int32_t *col,
                  const float *val, int64_t *rowptr, int32_t *col_out,
                  float *val_out) {
    for (int32_t i = 0; i <= m; ++i) rowptr[i] = 0;
    for (int64_t i = 0; i < nnz; ++i) {
        if (row[i] < 0 || row[i] >= m) return -1;
        rowptr[row[i] + 1]++;
    }
    for (int32_t i = 0; i < m; ++i) rowptr[i + 1] += rowptr[i];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    if (!fill) return -2;
    memcpy(fill, rowptr, sizeof(int64_t) * (size_t)(m + 1));
    for (int64_t i = 0; i < nnz; ++i) {
        int64_t d = fill[row[i]]++;
        col_out[d] = col[i];
        if (val_out) val_out[d] = val ? val[i] : 1.0f;
    }
    free(fill);
    return 0;
}

/* ======================================================================== */
/* Synthetic inputs (definitions owned by this build; the reference has no  */
/* generator).                                                              */
/* ======================================================================== */
int or_rmat_edges(int scale, int ef, double a, double b, double c, uint32_t seed,
                  int32_t *src, int32_t *dst) {
    or_mt19937 g;
    or_mt_seed(&g, seed);
    int64_t ne = (int64_t)ef << scale;
    double ab = a + b, abc = a + b + c;
    for (int64_t e = 0; e < ne; ++e) {
        int32_t u = 0, v = 0;
        for (int l = 0; l < scale; ++l) {
            double r = or_mt_res53(&g);
            int ub = 0, vb = 0;
            if (r < a) { ub = 0; vb = 0; }
            else if (r < ab) { ub = 0; vb = 1; }
            else if (r < abc) { ub = 1; vb = 0; }
            else { ub = 1; vb = 1; }
            u = (u << 1) | ub;
            v = (v << 1) | vb;
        }
        src[e] = u;
        dst[e] = v;
    }
    return 0;
}

/* open-addressing set of uint64 keys (for lattice adjacency tests) */
typedef struct {
    uint64_t *k;
    int64_t cap;
} or_set;

static uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

static int set_init(or_set *s, int64_t n) {
    int64_t cap = 16;
    while (cap < 2 * n + 16) cap <<= 1;
    s->k = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cap);
    if (!s->k) return -1;
    for (int64_t i = 0; i < cap; ++i) s->k[i] = ~0ULL;
    s->cap = cap;
    return 0;
}

/* returns 1 if inserted, 0 if present */
static int set_insert(or_set *s, uint64_t key) {
    int64_t h = (int64_t)(mix64(key) & (uint64_t)(s->cap - 1));
    while (s->k[h] != ~0ULL) {
        if (s->k[h] == key) return 0;
        h = (h + 1) & (s->cap - 1);
    }
    s->k[h] = key;
    return 1;
}

static int set_has(const or_set *s, uint64_t key) {
    int64_t h = (int64_t)(mix64(key) & (uint64_t)(s->cap - 1));
    while (s->k[h] != ~0ULL) {
        if (s->k[h] == key) return 1;
        h = (h + 1) & (s->cap - 1);
    }
    return 0;
}

int64_t or_lattice_edges(int32_t W, uint32_t seed, int32_t *eu, int32_t *ev) {
    int64_t n = (int64_t)W * W;
    int64_t E = 2LL * W * (W - 1);
    int32_t *lu = (int32_t *)malloc(sizeof(int32_t) * (size_t)E);
    int32_t *lv = (int32_t *)malloc(sizeof(int32_t) * (size_t)E);
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)E);
    unsigned char *gone = (unsigned char *)calloc((size_t)E, 1);
    if (!lu || !lv || !perm || !gone) return -1;
    int64_t k = 0;
    for (int64_t id = 0; id < n; ++id) {
        int64_t r = id / W, c = id % W;
        if (c < W - 1) { lu[k] = (int32_t)id; lv[k] = (int32_t)(id + 1); ++k; }
        if (r < W - 1) { lu[k] = (int32_t)id; lv[k] = (int32_t)(id + W); ++k; }
    }
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < E; ++i) perm[i] = i;
    for (int64_t i = E - 1; i >= 1; --i) {
        int64_t j = (int64_t)(or_mt_next(&g) % (uint32_t)(i + 1));
        int64_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    int64_t R = E / 10;       /* floor(0.10 * E) */
    int64_t L = E / 1000;     /* floor(0.001 * E) */
    for (int64_t i = 0; i < R; ++i) gone[perm[i]] = 1;
    or_set s;
    if (set_init(&s, E + L)) return -1;
    int64_t out = 0;
    for (int64_t i = 0; i < E; ++i) {
        if (gone[i]) continue;
        eu[out] = lu[i]; ev[out] = lv[i]; ++out;
        set_insert(&s, ((uint64_t)(uint32_t)lu[i] << 32) | (uint32_t)lv[i]);
    }
    int64_t added = 0;
    while (added < L) {
        uint32_t u = or_mt_next(&g) % (uint32_t)n;
        uint32_t v = or_mt_next(&g) % (uint32_t)n;
        if (u == v) continue;
        uint32_t a = u < v ? u : v, b = u < v ? v : u;
        uint64_t key = ((uint64_t)a << 32) | b;
        if (set_has(&s, key)) continue;
        set_insert(&s, key);
        eu[out] = (int32_t)a; ev[out] = (int32_t)b; ++out;
        ++added;
    }
    free(s.k); free(lu); free(lv); free(perm); free(gone);
    return out;
}

static int64_t find_in_row(const int64_t *rowptr, const int32_t *col, int32_t r,
                           int32_t c) {
    int64_t lo = rowptr[r], hi = rowptr[r + 1];
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (col[mid] < c) lo = mid + 1; else hi = mid;
    }
    return (lo < rowptr[r + 1] && col[lo] == c) ? lo : -1;
}

int or_assign_weights(int32_t m, const int64_t *rowptr, const int32_t *col,
                      uint32_t seed, float *val) {
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (int32_t u = 0; u < m; ++u) {
        for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            int32_t v = col[e];
            if (v <= u) continue;
            float w = (float)(1u + or_mt_next(&g) % 255u);
            val[e] = w;
            int64_t r = find_in_row(rowptr, col, v, u);
            if (r < 0) return -1; /* not symmetric */
            val[r] = w;
        }
    }
    return 0;
}

/* ======================================================================== */
/* BFS: restates SimpleReferenceBfs (testBfs.cpp:6-72) with stop = infinity */
/* ======================================================================== */
int32_t or_bfs(int32_t m, const int64_t *rowptr, const int32_t *col,
               int32_t src, int32_t *depth) {
    for (int32_t i = 0; i < m; ++i) depth[i] = -1;
    if (src < 0 || src >= m) return -1;
    int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
    if (!queue) return -2;
    int64_t head = 0, tail = 0;
    int32_t search_depth = 0;
    depth[src] = 0;
    queue[tail++] = src;
    while (head < tail) {
        int32_t u = queue[head++];
        int32_t nd = depth[u] + 1;
        for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            int32_t v = col[e];
            if (depth[v] == -1) {
                depth[v] = nd;
                if (nd > search_depth) search_depth = nd;
                queue[tail++] = v;
            }
        }
    }
    free(queue);
    return search_depth + 1; /* testBfs.cpp:66 */
}

/* ======================================================================== */
/* SSSP: restates SimpleReferenceSssp (testSssp.cpp:33-115).  A binary      */
/* min-heap of (vertex, key); all vertices are pushed up front with a junk  */
/* key (the reference uses h_csrValA[i], :51-53), the source with key 0.    */
/* A popped vertex relaxes its edges only if it has been reached (:86);     */
/* relaxations use the current tentative distance (:90), so the loop is     */
/* label-correcting and ends at the shortest-path fixed point.              */
/* ======================================================================== */
typedef struct {
    int32_t v;
    float d;
} or_hent;

typedef struct {
    or_hent *a;
    int64_t n, cap;
} or_heap;

static int heap_push(or_heap *h, int32_t v, float d) {
    if (h->n == h->cap) {
        int64_t nc = h->cap ? 2 * h->cap : 1024;
        or_hent *na = (or_hent *)realloc(h->a, sizeof(or_hent) * (size_t)nc);
        if (!na) return -1;
        h->a = na; h->cap = nc;
    }
    int64_t i = h->n++;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (h->a[p].d <= d) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i].v = v; h->a[i].d = d;
    return 0;
}

static or_hent heap_pop(or_heap *h) {
    or_hent top = h->a[0];
    or_hent last = h->a[--h->n];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, s = i;
        float sd = last.d;
        if (l < h->n && h->a[l].d < sd) { s = l; sd = h->a[l].d; }
        if (r < h->n && h->a[r].d < sd) { s = r; }
        if (s == i) break;
        h->a[i] = h->a[s];
        i = s;
    }
    if (h->n > 0) h->a[i] = last;
    return top;
}

int32_t or_sssp(int32_t m, const int64_t *rowptr, const int32_t *col,
                const float *w, int32_t src, float *dist) {
    if (src < 0 || src >= m) return -1;
    int64_t nnz = rowptr[m];
    or_heap h = {0, 0, 0};
    for (int32_t i = 0; i < m; ++i) {
        dist[i] = -1.0f;
        /* junk key as in testSssp.cpp:51-53; guarded when nnz < m */
        if (i != src && heap_push(&h, i, i < nnz ? w[i] : 0.0f)) return -2;
    }
    dist[src] = 0.0f;
    if (heap_push(&h, src, 0.0f)) return -2;
    while (h.n > 0) {
        or_hent t = heap_pop(&h);
        int32_t v = t.v;
        if (dist[v] == -1.0f) continue; /* never reached yet: skip (:86) */
        for (int64_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            int32_t nb = col[e];
            float alt = dist[v] + w[e];
            if (dist[nb] == -1.0f || alt < dist[nb]) {
                dist[nb] = alt;
                if (heap_push(&h, nb, alt)) return -2;
            }
        }
    }
    free(h.a);
    return 0;
}

/* ======================================================================== */
/* MIS                                                                       */
/* ======================================================================== */
static uint64_t mis_key(const uint32_t *prio, int32_t v) {
    /* larger key wins: priority first, then the smaller vertex id; the low
     * word stays below 2^31 so key + 1 never wraps */
    return ((uint64_t)prio[v] << 32) | (uint64_t)(0x7fffffffu - (uint32_t)v);
}

typedef struct {
    uint64_t key;
    int32_t v;
} or_kv;

static int kv_desc(const void *x, const void *y) {
    const or_kv *a = (const or_kv *)x, *b = (const or_kv *)y;
    if (a->key != b->key) return a->key > b->key ? -1 : 1;
    return 0;
}

int or_mis_greedy(int32_t m, const int64_t *rowptr, const int32_t *col,
                  const uint32_t *prio, int32_t *in_set) {
    or_kv *order = (or_kv *)malloc(sizeof(or_kv) * (size_t)(m ? m : 1));
    if (!order) return -1;
    for (int32_t v = 0; v < m; ++v) { order[v].key = mis_key(prio, v); order[v].v = v; }
    qsort(order, (size_t)m, sizeof(or_kv), kv_desc);
    for (int32_t v = 0; v < m; ++v) in_set[v] = -1;
    for (int32_t i = 0; i < m; ++i) {
        int32_t v = order[i].v;
        if (in_set[v] != -1) continue;
        in_set[v] = 1;
        for (int64_t e = rowptr[v]; e < rowptr[v + 1]; ++e)
            if (col[e] != v) in_set[col[e]] = 0;
    }
    free(order);
    return 0;
}

int or_mis_luby(int32_t m, const int64_t *rowptr, const int32_t *col,
                const uint32_t *prio, float delta, int32_t *in_set) {
    /* in_set: -1 undecided, 1 in, 0 out (spmspvMis.cuh:201-204) */
    unsigned char *cand = (unsigned char *)malloc((size_t)(m ? m : 1));
    uint64_t *best = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(m ? m : 1));
    if (!cand || !best) return -1;
    for (int32_t v = 0; v < m; ++v) in_set[v] = -1;
    double minimum = 1.0;
    int rounds = 0;
    for (int iter = 1; iter < 1000000; ++iter) {
        if (minimum >= 0) minimum -= delta;                     /* :235-236 */
        int64_t nc = 0;
        for (int32_t v = 0; v < m; ++v) {                       /* buildVector :53-61 */
            double p = (double)prio[v] / 4294967296.0;
            cand[v] = (in_set[v] == -1 && p > minimum) ? 1 : 0;
            nc += cand[v];
            best[v] = 0;                                        /* identity */
        }
        if (nc == 0) {
            if (minimum <= 0) break;                            /* :316-317 */
            continue;
        }
        ++rounds;
        /* expand candidates, max-reduce their keys onto neighbours (:258-293) */
        for (int32_t u = 0; u < m; ++u) {
            if (!cand[u]) continue;
            uint64_t ku = mis_key(prio, u) + 1; /* +1 keeps 0 as "no candidate" */
            for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
                int32_t k = col[e];
                if (k != u && ku > best[k]) best[k] = ku;
            }
        }
        /* updateMis (:63-68): a candidate joins if it beats every candidate nb */
        for (int32_t v = 0; v < m; ++v)
            if (cand[v] && mis_key(prio, v) + 1 > best[v]) in_set[v] = 1;
        /* updateNeighbor (:70-78): neighbours of joiners leave */
        for (int32_t u = 0; u < m; ++u) {
            if (!cand[u] || in_set[u] != 1) continue;
            for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e)
                if (col[e] != u) in_set[col[e]] = 0;
        }
    }
    free(cand); free(best);
    return rounds;
}

/* ======================================================================== */
/* Triangle counting: semantics of ewiseMultTc (tc.cuh:168-199) over the    */
/* lower-triangular operands built by buildLower (test/tc/main.cu:39-55).   */
/* ======================================================================== */
uint64_t or_tc(int32_t m, const int64_t *rowptr, const int32_t *col,
               int32_t *per_row) {
    uint64_t total = 0;
    for (int32_t r = 0; r < m; ++r) {
        int64_t rb = rowptr[r], re = rowptr[r + 1];
        int64_t rl = rb;
        while (rl < re && col[rl] < r) ++rl; /* L_r = [rb, rl) */
        int64_t cnt = 0;
        for (int64_t e = rb; e < rl; ++e) {
            int32_t k = col[e];
            /* intersect L_r (cols < k suffice) with L_k */
            int64_t i = rb, j = rowptr[k], je = rowptr[k + 1];
            while (i < rl && j < je && col[i] < k && col[j] < k) {
                if (col[i] == col[j]) { ++cnt; ++i; ++j; }
                else if (col[i] < col[j]) ++i;
                else ++j;
            }
        }
        if (per_row) per_row[r] = (int32_t)cnt;
        total += (uint64_t)cnt;
    }
    return total;
}

/* ======================================================================== */
/* Connected components (union-find with path halving)                      */
/* ======================================================================== */
static int32_t uf_find(int32_t *p, int32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

int or_components(int32_t m, const int64_t *rowptr, const int32_t *col,
                  int32_t *label) {
    for (int32_t v = 0; v < m; ++v) label[v] = v;
    for (int32_t u = 0; u < m; ++u)
        for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            int32_t a = uf_find(label, u), b = uf_find(label, col[e]);
            if (a == b) continue;
            if (a < b) label[b] = a; else label[a] = b;
        }
    for (int32_t v = 0; v < m; ++v) label[v] = uf_find(label, v);
    return 0;
}

/* ======================================================================== */
/* SpMSpV y = A^T x under a semiring (sparse mXv, mXv.cuh:249-389).        */
/* ======================================================================== */
int64_t or_spmspv(int32_t m, const int64_t *rowptr, const int32_t *col,
                  const float *aval, int semiring, int64_t nx,
                  const int32_t *x_ind, const float *x_val, const int32_t *mask,
                  int32_t *y_ind, float *y_val) {
    float *acc = (float *)malloc(sizeof(float) * (size_t)(m ? m : 1));
    unsigned char *hit = (unsigned char *)calloc((size_t)(m ? m : 1), 1);
    if (!acc || !hit) return -1;
    for (int64_t i = 0; i < nx; ++i) {
        int32_t u = x_ind[i];
        float xu = x_val ? x_val[i] : 1.0f;
        for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            int32_t v = col[e];
            float a = aval ? aval[e] : 1.0f;
            float t;
            switch (semiring) {
            case OR_SR_PLUS_TIMES: t = xu * a; break;  /* elementMult mXv.cuh:105 */
            case OR_SR_BOOL: t = 1.0f; break;
            case OR_SR_MIN_PLUS: t = xu + a; break;    /* spmspvSssp.cuh:86-89 */
            default: t = xu; break;                    /* gather prio, :281 */
            }
            if (!hit[v]) { hit[v] = 1; acc[v] = t; continue; }
            switch (semiring) {
            case OR_SR_PLUS_TIMES: acc[v] = acc[v] + t; break;
            case OR_SR_BOOL: acc[v] = 1.0f; break;
            case OR_SR_MIN_PLUS: if (t < acc[v]) acc[v] = t; break;
            default: if (t > acc[v]) acc[v] = t; break;
            }
        }
    }
    int64_t ny = 0;
    for (int32_t v = 0; v < m; ++v) {
        if (!hit[v]) continue;
        if (mask && mask[v]) continue;
        y_ind[ny] = v;
        y_val[ny] = acc[v];
        ++ny;
    }
    free(acc); free(hit);
    return ny;
}

/* ======================================================================== */
/* Gustavson SpGEMM with a dense accumulator                                 */
/* ======================================================================== */
static int i32_cmp(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return a < b ? -1 : (a > b);
}

int64_t or_spgemm(int32_t m, int32_t k, int32_t n,
                  const int64_t *a_rowptr, const int32_t *a_col, const float *a_val,
                  const int64_t *b_rowptr, const int32_t *b_col, const float *b_val,
                  int64_t *c_rowptr, int32_t *c_col, float *c_val) {
    (void)k;
    float *acc = (float *)calloc((size_t)(n ? n : 1), sizeof(float));
    int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    if (!acc || !mark || !touched) return -1;
    for (int32_t j = 0; j < n; ++j) mark[j] = -1;
    int64_t nnz = 0;
    c_rowptr[0] = 0;
    for (int32_t i = 0; i < m; ++i) {
        int32_t nt = 0;
        for (int64_t e = a_rowptr[i]; e < a_rowptr[i + 1]; ++e) {
            int32_t kk = a_col[e];
            float av = a_val ? a_val[e] : 1.0f;
            for (int64_t f = b_rowptr[kk]; f < b_rowptr[kk + 1]; ++f) {
                int32_t j = b_col[f];
                float bv = b_val ? b_val[f] : 1.0f;
                if (mark[j] != i) { mark[j] = i; acc[j] = 0.0f; touched[nt++] = j; }
                acc[j] += av * bv;
            }
        }
        qsort(touched, (size_t)nt, sizeof(int32_t), i32_cmp);
        if (c_col) {
            for (int32_t t = 0; t < nt; ++t) {
                c_col[nnz + t] = touched[t];
                if (c_val) c_val[nnz + t] = acc[touched[t]];
            }
        }
        nnz += nt;
        c_rowptr[i + 1] = nnz;
    }
    free(acc); free(mark); free(touched);
    return nnz;
}

/* ======================================================================== */
/* Column-by-column SpMSpM (mXm, tc.cuh:314-375): for each column j of B,    */
/* one sparse mXv (mXv.cuh:249-389) of A in CSC with that column as x:       */
/* expand every (k, B[k,j]) over column k of A into (row r, B[k,j]*A[r,k])  */
/* in column order (IntervalExpand/Gather, :300-307; elementMult :105), a    */
/* stable sort of the pairs by row (SortPairs, :324) and a reduce by key     */
/* with + (ReduceByKey, :332) -- a key whose products cancel stays, with     */
/* value 0.  Column j of C holds that column's output (the running total_nnz */
/* of tc.cuh:352-355 is the prefix of the per-column output counts).        */
/* The sums run sequentially in sorted order; the reference's segmented      */
/* reduce associates differently, which is exact on integer-valued inputs.   */
/* ======================================================================== */
typedef struct { int32_t row; int32_t pos; float val; } or_prod;

static int prod_cmp(const void *x, const void *y) {
    const or_prod *a = (const or_prod *)x, *b = (const or_prod *)y;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    return a->pos < b->pos ? -1 : (a->pos > b->pos); /* stable: keep expansion order */
}

int64_t or_mxm(int32_t p, int32_t q, int32_t r,
               const int64_t *a_colptr, const int32_t *a_row, const float *a_val,
               const int64_t *b_colptr, const int32_t *b_row, const float *b_val,
               int64_t *c_colptr, int32_t *c_row, float *c_val) {
    (void)p; (void)q;
    int64_t cap = 1;
    or_prod *buf = (or_prod *)malloc(sizeof(or_prod));
    if (!buf) return -1;
    int64_t nnz = 0;
    c_colptr[0] = 0;
    for (int32_t j = 0; j < r; ++j) {
        int64_t total = 0;
        for (int64_t f = b_colptr[j]; f < b_colptr[j + 1]; ++f) {
            const int32_t k = b_row[f];
            total += a_colptr[k + 1] - a_colptr[k];
        }
        if (total > cap) {
            free(buf);
            cap = total;
            buf = (or_prod *)malloc(sizeof(or_prod) * (size_t)cap);
            if (!buf) return -1;
        }
        int64_t t = 0;
        for (int64_t f = b_colptr[j]; f < b_colptr[j + 1]; ++f) {
            const int32_t k = b_row[f];
            const float xv = b_val ? b_val[f] : 1.0f;
            for (int64_t e = a_colptr[k]; e < a_colptr[k + 1]; ++e) {
                buf[t].row = a_row[e];
                buf[t].pos = (int32_t)t;
                buf[t].val = xv * (a_val ? a_val[e] : 1.0f);
                ++t;
            }
        }
        qsort(buf, (size_t)t, sizeof(or_prod), prod_cmp);
        for (int64_t s = 0; s < t;) {
            int64_t e = s + 1;
            float sum = buf[s].val;
            while (e < t && buf[e].row == buf[s].row) sum = sum + buf[e++].val;
            if (c_row) {
                c_row[nnz] = buf[s].row;
                if (c_val) c_val[nnz] = sum;
            }
            ++nnz;
            s = e;
        }
        c_colptr[j + 1] = nnz;
    }
    free(buf);
    return nnz;
}

/* ======================================================================== */
/* Row-block DCSC (csr_to_dcsc, matrix.cpp:286-468), restated step by step: */
/* part offsets from the row offsets (:313-315), a per-part stable sort of   */
/* the entries by column (:327, :413), segment ends where the column changes */
/* (lookRight mXv.cuh:36-41) or a part ends (specialScatter spgemm.cuh:      */
/* 34-38), an exclusive scan of the ends (:349), their positions (:351) and  */
/* columns (:358), ends turned into starts (shiftRightAdd spgemm.cuh:27-31,  */
/* :359-361), part offsets read from the scan at each part start (:369-371)  */
/* and the rows / values gathered through the sort's permutation (:405-424). */
/* Parts that start at or past the last entry read the segment count.        */
/* ======================================================================== */
typedef struct { int32_t key; int32_t pos; } or_kp;

static int kp_cmp(const void *x, const void *y) {
    const or_kp *a = (const or_kp *)x, *b = (const or_kp *)y;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->pos < b->pos ? -1 : (a->pos > b->pos);
}

int64_t or_dcsc(int32_t m, const int64_t *rowptr, const int32_t *col, const float *val,
                int32_t ps, int32_t pn, int64_t *part_ptr, int32_t *col_ind,
                int64_t *col_off, int32_t *row_ind, float *val_out) {
    const int64_t nnz = rowptr[m];
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * ((size_t)pn + 1));
    or_kp *kp = (or_kp *)malloc(sizeof(or_kp) * (size_t)(nnz ? nnz : 1));
    int32_t *rows = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    unsigned char *end = (unsigned char *)calloc((size_t)(nnz ? nnz : 1), 1);
    int64_t *scan = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nnz + 1));
    if (!off || !kp || !rows || !end || !scan) return -1;
    for (int32_t p = 0; p < pn; ++p) off[p] = rowptr[(int64_t)p * ps];
    off[pn] = nnz;
    for (int32_t r = 0; r < m; ++r)
        for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) rows[e] = r;
    for (int64_t e = 0; e < nnz; ++e) { kp[e].key = col[e]; kp[e].pos = (int32_t)e; }
    for (int32_t p = 0; p < pn; ++p)
        qsort(kp + off[p], (size_t)(off[p + 1] - off[p]), sizeof(or_kp), kp_cmp);
    for (int64_t i = 0; i + 1 < nnz; ++i) end[i] = kp[i].key != kp[i + 1].key;
    for (int32_t p = 1; p <= pn; ++p)
        if (off[p] > 0) end[off[p] - 1] = 1;
    scan[0] = 0;
    for (int64_t i = 0; i < nnz; ++i) scan[i + 1] = scan[i] + end[i];
    const int64_t L = scan[nnz];
    int64_t k = 0;
    col_off[0] = 0;
    for (int64_t i = 0; i < nnz; ++i) {
        if (!end[i]) continue;
        col_ind[k] = kp[i].key;
        col_off[k + 1] = i + 1;
        ++k;
    }
    for (int32_t p = 0; p < pn; ++p) part_ptr[p] = scan[off[p]];
    part_ptr[pn] = L;
    for (int64_t i = 0; i < nnz; ++i) {
        row_ind[i] = rows[kp[i].pos];
        if (val_out) val_out[i] = val ? val[kp[i].pos] : 1.0f;
    }
    free(off); free(kp); free(rows); free(end); free(scan);
    return L;
}

/* ======================================================================== */
/* Result invariants (checks, not restatements)                             */
/* ======================================================================== */
/* Invariants that fix a BFS depth array exactly on a graph whose CSR equals
 * its transpose (the depth semantics of testBfs.cpp:6-72): d[src] = 0,
 * every other entry >= -1; every edge joins two reached or two unreached
 * vertices and spans at most one level; every reached v != src has a
 * neighbour one level up.  By induction on the level these leave exactly
 * one array.  Returns 0, or the number of the first failed check. */
int32_t or_check_bfs(int32_t m, const int64_t *rowptr, const int32_t *col, const int32_t *d,
                     int32_t src) {
    if (src < 0 || src >= m || d[src] != 0) return 1;
    uint8_t *par = (uint8_t *)calloc((size_t)(m ? m : 1), 1);
    if (!par) return -1;
    int32_t bad = 0;
    for (int32_t v = 0; v < m && !bad; ++v) {
        const int32_t dv = d[v];
        if (dv < -1) { bad = 5; break; }
        for (int64_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            const int32_t u = col[e], du = d[u];
            if ((dv >= 0) != (du >= 0)) { bad = 2; break; }
            if (dv >= 0 && (du > dv + 1 || dv > du + 1)) { bad = 3; break; }
            if (dv >= 0 && du == dv + 1) par[u] = 1; /* v is one level above u */
        }
    }
    for (int32_t v = 0; v < m && !bad; ++v)
        if (d[v] > 0 && !par[v]) bad = 4;
    free(par);
    return bad;
}

/* Invariants that fix SSSP distances exactly for positive weights (the
 * semantics of testSssp.cpp:33-115 in FP32): d[src] = 0; every reached u
 * and edge (u, v) satisfy d[v] <= fl32(d[u] + w) (so d is at most the
 * shortest left-fold sum, and unreached vertices have no reached
 * in-neighbour); every reached v != src is tight on some in-edge,
 * d[v] == fl32(d[u] + w) (following tight edges strictly lowers d, so d[v]
 * is the sum of a real path and at least the shortest).  Unreached = inf
 * (the caller's sentinel).  Returns 0, or the number of the first failed
 * check. */
int32_t or_check_sssp(int32_t m, const int64_t *rowptr, const int32_t *col, const float *w,
                      const float *d, int32_t src, float inf) {
    if (src < 0 || src >= m || d[src] != 0.0f) return 1;
    uint8_t *tight = (uint8_t *)calloc((size_t)(m ? m : 1), 1);
    if (!tight) return -1;
    int32_t bad = 0;
    for (int32_t u = 0; u < m && !bad; ++u) {
        const float du = d[u];
        if (!(du >= 0.0f)) { bad = 4; break; }
        if (du == inf) continue;
        for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            const int32_t v = col[e];
            const float c = du + w[e];
            if (d[v] > c) { bad = 2; break; }
            if (d[v] == c) tight[v] = 1;
        }
    }
    for (int32_t v = 0; v < m && !bad; ++v)
        if (v != src && d[v] != inf && !tight[v]) bad = 3;
    free(tight);
    return bad;
}
