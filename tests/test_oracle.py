"""CPU tests that pin the oracle (oracle/) before anything is compared to it.

Each oracle function is checked against something it was not derived from:
the reference's own mt19937ar.cpp compiled into oracle/_ref, the reference's
MatrixMarket fixtures with known answers, independent library routines
(numpy's MT19937, scipy.sparse.csgraph, networkx, dense numpy products) and
closed forms.  A few tests plant a mistake in an oracle copy to show that the
checks would catch it.
"""
import hashlib
import json
from itertools import combinations
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

from oracle import oracle as O

GOLD = json.loads((Path(__file__).parent / "golden" / "golden.json").read_text())
MTX = Path(__file__).parent / "golden" / "mtx"


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def to_scipy(m, rowptr, col, val=None):
    v = np.ones(len(col)) if val is None else val.astype(np.float64)
    return sp.csr_matrix((v, col, rowptr), shape=(m, m))


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    a = rng.random((n, n)) < p
    a = np.triu(a, 1)
    u, v = np.nonzero(a)
    m, rp, c = (n,) + O.edges_to_csr(n, u.astype(np.int32), v.astype(np.int32))
    return m, rp, c


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


# ------------------------------------------------------------------ MT19937
def test_mt_known_answers():
    g = GOLD["mt19937"]
    s = O.mt_stream(5489, 10000)
    assert s[0] == g["seed5489_first"] == 3499211612
    assert s[-1] == g["seed5489_10000th"] == 4123659995  # C++ [rand.predef]
    assert O.mt_stream(1, 8).tolist() == g["seed1_first8"]
    assert O.mt_stream(2, 8).tolist() == g["seed2_first8"]
    assert O.mt_res53_stream(1, 4).tolist() == g["res53_seed1_first4"]


def test_mt_matches_reference_build():
    ref = O.ref_mt()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    init, i32, r53 = ref
    for seed in (5489, 1, 2, 3, 123456789):
        init(seed)
        want = [i32() for _ in range(2000)]
        assert O.mt_stream(seed, 2000).tolist() == want
        init(seed)
        want = [r53() for _ in range(700)]
        assert O.mt_res53_stream(seed, 700).tolist() == want


# ---------------------------------------------------------- graph building
@pytest.mark.parametrize("name", sorted(GOLD["fixtures"]))
def test_fixture_graph_build(name):
    g = GOLD["fixtures"][name]
    m, rp, c, _ = O.mtx_graph((MTX / f"{name}.mtx").read_text())
    assert m == g["n"] and len(c) == g["nnz"]
    assert rp.tolist() == g["rowptr"] and c.tolist() == g["col"]


def test_mtx_column_swap_and_values():
    txt = "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 2 5.5\n3 1 2\n"
    m, n, row, col, val, k, weighted = O.parse_mtx(txt)
    assert (m, n, k, weighted) == (3, 3, 2, True)
    # first index -> col, second -> row (reference readMtx)
    assert col[:2].tolist() == [0, 2] and row[:2].tolist() == [1, 0]
    assert val[:2].tolist() == [5.5, 2.0]


def test_symmetrize_dedup_selfloops_and_values():
    row = np.zeros(20, np.int32); col = np.zeros(20, np.int32); val = np.zeros(20, np.float32)
    edges = [(0, 1, 3.0), (1, 0, 7.0), (2, 2, 9.0), (1, 2, 4.0), (0, 1, 8.0)]
    for i, (r, c, v) in enumerate(edges):
        row[i], col[i], val[i] = r, c, v
    r, c, v = O.symmetrize(row, col, val, len(edges))
    got = list(zip(r.tolist(), c.tolist(), v.tolist()))
    # (0,1) keeps its first value 3, (1,0) its first value 7, no (2,2)
    assert got == [(0, 1, 3.0), (1, 0, 7.0), (1, 2, 4.0), (2, 1, 4.0)]


def test_coo_to_csr_keeps_row_order():
    row = np.array([2, 0, 2, 1, 0], np.int32)
    col = np.array([5, 4, 1, 3, 2], np.int32)
    rp, c, _ = O.coo_to_csr(3, row, col)
    assert rp.tolist() == [0, 2, 3, 5]
    assert c.tolist() == [4, 2, 3, 5, 1]


def test_rmat_generator_known_answer():
    g = GOLD["rmat10"]
    s, d = O.rmat_edges(10)
    assert sha(s) == g["edges_u_sha"] and sha(d) == g["edges_v_sha"]
    assert [[int(a), int(b)] for a, b in zip(s[:8], d[:8])] == g["first_edges"]
    m, rp, c = O.rmat_graph(10)
    assert sha(rp.astype(np.int32)) == g["rowptr_sha"] and sha(c) == g["col_sha"]


def test_rmat_quadrant_walk_against_reference_rng():
    ref = O.ref_mt()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    init, _, r53 = ref
    init(7)
    s, d = O.rmat_edges(6, ef=2, seed=7)
    for e in range(len(s)):
        u = v = 0
        for _ in range(6):
            r = r53()
            ub = 1 if r >= 0.57 + 0.19 else 0
            vb = 1 if (0.57 <= r < 0.57 + 0.19) or r >= 0.57 + 0.19 + 0.19 else 0
            u, v = 2 * u + ub, 2 * v + vb
        assert (s[e], d[e]) == (u, v)


@pytest.mark.parametrize("W", [32, 64])
def test_lattice_generator_known_answer(W):
    g = GOLD[f"lattice{W}"]
    eu, ev = O.lattice_edges(W)
    assert len(eu) == g["n_edges"]
    assert sha(eu) == g["edges_u_sha"] and sha(ev) == g["edges_v_sha"]
    E = 2 * W * (W - 1)
    assert len(eu) == E - E // 10 + E // 1000


def test_weights_known_answer_and_symmetry():
    g = GOLD["rmat10"]
    m, rp, c = O.rmat_graph(10)
    w = O.assign_weights(m, rp, c)
    assert sha(w) == g["weights_sha"]
    A = to_scipy(m, rp, c, w)
    assert (A != A.T).nnz == 0
    assert w.min() >= 1 and w.max() <= 255 and np.all(w == np.round(w))


# ---------------------------------------------------------------- BFS / SSSP
@pytest.mark.parametrize("name", sorted(GOLD["fixtures"]))
def test_bfs_fixture(name):
    g = GOLD["fixtures"][name]
    m, rp, c, _ = O.mtx_graph((MTX / f"{name}.mtx").read_text())
    depth, sd = O.bfs(m, rp, c, 0)
    assert depth.tolist() == g["bfs"]
    assert sd == max(g["bfs"]) + 1  # search depth = levels, testBfs.cpp:66


def test_bfs_survey_kats():
    # networkx-derived answers recorded in the survey of the reference
    kat = {"test_bc": [0, 1, 1, 1, 2, 2, 2], "test_cc": [0, 1, 1, 1, 2, 2, 2, -1, -1, -1, -1],
           "test_mesh": [0, 1, 2, 1, 2, 3, 2, 3, 4], "test_mis": [0, 1, 1, 2, 2, 3, 1, 2],
           "test_pr": [0, 1, 1, 1], "test_spgemm": [0, 1, 1, 1]}
    for name, want in kat.items():
        m, rp, c, _ = O.mtx_graph((MTX / f"{name}.mtx").read_text())
        assert O.bfs(m, rp, c, 0)[0].tolist() == want


@pytest.mark.parametrize("key", ["rmat10", "rmat14", "lattice32", "lattice64"])
def test_bfs_sssp_tc_mis_synthetic(key):
    g = GOLD[key]
    if key.startswith("rmat"):
        m, rp, c = O.rmat_graph(int(key[4:]))
    else:
        m, rp, c = O.lattice_graph(int(key[7:]))
    assert sha(rp.astype(np.int32)) == g["rowptr_sha"] and sha(c) == g["col_sha"]
    depth, _ = O.bfs(m, rp, c, g["source"])
    assert sha(depth) == g["bfs_sha"]
    w = O.assign_weights(m, rp, c)
    dist = O.sssp(m, rp, c, w, g["source"])
    assert sha(dist) == g["sssp_sha"]
    total, rows = O.tc(m, rp, c)
    assert total == g["triangles"] and sha(rows) == g["tc_rows_sha"]
    mis = O.mis_greedy(m, rp, c, O.mis_priorities(m))
    assert sha(mis) == g["mis_sha"]
    lab = O.components(m, rp, c)
    assert len(np.unique(lab)) == g["components"]


def test_bfs_closed_forms():
    # path graph: depth = |i - src|
    n = 50
    m, rp, c = (n,) + O.edges_to_csr(n, np.arange(n - 1, dtype=np.int32),
                                     np.arange(1, n, dtype=np.int32))
    assert O.bfs(m, rp, c, 17)[0].tolist() == [abs(i - 17) for i in range(n)]
    # full lattice (no removals): Manhattan distance from the corner
    W = 9
    u = [i for i in range(W * W) if i % W < W - 1] + [i for i in range(W * W - W)]
    v = [i + 1 for i in range(W * W) if i % W < W - 1] + [i + W for i in range(W * W - W)]
    m, rp, c = (W * W,) + O.edges_to_csr(W * W, np.array(u, np.int32), np.array(v, np.int32))
    assert O.bfs(m, rp, c, 0)[0].tolist() == [i // W + i % W for i in range(W * W)]


@pytest.mark.parametrize("seed", range(6))
def test_bfs_sssp_vs_scipy_random(seed):
    m, rp, c = random_graph(120, 0.03, seed)
    w = O.assign_weights(m, rp, c, seed=seed + 10)
    A = to_scipy(m, rp, c)
    for src in (0, 7, 119):
        d = csg.shortest_path(A, unweighted=True, indices=src)
        assert O.bfs(m, rp, c, src)[0].tolist() == np.where(np.isinf(d), -1, d).astype(int).tolist()
        d = csg.dijkstra(to_scipy(m, rp, c, w), indices=src)
        assert O.sssp(m, rp, c, w, src).tolist() == np.where(np.isinf(d), -1, d).tolist()


def test_sssp_invariants():
    m, rp, c = O.lattice_graph(32)
    w = O.assign_weights(m, rp, c)
    d = O.sssp(m, rp, c, w, 5)
    rows = np.repeat(np.arange(m), np.diff(rp))
    reach = d[rows] >= 0
    # triangle inequality on every edge, and each reached v != src is tight on some edge
    assert np.all(d[c[reach]] <= d[rows[reach]] + w[reach])
    tight = np.zeros(m, bool)
    tight[c[reach & (d[c] == d[rows] + w)]] = True
    assert np.all(tight[(d > 0)])


def test_result_checkers_accept_exact_and_reject_planted_errors():
    """check_bfs / check_sssp (the full-size parity checks): the oracle's
    answers pass; each planted mistake -- a depth or distance one off in
    either direction, a reached vertex reported unreached, an unreached one
    reported reached -- fails."""
    rng = np.random.default_rng(8)
    for m, rp, c in (O.rmat_graph(10), O.lattice_graph(40)):
        w = O.assign_weights(m, rp, c)
        deg = np.diff(rp)
        for src in (int(np.argmax(deg)), int(np.nonzero(deg)[0][-1])):
            d, _ = O.bfs(m, rp, c, src)
            assert O.check_bfs(m, rp, c, d, src) == 0
            reached = np.nonzero(d > 0)[0]
            for delta in (1, -1):
                bad = d.copy()
                bad[rng.choice(reached)] += delta
                assert O.check_bfs(m, rp, c, bad, src) != 0
            bad = d.copy()
            bad[rng.choice(reached)] = -1
            assert O.check_bfs(m, rp, c, bad, src) != 0
            s = O.sssp(m, rp, c, w, src)
            s = np.where(s == -1, np.float32(1.70141e38), s).astype(np.float32)
            assert O.check_sssp(m, rp, c, w, s, src) == 0
            for delta in (1.0, -1.0):
                bad = s.copy()
                bad[rng.choice(reached)] += delta
                assert O.check_sssp(m, rp, c, w, bad, src) != 0
            bad = s.copy()
            bad[rng.choice(reached)] = np.float32(1.70141e38)
            assert O.check_sssp(m, rp, c, w, bad, src) != 0
    # an isolated vertex reported reached
    m = 6
    rp, c = O.edges_to_csr(m, np.array([0, 1], np.int32), np.array([1, 2], np.int32))
    d, _ = O.bfs(m, rp, c, 0)
    assert O.check_bfs(m, rp, c, d, 0) == 0
    bad = d.copy()
    bad[5] = 3
    assert O.check_bfs(m, rp, c, bad, 0) != 0


@pytest.mark.parametrize("lo,hi", [(1, 10 ** 8), (10 ** 9, 10 ** 10)])
def test_sssp_fp32_sums_with_large_weights(lo, hi):
    """Weights whose sums pass 2^24 round in FP32.  The oracle's distances
    are then the smallest left-fold FP32 path sums: they equal the fixed
    point of d[v] = min(d[v], fl32(d[u] + w)) iterated over all edges from
    d[src] = 0 (a vectorised Bellman-Ford in numpy float32, independent of
    the heap order), and they differ from the float64 shortest paths."""
    rng = np.random.default_rng(lo % 1000)
    m, rp, c = O.rmat_graph(8)
    w = rng.integers(lo, hi + 1, len(c)).astype(np.float32)
    rows = np.repeat(np.arange(m), np.diff(rp))
    differs = 0
    for src in (0, 17, 100):
        d = O.sssp(m, rp, c, w, src)
        bf = np.full(m, np.inf, np.float32)
        bf[src] = 0
        while True:
            cand = (bf[rows] + w).astype(np.float32)
            nb = bf.copy()
            np.minimum.at(nb, c, cand)
            if np.array_equal(nb, bf):
                break
            bf = nb
        want = np.where(np.isinf(bf), np.float32(-1), bf)
        assert np.array_equal(d, want), src
        import scipy.sparse as sp
        import scipy.sparse.csgraph as cg
        A64 = sp.csr_matrix((w.astype(np.float64), c, rp), shape=(m, m))
        d64 = cg.dijkstra(A64, indices=src)
        differs += int(np.sum(np.isfinite(d64) & (d64 != d.astype(np.float64))))
    assert differs > 0  # the test does exercise rounding


# ------------------------------------------------------------------------ MIS
def test_mis_independent_maximal_and_luby_fixed_point():
    for seed in range(5):
        m, rp, c = random_graph(200, 0.04, seed)
        prio = O.mis_priorities(m, seed=5489 + seed)
        g = O.mis_greedy(m, rp, c, prio)
        rows = np.repeat(np.arange(m), np.diff(rp))
        assert not np.any((g[rows] == 1) & (g[c] == 1))            # independent
        covered = np.zeros(m, bool)
        covered[c[g[rows] == 1]] = True
        assert np.all((g == 1) | covered)                           # maximal
        for delta in (0.1, 0.5, 2.0):
            lub, rounds = O.mis_luby(m, rp, c, prio, delta)
            assert lub.tolist() == g.tolist() and rounds > 0


def test_mis_ties_break_to_lower_id():
    # all priorities equal: greedy by id = the reference CPU MIS order
    m, rp, c = random_graph(80, 0.05, 3)
    prio = np.full(m, 7, np.uint32)
    g = O.mis_greedy(m, rp, c, prio)
    st = [-1] * m
    for v in range(m):
        if st[v] == -1:
            st[v] = 1
            for e in range(rp[v], rp[v + 1]):
                st[c[e]] = 0
    assert g.tolist() == st
    assert O.mis_luby(m, rp, c, prio)[0].tolist() == st


# ------------------------------------------------------------------------- TC
def test_tc_fixture_kats():
    kat = {"test_bc": 7, "test_cc": 9, "test_mesh": 0, "test_mis": 0, "test_pr": 4,
           "test_spgemm": 1, "chesapeake": 194}
    for name, want in kat.items():
        m, rp, c, _ = O.mtx_graph((MTX / f"{name}.mtx").read_text())
        total, rows = O.tc(m, rp, c)
        assert total == want == GOLD["fixtures"][name]["nx_triangles"]
        assert rows.tolist() == GOLD["fixtures"][name]["tc_rows"]


def test_tc_closed_forms_and_bruteforce():
    n = 12  # K_n: C(n,3) triangles, row r has C(r,2)
    u, v = zip(*combinations(range(n), 2))
    m, rp, c = (n,) + O.edges_to_csr(n, np.array(u, np.int32), np.array(v, np.int32))
    total, rows = O.tc(m, rp, c)
    assert total == n * (n - 1) * (n - 2) // 6
    assert rows.tolist() == [r * (r - 1) // 2 for r in range(n)]
    for seed in range(4):
        m, rp, c = random_graph(40, 0.2, seed)
        A = to_scipy(m, rp, c).toarray().astype(np.int64)
        total, rows = O.tc(m, rp, c)
        assert total == np.trace(A @ A @ A) // 6
        want = [sum(1 for i, j in combinations(range(r), 2) if A[r, i] and A[r, j] and A[i, j])
                for r in range(m)]
        assert rows.tolist() == want


# ------------------------------------------------------------------ primitives
@pytest.mark.parametrize("sr", [O.SR_PLUS_TIMES, O.SR_BOOL, O.SR_MIN_PLUS, O.SR_MAX_FIRST])
def test_spmspv_vs_dense(sr):
    rng = np.random.default_rng(sr)
    m, rp, c = random_graph(90, 0.06, sr)
    a = rng.integers(1, 9, len(c)).astype(np.float32)     # integer-valued: exact
    x_ind = np.sort(rng.choice(m, 20, replace=False)).astype(np.int32)
    x_val = rng.integers(1, 50, 20).astype(np.float32)
    mask = (rng.random(m) < 0.3).astype(np.int32)
    yi, yv = O.spmspv(m, rp, c, a, sr, x_ind, x_val, mask)
    A = to_scipy(m, rp, c, a).toarray()
    P = to_scipy(m, rp, c).toarray() > 0
    x = np.zeros(m); x[x_ind] = x_val
    xin = np.zeros(m, bool); xin[x_ind] = True
    hit = (P[xin].sum(0) > 0)
    if sr == O.SR_PLUS_TIMES:
        y = A.T @ x
    elif sr == O.SR_BOOL:
        y = hit.astype(float)
    elif sr == O.SR_MIN_PLUS:
        y = np.where(P[xin], x[xin][:, None] + A[xin], np.inf).min(0)
    else:
        y = np.where(P[xin], x[xin][:, None], -np.inf).max(0)
    keep = hit & (mask == 0)
    assert yi.tolist() == np.nonzero(keep)[0].tolist()
    assert yv.tolist() == y[keep].astype(np.float32).tolist()


def test_spgemm_vs_scipy():
    rng = np.random.default_rng(1)
    for seed in range(3):
        A = sp.random(60, 45, 0.08, random_state=seed, format="csr")
        B = sp.random(45, 70, 0.1, random_state=seed + 9, format="csr")
        A.data = rng.integers(1, 7, A.nnz).astype(np.float64)
        B.data = rng.integers(1, 7, B.nnz).astype(np.float64)
        A.sort_indices(); B.sort_indices()
        rp, cc, vv = O.spgemm(60, 45, 70, A.indptr.astype(np.int64), A.indices.astype(np.int32),
                              A.data.astype(np.float32), B.indptr.astype(np.int64),
                              B.indices.astype(np.int32), B.data.astype(np.float32))
        Cm = (A @ B).tocsr(); Cm.sort_indices()
        Cm.eliminate_zeros()
        assert rp.tolist() == Cm.indptr.tolist()
        assert cc.tolist() == Cm.indices.tolist()
        assert vv.tolist() == Cm.data.astype(np.float32).tolist()


def _csc(M):
    M = sp.csc_matrix(M)
    M.sort_indices()
    return (M.indptr.astype(np.int64), M.indices.astype(np.int32),
            M.data.astype(np.float32))


def test_mxm_vs_scipy_with_cancellation():
    """Column-by-column C = A*B: the pattern is that of |A|*|B| (a sum that
    cancels keeps its entry, value 0, as the reduce by key does), the values
    are scipy's; A is not square and not symmetric, so a transposed or
    swapped operand fails."""
    rng = np.random.default_rng(4)
    for seed in range(3):
        A = sp.random(50, 40, 0.1, random_state=seed, format="csc")
        B = sp.random(40, 65, 0.12, random_state=seed + 5, format="csc")
        A.data = rng.integers(-3, 4, A.nnz).astype(np.float64)  # zeros and signs: cancels
        B.data = rng.integers(-3, 4, B.nnz).astype(np.float64)
        cp, ri, cv = O.mxm(50, 40, 65, *_csc(A), *_csc(B))
        ones = lambda M: sp.csc_matrix((np.ones(M.nnz), M.indices, M.indptr), shape=M.shape)
        pat = (ones(A) @ ones(B)).tocsc()  # structural: stored zeros count
        pat.sort_indices()
        dense = (A @ B).toarray()
        assert cp.tolist() == pat.indptr.tolist()
        assert ri.tolist() == pat.indices.tolist()
        want = np.array([dense[r, j] for j in range(65)
                         for r in pat.indices[pat.indptr[j]:pat.indptr[j + 1]]], np.float32)
        assert cv.tolist() == want.tolist()
        assert (cv == 0).any()  # the cancellation case occurred


def test_mxm_hand_cases():
    # [1 1] * [1; -1]: one product pair that cancels -> one entry, value 0
    cp, ri, cv = O.mxm(1, 2, 1, np.array([0, 1, 2]), np.array([0, 0]), np.array([1.0, 1.0]),
                       np.array([0, 2]), np.array([0, 1]), np.array([1.0, -1.0]))
    assert cp.tolist() == [0, 1] and ri.tolist() == [0] and cv.tolist() == [0.0]
    # identity times B is B; an empty column of B gives an empty column of C
    B = sp.csc_matrix(np.array([[0, 2, 0], [5, 0, 0], [0, 7, 0]], float))
    cp, ri, cv = O.mxm(3, 3, 3, *_csc(sp.identity(3)), *_csc(B))
    assert cp.tolist() == B.indptr.tolist() and ri.tolist() == B.indices.tolist()
    assert cv.tolist() == B.data.tolist()


def test_mxm_bridges_spmspv_and_spgemm():
    """Column j of the column-by-column product is the SpMSpV of A's CSC with
    column j of B (mXm's loop body), and the whole product read as CSR is the
    row-wise SpGEMM of the transposes, B^T * A^T."""
    m, rp, c = random_graph(80, 0.07, 6)
    rng = np.random.default_rng(6)
    a = rng.integers(1, 9, len(c)).astype(np.float32)
    B = sp.random(80, 30, 0.1, random_state=2, format="csc")
    B.data = rng.integers(1, 9, B.nnz).astype(np.float64)
    bp, bi, bv = _csc(B)
    cp, ri, cv = O.mxm(80, 80, 30, rp, c, a, bp, bi, bv)
    for j in range(30):
        yi, yv = O.spmspv(m, rp, c, a, O.SR_PLUS_TIMES, bi[bp[j]:bp[j + 1]], bv[bp[j]:bp[j + 1]])
        assert ri[cp[j]:cp[j + 1]].tolist() == yi.tolist()
        assert cv[cp[j]:cp[j + 1]].tolist() == yv.tolist()
    srp, sc, sv = O.spgemm(30, 80, 80, bp, bi, bv, rp, c, a)
    assert srp.tolist() == cp.tolist() and sc.tolist() == ri.tolist()
    assert sv.tolist() == cv.tolist()


def dcsc_by_blocks(A, ps, pn):
    """Row-block DCSC from scipy: each block of rows converted to CSC, its
    empty columns dropped (an independent route to the same arrays)."""
    m = A.shape[0]
    pp, ci, co, ri, vv = [0], [], [0], [], []
    for p in range(pn):
        lo, hi = p * ps, (m if p == pn - 1 else (p + 1) * ps)
        B = sp.csc_matrix(A[lo:hi])
        B.sort_indices()
        for j in np.nonzero(np.diff(B.indptr))[0]:
            s, e = B.indptr[j], B.indptr[j + 1]
            ci.append(j)
            ri.extend((B.indices[s:e] + lo).tolist())
            vv.extend(B.data[s:e].tolist())
            co.append(co[-1] + e - s)
        pp.append(len(ci))
    return pp, ci, co, ri, vv


@pytest.mark.parametrize("ps,pn", [(16, 8), (10, 13), (128, 1), (1, 128), (40, 4)])
def test_dcsc_vs_scipy_blocks(ps, pn):
    """Parts of equal rows, a last part taking the rest (10 x 13 over 128 rows
    leaves 8 rows to the last part), one part (the transpose, compressed) and
    one row per part; a random graph with empty rows and distinct values."""
    m, rp, c = random_graph(128, 0.04, 8)
    v = np.arange(1, len(c) + 1, dtype=np.float32)
    A = sp.csr_matrix((v.astype(np.float64), c, rp), shape=(m, m))
    got = O.dcsc(m, rp, c, v, ps, pn)
    want = dcsc_by_blocks(A, ps, pn)
    for g, w in zip(got, want):
        assert g.tolist() == list(w)


def test_dcsc_hand_case():
    # rows 0-1 | 2-3 of [[0 5 0 6], [7 0 0 8], [0 0 0 0], [0 9 1 0]]
    rp = np.array([0, 2, 4, 4, 6])
    c = np.array([1, 3, 0, 3, 1, 2], np.int32)
    v = np.array([5, 6, 7, 8, 9, 1], np.float32)
    pp, ci, co, ri, vv = O.dcsc(4, rp, c, v, 2, 2)
    assert pp.tolist() == [0, 3, 5]
    assert ci.tolist() == [0, 1, 3, 1, 2]
    assert co.tolist() == [0, 1, 2, 4, 5, 6]
    assert ri.tolist() == [1, 0, 0, 1, 3, 3]
    assert vv.tolist() == [7, 5, 6, 8, 9, 1]


# ----------------------------------------------- planted mistakes are caught
def test_planted_bfs_mistake_is_caught():
    """An off-by-one in depth (a plausible bug) must fail the scipy pin."""
    m, rp, c = random_graph(60, 0.05, 2)
    depth, _ = O.bfs(m, rp, c, 0)
    bad = depth.copy()
    bad[bad > 1] += 1
    d = csg.shortest_path(to_scipy(m, rp, c), unweighted=True, indices=0)
    ref = np.where(np.isinf(d), -1, d).astype(int)
    assert depth.tolist() == ref.tolist()
    assert bad.tolist() != ref.tolist()


def test_planted_tc_mistake_is_caught():
    """Counting L_r cap L_k without the k bound over-counts: the closed form fails."""
    n = 8
    u, v = zip(*combinations(range(n), 2))
    m, rp, c = (n,) + O.edges_to_csr(n, np.array(u, np.int32), np.array(v, np.int32))
    _, rows = O.tc(m, rp, c)
    L_rp, L_c = O.lower(m, rp, c)
    wrong = [sum(len(set(L_c[L_rp[r]:L_rp[r + 1]]) | set(L_c[L_rp[k]:L_rp[k + 1]]))
                 for k in L_c[L_rp[r]:L_rp[r + 1]]) for r in range(m)]
    assert rows.tolist() == [r * (r - 1) // 2 for r in range(n)]
    assert wrong != rows.tolist()


def test_mxv_dense_vs_scipy():
    """Dense mXv: y = A^T (x restricted to x > 0.5), every y entry written
    (zeros included), total = the frontier rows' lengths.  A is rectangular
    (a transposed operand fails), x holds 0.5 exactly (>= instead of > fails)
    and 0.75 (exact sums: the products are multiples of 0.25),
    negative and small positive entries (not in the frontier) and a frontier
    row with no entries."""
    rng = np.random.default_rng(11)
    for seed in range(3):
        A = sp.random(60, 45, 0.08, random_state=seed, format="csr")
        A.data = rng.integers(1, 9, A.nnz).astype(np.float64)
        A.sort_indices()
        rp = A.indptr.astype(np.int64)
        x = rng.choice([0.0, 0.3, 0.5, 0.75, 1.0, 2.0, -3.0], 60).astype(np.float32)
        x[np.diff(rp) == 0] = 2.0  # rows without entries in the frontier
        y, total = O.mxv_dense(60, 45, rp, A.indices.astype(np.int32),
                               A.data.astype(np.float32), x, np.full(45, 7, np.float32))
        xf = np.where(x > 0.5, x, 0).astype(np.float64)
        assert y.tolist() == (A.T @ xf).astype(np.float32).tolist()
        assert total == int(np.diff(rp)[x > 0.5].sum())


def test_mxv_dense_hand_cases():
    # A = [[1 2] [0 3]], x = [1, 2]: y = [1, 2 + 6] = [1, 8]; total = 3 entries
    rp, col, val = np.array([0, 2, 3]), np.array([0, 1], np.int32), None
    col = np.array([0, 1, 1], np.int32)
    val = np.array([1, 2, 3], np.float32)
    y, t = O.mxv_dense(2, 2, rp, col, val, np.array([1, 2], np.float32))
    assert y.tolist() == [1.0, 8.0] and t == 3
    # x = [0.5, 0.2]: no frontier -> y untouched, 0
    y, t = O.mxv_dense(2, 2, rp, col, val, np.array([0.5, 0.2], np.float32),
                       np.array([4, 5], np.float32))
    assert y.tolist() == [4.0, 5.0] and t == 0
    # frontier row 1 only: y[0] is zeroed even though nothing lands there
    y, t = O.mxv_dense(2, 2, rp, col, val, np.array([0, 1], np.float32),
                       np.array([4, 5], np.float32))
    assert y.tolist() == [0.0, 3.0] and t == 1


@pytest.mark.parametrize("n,p,seed", [(2000, 0.0005, 1), (3000, 0.0008, 13), (500, 0.01, 4),
                                      (1, 0.0, 0)])
def test_components_vs_scipy(n, p, seed):
    """Union-find labels (smallest vertex id of each component) against
    scipy's connected components relabelled the same way, on graphs with many
    small components and isolated vertices; labels by the largest id of each
    component (what hooking the smaller root under the larger would give)
    must not match."""
    from tests import graphs as T
    g = T.random_graph(n, p, seed)
    a = sp.csr_matrix((np.ones(len(g.col)), g.col, g.rowptr), shape=(g.m, g.m))
    _, lab = csg.connected_components(a, directed=False)
    want = np.empty(g.m, np.int64)
    for c in np.unique(lab):
        members = np.nonzero(lab == c)[0]
        want[members] = members.min()
    got = O.components(g.m, g.rowptr, g.col)
    assert np.array_equal(got, want)
    if g.m > 1 and len(np.unique(lab)) < g.m:
        wrong = want.copy()
        for c in np.unique(lab):
            members = np.nonzero(lab == c)[0]
            wrong[members] = members.max()
        assert not np.array_equal(got, wrong)
