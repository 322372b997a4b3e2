"""The SpMSpV pipeline primitives and the general hash SpGEMM on the GPU:
scan, stable radix sort, SpMSpV under the four semirings on both execution
paths, SpGEMM, and device-side graph ingestion / transpose, each against the
oracle or a closed form."""
import numpy as np
import pytest
import scipy.sparse as sp

import apspie_amd as A
from oracle import oracle as O
from tests import graphs as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with A.Context(0) as c:
        yield c


@pytest.mark.parametrize("n", [0, 1, 5, 2047, 2048, 2049, 100_000, 3_000_001])
def test_scan_exclusive(ctx, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 100, n).astype(np.int32)
    out, tot = ctx.scan_exclusive(x)
    want = np.concatenate([[0], np.cumsum(x, dtype=np.int64)[:-1]]) if n else np.zeros(0)
    assert np.array_equal(out, want.astype(np.int32)) and tot == int(x.sum())


@pytest.mark.parametrize("n,bits", [(1, 8), (100, 8), (5000, 13), (1_000_003, 22), (300_000, 32)])
def test_radix_sort_stable(ctx, n, bits):
    rng = np.random.default_rng(bits)
    hi = (1 << bits) if bits < 32 else (1 << 32)
    keys = rng.integers(0, min(hi, 5000 if n > 1000 else hi), n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = ctx.sort_pairs(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
    k2, _ = ctx.sort_pairs(keys, None, bits)
    assert np.array_equal(k2, keys[order])


SEMIRINGS = [A.SR_PLUS_TIMES, A.SR_BOOL, A.SR_MIN_PLUS, A.SR_MAX_FIRST]


@pytest.mark.parametrize("path", [A.PATH_SORT, A.PATH_ATOMIC, A.PATH_AUTO])
@pytest.mark.parametrize("sr", SEMIRINGS)
def test_spmspv_vs_oracle(ctx, sr, path):
    rng = np.random.default_rng(sr * 7 + path)
    for g in (T.rmat(12), T.fixture("chesapeake"), T.lattice(32)):
        a = rng.integers(1, 9, len(g.col)).astype(np.float32)
        G = ctx.upload(g.rowptr32, g.col, a)
        for nx in (1, 17, g.m // 3):
            x_ind = np.sort(rng.choice(g.m, nx, replace=False)).astype(np.int32)
            x_val = rng.integers(-20, 50, nx).astype(np.float32)
            mask = (rng.random(g.m) < 0.25).astype(np.int32)
            for mk in (None, mask):
                yi, yv = ctx.spmspv(G, sr, x_ind, x_val, mk, path)
                wi, wv = O.spmspv(g.m, g.rowptr, g.col, a, sr, x_ind, x_val, mk)
                assert np.array_equal(yi, wi) and np.array_equal(yv, wv), (g, sr, path, nx)
        G.close()


def test_spmspv_sort_path_is_sequential_fold(ctx):
    """Non-integer (+,x): the sort path folds each run in expansion order,
    so it is bit-identical to the oracle's sequential fold."""
    g = T.rmat(11)
    rng = np.random.default_rng(0)
    a = rng.random(len(g.col)).astype(np.float32)
    G = ctx.upload(g.rowptr32, g.col, a)
    x_ind = np.sort(rng.choice(g.m, 300, replace=False)).astype(np.int32)
    x_val = rng.random(300).astype(np.float32)
    yi, yv = ctx.spmspv(G, A.SR_PLUS_TIMES, x_ind, x_val, None, A.PATH_SORT)
    wi, wv = O.spmspv(g.m, g.rowptr, g.col, a, A.SR_PLUS_TIMES, x_ind, x_val)
    assert np.array_equal(yi, wi) and np.array_equal(yv, wv)
    yi, yv = ctx.spmspv(G, A.SR_PLUS_TIMES, x_ind, x_val, None, A.PATH_ATOMIC)
    assert np.array_equal(yi, wi) and np.allclose(yv, wv, rtol=1e-5, atol=1e-6)
    G.close()


def test_spmspv_empty_and_zero_degree(ctx):
    g = T.random_graph(500, 0.002, 9)
    G = ctx.upload(g.rowptr32, g.col, np.ones(len(g.col), np.float32))
    zero = np.nonzero(g.deg == 0)[0][:5].astype(np.int32)
    yi, yv = ctx.spmspv(G, A.SR_BOOL, zero)
    assert len(yi) == 0
    yi, yv = ctx.spmspv(G, A.SR_BOOL, np.zeros(0, np.int32))
    assert len(yi) == 0
    G.close()


def test_mxv_dense_vs_oracle(ctx):
    """Dense mXv through the C ABI: frontier = x above 0.5 (0.5 itself
    excluded), y zeroed then summed, the frontier edge count returned; a
    rectangular matrix, R-MAT and a lattice, at several frontier sizes; an
    empty frontier leaves y untouched."""
    import scipy.sparse as sp
    rng = np.random.default_rng(5)
    R = sp.random(700, 450, 0.02, random_state=3, format="csr")
    R.sort_indices()
    cases = [(700, 450, R.indptr.astype(np.int64), R.indices.astype(np.int32))]
    for g in (T.rmat(12), T.lattice(32)):
        cases.append((g.m, g.m, g.rowptr, g.col))
    for m, n, rp, col in cases:
        a = rng.integers(1, 9, len(col)).astype(np.float32)
        G = ctx.upload(rp.astype(np.int32), col, a, n=n)
        for frac in (0.001, 0.05, 0.6):
            x = np.where(rng.random(m) < frac, rng.integers(1, 20, m), 0).astype(np.float32)
            x[rng.random(m) < 0.05] = 0.5
            y0 = np.full(n, 3, np.float32)
            got, total = ctx.mxv_dense(G, x, y0.copy())
            want, wt = O.mxv_dense(m, n, rp, col, a, x, y0)
            assert np.array_equal(got, want) and total == wt, (m, n, frac)
        got, total = ctx.mxv_dense(G, np.full(m, 0.5, np.float32), np.full(n, 3, np.float32))
        assert total == 0 and (got == 3).all()
        G.close()


def _scipy_csr(g_or_mat):
    return g_or_mat


@pytest.mark.parametrize("case", ["rect", "square_int", "rmat_sq", "hub", "rmat12_sq",
                                  "wide_global"])
def test_spgemm_vs_oracle(ctx, case):
    """All accumulator tiers: wave and block LDS tables, the dense accumulator
    (a row of 7000 columns in 'hub', R-MAT-12 squared) and the global table
    (a row of 50000 possible columns in 'wide_global')."""
    rng = np.random.default_rng(3)
    if case == "rmat12_sq":
        g = T.rmat(12)
        Am = sp.csr_matrix((np.ones(len(g.col)), g.col, g.rowptr), shape=(g.m, g.m))
        Bm = Am
    elif case == "wide_global":
        Am = sp.lil_matrix((20, 300))
        Am[0, :] = 1
        Am[1, :10] = 1
        Am = Am.tocsr()
        Bm = sp.random(300, 50000, 0.01, random_state=6, format="csr")
    elif case == "rect":
        Am = sp.random(300, 200, 0.05, random_state=1, format="csr")
        Bm = sp.random(200, 400, 0.05, random_state=2, format="csr")
    elif case == "square_int":
        Am = sp.random(1000, 1000, 0.01, random_state=3, format="csr")
        Bm = Am.T.tocsr()
    elif case == "rmat_sq":
        g = T.rmat(10)
        Am = sp.csr_matrix((np.ones(len(g.col)), g.col, g.rowptr), shape=(g.m, g.m))
        Bm = Am
    else:  # one row with a huge product count (global spill tier)
        Am = sp.lil_matrix((50, 6000))
        Am[0, :] = 1
        Am[1, :40] = 1
        Am = Am.tocsr()
        Bm = sp.random(6000, 7000, 0.002, random_state=5, format="csr")
    for M in (Am, Bm):
        M.sort_indices()
        M.data = rng.integers(1, 5, M.nnz).astype(np.float64)
    Ga = ctx.upload(Am.indptr.astype(np.int32), Am.indices.astype(np.int32),
                    Am.data.astype(np.float32), n=Am.shape[1])
    Gb = ctx.upload(Bm.indptr.astype(np.int32), Bm.indices.astype(np.int32),
                    Bm.data.astype(np.float32), n=Bm.shape[1])
    Gc = ctx.spgemm(Ga, Gb)
    rp, cc, vv = Gc.download(values=True)
    wrp, wc, wv = O.spgemm(Am.shape[0], Am.shape[1], Bm.shape[1], Am.indptr.astype(np.int64),
                           Am.indices.astype(np.int32), Am.data.astype(np.float32),
                           Bm.indptr.astype(np.int64), Bm.indices.astype(np.int32),
                           Bm.data.astype(np.float32))
    assert np.array_equal(rp, wrp.astype(np.int32))
    assert np.array_equal(cc, wc) and np.array_equal(vv, wv)
    for G in (Ga, Gb, Gc):
        G.close()


def _csc_graph(ctx, M):
    M = sp.csc_matrix(M)
    M.sort_indices()
    return ctx.upload(M.indptr.astype(np.int32), M.indices.astype(np.int32),
                      M.data.astype(np.float32), n=M.shape[0])


@pytest.mark.parametrize("case", ["rect_cancel", "empty_cols", "rmat10_weighted", "wide_col"])
def test_mxm_vs_oracle(ctx, case):
    """Column-by-column C = A*B in CSC (mXm) against the oracle's column loop:
    non-square operands with signed values (sums that cancel keep their entry),
    empty columns on both sides, a weighted R-MAT squared, and a column of B
    whose products spill past the LDS tables."""
    rng = np.random.default_rng(11)
    if case == "rect_cancel":
        Am = sp.random(300, 200, 0.05, random_state=1, format="csc")
        Bm = sp.random(200, 400, 0.05, random_state=2, format="csc")
        vals = lambda k: rng.integers(-3, 4, k)
    elif case == "empty_cols":
        Am = sp.random(120, 90, 0.04, random_state=3, format="lil")
        Am[:, 10:30] = 0
        Bm = sp.random(90, 150, 0.05, random_state=4, format="lil")
        Bm[:, 50:80] = 0
        Am, Bm = sp.csc_matrix(Am), sp.csc_matrix(Bm)
        vals = lambda k: rng.integers(1, 6, k)
    elif case == "rmat10_weighted":
        g = T.rmat(10)
        w = O.assign_weights(g.m, g.rowptr, g.col)
        Am = sp.csr_matrix((w, g.col, g.rowptr), shape=(g.m, g.m)).tocsc()
        Bm = Am
        vals = None
    else:
        Am = sp.random(6000, 300, 0.02, random_state=5, format="csc")
        Bm = sp.lil_matrix((300, 40))
        Bm[:, 0] = 1
        Bm[::2, 1] = 1
        Bm = sp.csc_matrix(Bm)
        vals = lambda k: rng.integers(1, 6, k)
    if vals is not None:
        for M in (Am, Bm):
            M.data = vals(M.nnz).astype(np.float64)
    Ga, Gb = _csc_graph(ctx, Am), _csc_graph(ctx, Bm)
    Gc = ctx.mxm(Ga, Gb)
    assert (Gc.m, Gc.n) == (Bm.shape[1], Am.shape[0])
    cp, ri, cv = Gc.download(values=True)
    A_ = sp.csc_matrix(Am); A_.sort_indices()
    B_ = sp.csc_matrix(Bm); B_.sort_indices()
    wcp, wri, wcv = O.mxm(Am.shape[0], Am.shape[1], Bm.shape[1], A_.indptr, A_.indices,
                          A_.data.astype(np.float32), B_.indptr, B_.indices,
                          B_.data.astype(np.float32))
    assert np.array_equal(cp, wcp.astype(np.int32))
    assert np.array_equal(ri, wri) and np.array_equal(cv, wcv)
    if case == "rect_cancel":
        assert (cv == 0).any()
    for G in (Ga, Gb, Gc):
        G.close()


def test_mxm_columns_equal_spmspv(ctx):
    """The bridge between the two engines: each column of mXm equals the SpMSpV
    of A's CSC with that column of B, computed on the device by apspie_spmspv."""
    g = T.rmat(9)
    rng = np.random.default_rng(12)
    a = rng.integers(1, 9, len(g.col)).astype(np.float32)
    Ga = ctx.upload(g.rowptr.astype(np.int32), g.col, a)
    Bm = sp.random(g.m, 24, 0.05, random_state=7, format="csc")
    Bm.data = rng.integers(1, 9, Bm.nnz).astype(np.float64)
    Bm.sort_indices()
    Gb = _csc_graph(ctx, Bm)
    Gc = ctx.mxm(Ga, Gb)
    cp, ri, cv = Gc.download(values=True)
    for j in range(24):
        s, e = Bm.indptr[j], Bm.indptr[j + 1]
        yi, yv = ctx.spmspv(Ga, A.SR_PLUS_TIMES, Bm.indices[s:e], Bm.data[s:e])
        assert np.array_equal(ri[cp[j]:cp[j + 1]], yi)
        assert np.array_equal(cv[cp[j]:cp[j + 1]], yv)
    for G in (Ga, Gb, Gc):
        G.close()


def test_mxm_rejects_bad_shapes(ctx):
    Ga = _csc_graph(ctx, sp.random(5, 4, 0.5, random_state=1))   # A is 5 x 4
    Gb = _csc_graph(ctx, sp.random(3, 6, 0.5, random_state=2))   # B is 3 x 6
    with pytest.raises(A.ApspieError):
        ctx.mxm(Ga, Gb)
    for G in (Ga, Gb):
        G.close()


@pytest.mark.parametrize("case,ps,pn", [("rmat10", 128, 8), ("rmat10", 100, 11),
                                       ("random_empty", 7, 40), ("lattice32", 1024, 1),
                                       ("lattice32", 1, 1024), ("no_values", 64, 4),
                                       ("no_entries", 3, 2)])
def test_dcsc_vs_oracle(ctx, case, ps, pn):
    """Row-block DCSC on the device against the oracle's step-by-step
    restatement of csr_to_dcsc: equal parts, a last part taking the rest
    (11 parts of 100 rows over 1024), rows with no entries, one part (the
    compressed transpose), one row per part, a graph without values and one
    without entries."""
    if case == "rmat10":
        g = T.rmat(10)
        m, rp, c = g.m, g.rowptr, g.col
        v = O.assign_weights(m, rp, c)
    elif case == "random_empty":
        g = T.random_graph(300, 0.01, 5)
        m, rp, c = g.m, g.rowptr, g.col
        v = np.arange(1, len(c) + 1, dtype=np.float32)
    elif case == "lattice32":
        g = T.lattice(32)
        m, rp, c = g.m, g.rowptr, g.col
        v = np.arange(len(c), dtype=np.float32)
    elif case == "no_values":
        g = T.rmat(8)
        m, rp, c, v = g.m, g.rowptr, g.col, None
    else:
        m, rp, c, v = 5, np.zeros(6, np.int64), np.zeros(0, np.int32), None
    G = ctx.upload(np.asarray(rp, np.int32), np.asarray(c, np.int32), v, n=m)
    pp, ci, co, ri, vv = G.to_dcsc(ps, pn, values=v is not None)
    wpp, wci, wco, wri, wvv = O.dcsc(m, rp, c, v, ps, pn)
    assert pp.tolist() == wpp.tolist()
    assert ci.tolist() == wci.tolist() and co.tolist() == wco.tolist()
    assert ri.tolist() == wri.tolist()
    if v is not None:
        assert np.array_equal(vv, wvv)
    G.close()


def test_dcsc_rejects_bad_parts(ctx):
    g = T.rmat(8)
    G = ctx.upload(g.rowptr.astype(np.int32), g.col)
    for ps, pn in ((0, 4), (4, 0), (64, 6)):  # 5 * 64 > 256 rows: a part past the end
        with pytest.raises(A.ApspieError):
            G.to_dcsc(ps, pn, values=False)
    G.close()


def _masked_want(Am, Bm, Mm):
    """The oracle's full product, read at the mask's entries."""
    wrp, wc, wv = O.spgemm(Am.shape[0], Am.shape[1], Bm.shape[1], Am.indptr.astype(np.int64),
                           Am.indices.astype(np.int32), Am.data.astype(np.float32),
                           Bm.indptr.astype(np.int64), Bm.indices.astype(np.int32),
                           Bm.data.astype(np.float32))
    vals, hit = [], 0
    for i in range(Mm.shape[0]):
        row = dict(zip(wc[wrp[i]:wrp[i + 1]].tolist(), wv[wrp[i]:wrp[i + 1]].tolist()))
        for j in Mm.indices[Mm.indptr[i]:Mm.indptr[i + 1]].tolist():
            vals.append(row.get(j, 0.0))
            hit += j in row
    return np.array(vals, np.float32), hit


@pytest.mark.parametrize("case", ["rect", "rmat_tc", "long_mask_row", "empty_rows"])
def test_spgemm_masked_vs_oracle(ctx, case):
    """(A*B) o M on M's pattern equals the oracle's full product read at M's
    entries; on L, L, L the sum is the triangle count."""
    rng = np.random.default_rng(8)
    if case == "rect":
        Am = sp.random(300, 200, 0.05, random_state=1, format="csr")
        Bm = sp.random(200, 400, 0.05, random_state=2, format="csr")
        Mm = sp.random(300, 400, 0.1, random_state=3, format="csr")
    elif case == "rmat_tc":
        g = T.rmat(10)
        L = sp.tril(sp.csr_matrix((np.ones(len(g.col)), g.col, g.rowptr), shape=(g.m, g.m)),
                    k=-1).tocsr()
        Am = Bm = Mm = L
    elif case == "long_mask_row":  # a mask row beyond the LDS table (global tier)
        Am = sp.random(20, 300, 0.3, random_state=4, format="csr")
        Bm = sp.random(300, 9000, 0.01, random_state=5, format="csr")
        Mm = sp.lil_matrix((20, 9000))
        Mm[0, :] = 1
        Mm[1, ::3] = 1
        Mm = Mm.tocsr()
    else:
        Am = sp.random(100, 100, 0.02, random_state=6, format="csr")
        Bm = sp.random(100, 100, 0.02, random_state=7, format="csr")
        Mm = sp.random(100, 100, 0.05, random_state=8, format="csr")
        Mm = sp.csr_matrix(Mm.multiply(np.arange(100)[:, None] % 3 == 0))
    for X in (Am, Bm, Mm):
        X.sort_indices()
    if case != "rmat_tc":
        for X in (Am, Bm):
            X.data = rng.integers(1, 5, X.nnz).astype(np.float64)
    Gs = [ctx.upload(X.indptr.astype(np.int32), X.indices.astype(np.int32),
                     X.data.astype(np.float32), n=X.shape[1]) for X in (Am, Bm, Mm)]
    got, nnz = ctx.spgemm_masked(*Gs)
    want, hit = _masked_want(Am, Bm, Mm)
    assert np.array_equal(got, want) and nnz == hit, case
    if case == "rmat_tc":
        assert int(got.sum()) == T.GOLD["rmat10"]["triangles"]
    for G in Gs:
        G.close()


def test_spgemm_masked_rejects_bad_shapes(ctx):
    A1 = ctx.upload(np.array([0, 1], np.int32), np.array([0], np.int32), n=2)
    B1 = ctx.upload(np.array([0, 1, 1], np.int32), np.array([1], np.int32), n=3)
    M_bad = ctx.upload(np.array([0, 1], np.int32), np.array([0], np.int32), n=2)
    with pytest.raises(A.ApspieError):
        ctx.spgemm_masked(A1, B1, M_bad)  # mask is 1x2, product is 1x3
    for G in (A1, B1, M_bad):
        G.close()


@pytest.mark.parametrize("name", T.FIXTURES)
def test_graph_from_coo_matches_host_build(ctx, name):
    path = T.GOLDEN_DIR / "mtx" / f"{name}.mtx"
    m, n, row, col, val = A.read_mtx(path)
    nv = max(m, n)
    G = ctx.from_coo(nv, row, col, val, symmetrize=True)
    rp, cc, _ = G.download()
    g = T.fixture(name)
    assert np.array_equal(rp, g.rowptr32) and np.array_equal(cc, g.col)
    G.close()


def test_graph_from_coo_values_and_buildmatrix_order(ctx):
    row = np.array([0, 1, 2, 1, 0, 2], np.int32)
    col = np.array([1, 0, 2, 2, 1, 0], np.int32)
    val = np.array([3, 7, 9, 4, 8, 5], np.float32)
    G = ctx.from_coo(3, row, col, val, symmetrize=True)
    rp, cc, vv = G.download(values=True)
    r2, c2, v2 = O.symmetrize(np.concatenate([row, np.zeros(6, np.int32)]),
                              np.concatenate([col, np.zeros(6, np.int32)]),
                              np.concatenate([val, np.zeros(6, np.float32)]), 6)
    wrp, wc, wv = O.coo_to_csr(3, r2, c2, v2)
    assert np.array_equal(rp, wrp.astype(np.int32)) and np.array_equal(cc, wc)
    assert np.array_equal(vv, wv)
    G.close()
    # no symmetrize: buildMatrix keeps input order inside a row, loops and repeats
    G = ctx.from_coo(3, row, col, val, symmetrize=False)
    rp, cc, vv = G.download(values=True)
    wrp, wc, wv = O.coo_to_csr(3, row, col, val)
    assert np.array_equal(rp, wrp.astype(np.int32)) and np.array_equal(cc, wc)
    assert np.array_equal(vv, wv)
    G.close()


def test_graph_from_coo_large(ctx):
    s, d = O.rmat_edges(14)
    G = ctx.from_coo(1 << 14, s, d, None, symmetrize=True)
    rp, cc, _ = G.download()
    g = T.rmat(14)
    assert np.array_equal(rp, g.rowptr32) and np.array_equal(cc, g.col)
    G.close()


def test_transpose(ctx):
    Mx = sp.random(500, 300, 0.02, random_state=4, format="csr")
    Mx.sort_indices()
    Mx.data = np.arange(1, Mx.nnz + 1, dtype=np.float64)
    G = ctx.upload(Mx.indptr.astype(np.int32), Mx.indices.astype(np.int32),
                   Mx.data.astype(np.float32), n=300)
    T_ = G.transpose()
    rp, cc, vv = T_.download(values=True)
    Wt = Mx.T.tocsr()
    Wt.sort_indices()
    assert np.array_equal(rp, Wt.indptr) and np.array_equal(cc, Wt.indices)
    assert np.array_equal(vv, Wt.data.astype(np.float32))
    G.close()
    T_.close()


def test_hbm_budget_is_enforced(ctx):
    g = T.rmat(12)
    with A.Context(0) as c2:
        c2.set_hbm_budget(1000)
        with pytest.raises(A.ApspieError) as e:
            c2.upload(g.rowptr32, g.col)
        assert e.value.status == -4
        c2.set_hbm_budget(0)
        G = c2.upload(g.rowptr32, g.col)
        assert c2.hbm_used() >= G.hbm_bytes > 0
        G.close()


@pytest.mark.parametrize("case", ["rmat", "random", "lattice", "fixtures"])
def test_components_vs_oracle(ctx, case):
    """Connected-component labels (smallest vertex id per component) equal
    the oracle's union-find on graphs with many small components, isolated
    vertices, long paths and the reference's fixtures (the unions race
    across the whole device: several seeds and sizes)."""
    if case == "fixtures":
        gs = [T.fixture(n) for n in T.FIXTURES]
    elif case == "random":
        gs = [T.random_graph(n, p, s) for n, p in ((5000, 0.0003), (20000, 0.00008))
              for s in range(1, 6)]
    elif case == "rmat":
        gs = [T.rmat(12), T.rmat(16)]
    else:
        gs = [T.lattice(64), T.lattice(256)]
    for g in gs:
        G = ctx.upload(g.rowptr32, g.col)
        got = ctx.components(G)
        assert np.array_equal(got, O.components(g.m, g.rowptr, g.col)), g
        assert np.array_equal(ctx.components(G), got)  # cached with the graph
        G.close()


def test_components_need_symmetric(ctx):
    rp = np.array([0, 1, 1], np.int32)
    G = ctx.upload(rp, np.array([1], np.int32))
    with pytest.raises(A.ApspieError):
        ctx.components(G)
    G.close()


@pytest.mark.parametrize("wave_ub,block_ub,dense", [(64, 1024, 1), (128, 2048, 0),
                                                     (256, 4096, 1), (1, 1, 0), (256, 4096, 0)])
def test_spgemm_tier_bounds_exact(ctx, wave_ub, block_ub, dense):
    """The tuner's SpGEMM knobs only move rows between tiers: every setting
    (down to every row in the dense or global tier) gives the oracle's
    product on R-MAT-12 squared and on a matrix with a very long row."""
    ctx.set_option("spgemm_wave_ub", wave_ub)
    ctx.set_option("spgemm_block_ub", block_ub)
    ctx.set_option("spgemm_dense", dense)
    try:
        g = T.rmat(12)
        Am = sp.csr_matrix((np.ones(len(g.col)), g.col, g.rowptr), shape=(g.m, g.m))
        Hm = sp.lil_matrix((50, 6000))
        Hm[0, :] = 1
        Hm[1, :40] = 1
        Hm = Hm.tocsr()
        Bh = sp.random(6000, 7000, 0.002, random_state=5, format="csr")
        Bh.data = np.random.default_rng(8).integers(1, 5, Bh.nnz).astype(np.float64)  # exact sums
        for A_, B_ in ((Am, Am), (Hm, Bh)):
            for M in (A_, B_):
                M.sort_indices()
            a = A_.data.astype(np.float32)
            b = B_.data.astype(np.float32) if B_ is not A_ else a
            Ga = ctx.upload(A_.indptr.astype(np.int32), A_.indices.astype(np.int32), a,
                            n=A_.shape[1])
            Gb = Ga if B_ is A_ else ctx.upload(B_.indptr.astype(np.int32),
                                                B_.indices.astype(np.int32), b, n=B_.shape[1])
            Gc = ctx.spgemm(Ga, Gb)
            rp, cc, vv = Gc.download(values=True)
            wrp, wc, wv = O.spgemm(A_.shape[0], A_.shape[1], B_.shape[1],
                                   A_.indptr.astype(np.int64), A_.indices.astype(np.int32), a,
                                   B_.indptr.astype(np.int64), B_.indices.astype(np.int32), b)
            assert np.array_equal(rp, wrp.astype(np.int32)) and np.array_equal(cc, wc)
            assert np.array_equal(vv, wv)
            for G in {id(Ga): Ga, id(Gb): Gb, id(Gc): Gc}.values():
                G.close()
    finally:
        ctx.set_option("spgemm_wave_ub", 256)
        ctx.set_option("spgemm_block_ub", 4096)
        ctx.set_option("spgemm_dense", 1)


@pytest.mark.parametrize("medium_row", [128, 256, 1024, 2048])
def test_tc_tier_bound_exact(ctx, medium_row):
    """TC's tuned medium / big tier bound moves rows between the 256- and
    1024-thread tiers only: the count and the per-row counts stay the
    oracle's (R-MAT-14 has rows in every tier)."""
    ctx.set_option("tc_medium_row", medium_row)
    try:
        g = T.rmat(14)
        G = ctx.upload(g.rowptr32, g.col)
        total, rows = ctx.tc(G)
        wt, wr = O.tc(g.m, g.rowptr, g.col)
        assert total == wt and np.array_equal(rows, wr)
        G.close()
    finally:
        ctx.set_option("tc_medium_row", 512)


def test_autotune_hash_and_table_round_trip(tmp_path):
    """autotune_hash picks tier bounds among its candidates and the results
    stay exact; the tuned options saved for (this GPU, the graph's class) come
    back on a fresh context for a graph of the same class, not for another
    class; saving again replaces the entry; a malformed entry is refused."""
    g = T.rmat(12)
    path = tmp_path / "tune.txt"
    with A.Context(0) as c:
        G = c.upload(g.rowptr32, g.col)
        c.autotune_hash(G)
        picked = {k: c.get_option(k) for k in ("tc_medium_row", "spgemm_wave_ub",
                                                 "spgemm_block_ub", "spgemm_dense")}
        assert picked["tc_medium_row"] in (256, 512, 1024, 2048)
        assert picked["spgemm_wave_ub"] in (64, 128, 256)
        assert picked["spgemm_block_ub"] in (1024, 2048, 4096)
        assert picked["spgemm_dense"] in (0, 1)
        total, rows = c.tc(G)
        wt, wr = O.tc(g.m, g.rowptr, g.col)
        assert total == wt and np.array_equal(rows, wr)
        c.set_option("bfs_alpha", 28)
        c.tune_save(G, path)
        c.tune_save(G, path)  # same (GPU, class): still one entry
        lines = path.read_text().splitlines()
        assert len(lines) == 1 and "class=v12_" in lines[0] and "bfs_alpha=28" in lines[0]
        G.close()
    with A.Context(0) as c:
        G = c.upload(g.rowptr32, g.col)  # same graph, fresh context
        assert c.get_option("bfs_alpha") == 14
        assert c.tune_load(G, path)
        assert c.get_option("bfs_alpha") == 28
        for k, v in picked.items():
            assert c.get_option(k) == v
        other = T.lattice(32)
        H = c.upload(other.rowptr32, other.col)
        assert not c.tune_load(H, path)  # another class: nothing applied
        assert not c.tune_load(G, tmp_path / "missing.txt")
        H.close()
        bad = tmp_path / "bad.txt"
        bad.write_text(lines[0].replace("bfs_alpha=28", "bfs_alpha=x"))
        with pytest.raises(A.ApspieError):
            c.tune_load(G, bad)
        bad.write_text(lines[0].replace("bfs_alpha=28", "bfs_tile=3000"))
        with pytest.raises(A.ApspieError):
            c.tune_load(G, bad)
        G.close()
