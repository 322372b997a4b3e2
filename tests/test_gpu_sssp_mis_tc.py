"""SSSP, MIS and triangle counting on the GPU through the C ABI, bit-exact
against the oracle (testSssp.cpp Dijkstra restated, greedy MIS closed form,
ewiseMultTc semantics) and the golden fixtures."""
import numpy as np
import pytest

import apspie_amd as A
from oracle import oracle as O
from tests import graphs as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with A.Context(0) as c:
        yield c


def to_ref_sentinel(d):
    """oracle marks unreached with -1 (testSssp.cpp:49); the GPU path with
    the reference sentinel 1.70141e38 (sssp.cuh:48)."""
    return np.where(d == -1, A.SSSP_INF, d).astype(np.float32)


# ---------------------------------------------------------------------- SSSP
@pytest.mark.parametrize("name", T.FIXTURES)
def test_sssp_fixtures(ctx, name):
    g = T.fixture(name)
    G = ctx.upload(g.rowptr32, g.col, g.w)
    got = ctx.sssp(G, 0)
    want = to_ref_sentinel(np.array(T.GOLD["fixtures"][name]["sssp"], np.float32))
    assert np.array_equal(got, want)
    G.close()


@pytest.mark.parametrize("delta", [-1, 0])
@pytest.mark.parametrize("key", ["rmat10", "rmat14", "lattice32", "lattice64"])
def test_sssp_synthetic_vs_oracle(ctx, key, delta):
    """The default (-1: near-far with the width chosen from the weights) and
    the Bellman-Ford rounds (0)."""
    g = T.rmat(int(key[4:])) if key.startswith("rmat") else T.lattice(int(key[7:]))
    ctx.set_option("sssp_delta", delta)
    try:
        G = ctx.upload(g.rowptr32, g.col, g.w)
        for src in [T.GOLD[key]["source"]] + T.sources(g, 4):
            want = to_ref_sentinel(O.sssp(g.m, g.rowptr, g.col, g.w, src))
            got = ctx.sssp(G, src)
            assert np.array_equal(got, want), (key, src)
        G.close()
    finally:
        ctx.set_option("sssp_delta", -1)


@pytest.mark.parametrize("delta", [1, 7, 64, 1000, 1 << 20])
@pytest.mark.parametrize("key", ["rmat10", "rmat14", "lattice64", "test_cc", "directed"])
def test_sssp_near_far_vs_oracle(ctx, key, delta):
    """Near-far rounds (sssp_delta > 0) reach the same fixed point as the
    Bellman-Ford frontier, for thresholds from below the smallest weight to
    above every distance; also on a directed graph and the fixtures."""
    if key == "directed":
        rng = np.random.default_rng(31)
        n = 3000
        u = rng.integers(0, n, 20000)
        v = rng.integers(0, n, 20000)
        k = np.unique((u * n + v)[u != v])
        rp, c, _ = O.coo_to_csr(n, (k // n).astype(np.int32), (k % n).astype(np.int32))
        w = (1 + rng.integers(0, 255, len(c))).astype(np.float32)
        m, rowptr, col, wt = n, rp, c, w
    else:
        g = T.fixture(key) if key == "test_cc" else (
            T.rmat(int(key[4:])) if key.startswith("rmat") else T.lattice(int(key[7:])))
        m, rowptr, col, wt = g.m, g.rowptr, g.col, g.w
    ctx.set_option("sssp_delta", delta)
    try:
        G = ctx.upload(rowptr.astype(np.int32), col, wt)
        deg = np.diff(rowptr)
        srcs = [int(s) for s in np.nonzero(deg > 0)[0][[0, -1]]] + [int(np.argmax(deg))]
        for src in srcs:
            want = to_ref_sentinel(O.sssp(m, rowptr, col, wt, src))
            got = ctx.sssp(G, src)
            assert np.array_equal(got, want), (key, delta, src)
        G.close()
    finally:
        ctx.set_option("sssp_delta", -1)


def wheel(n):
    """hub n-1 joined to every rim vertex, rim a path: one row spans many tiles."""
    rim = np.arange(n - 1)
    u = np.concatenate([rim[:-1], rim])
    v = np.concatenate([rim[1:], np.full(n - 1, n - 1)])
    rp, c = O.edges_to_csr(n, u.astype(np.int32), v.astype(np.int32))
    return rp, c


@pytest.mark.parametrize("pull_div", [0, 1, 2, 1 << 20])
@pytest.mark.parametrize("key", ["rmat10", "rmat14", "lattice64", "star", "wheel", "test_cc"])
def test_sssp_pull_rounds_vs_oracle(ctx, key, pull_div):
    """Rounds pulled over the frontier bitmap (symmetric structure and
    weights) reach the oracle's distances: push only (0), pull on big rounds
    (1, 2) and pull on every round (2^20); the star and the wheel put hubs of
    3000 and 20000 entries into the tiled long-row pass, and a pulled round
    reads every CSR entry."""
    if key == "wheel":
        rp, c = wheel(20000)
        m, rowptr, col = 20000, rp, c
        wt = O.assign_weights(m, rp, c)
    else:
        g = {"rmat10": lambda: T.rmat(10), "rmat14": lambda: T.rmat(14),
             "lattice64": lambda: T.lattice(64), "star": lambda: T.star_graph(3000),
             "test_cc": lambda: T.fixture("test_cc")}[key]()
        m, rowptr, col, wt = g.m, g.rowptr, g.col, g.w
        if key == "star":
            wt = O.assign_weights(m, rowptr, col)
    ctx.set_option("sssp_pull_div", pull_div)
    ctx.set_option("sssp_delta", 0)  # pulls apply to the Bellman-Ford rounds
    try:
        G = ctx.upload(rowptr.astype(np.int32), col, wt)
        deg = np.diff(rowptr)
        srcs = [int(s) for s in np.nonzero(deg > 0)[0][[0, -1]]] + [int(np.argmax(deg))]
        for src in srcs:
            want = to_ref_sentinel(O.sssp(m, rowptr, col, wt, src))
            got = ctx.sssp(G, src)
            assert np.array_equal(got, want), (key, pull_div, src)
            st = ctx.stats()
            if pull_div == 0:  # push only (the default)
                assert st["edges_checked"] == st["edges_traversed"]
            if pull_div == 1 << 20:
                assert st["edges_checked"] == st["iterations"] * len(col)
        G.close()
    finally:
        ctx.set_option("sssp_pull_div", 0)
        ctx.set_option("sssp_delta", -1)


def test_sssp_pull_needs_symmetric_weights(ctx):
    """A symmetric pattern with w(u, v) != w(v, u), and a directed graph, are
    never pulled (a pull reads w(v, u) for the edge u -> v)."""
    g = T.rmat(12)
    rng = np.random.default_rng(5)
    w = (1 + rng.integers(0, 63, len(g.col))).astype(np.float32)
    rng = np.random.default_rng(17)
    n = 3000
    u = rng.integers(0, n, 30000)
    v = rng.integers(0, n, 30000)
    k = np.unique((u * n + v)[u != v])
    drp, dc, _ = O.coo_to_csr(n, (k // n).astype(np.int32), (k % n).astype(np.int32))
    dw = (1 + rng.integers(0, 255, len(dc))).astype(np.float32)
    ctx.set_option("sssp_pull_div", 1 << 20)
    ctx.set_option("sssp_delta", 0)
    try:
        for m, rowptr, col, wt in [(g.m, g.rowptr, g.col, w), (n, drp, dc, dw)]:
            G = ctx.upload(rowptr.astype(np.int32), col, wt)
            for src in T.sources(T.G("x", m, rowptr, col, weights=False), 3):
                want = to_ref_sentinel(O.sssp(m, rowptr, col, wt, src))
                assert np.array_equal(ctx.sssp(G, src), want), (m, src)
                st = ctx.stats()
                assert st["edges_checked"] == st["edges_traversed"]
            G.close()
    finally:
        ctx.set_option("sssp_pull_div", 0)
        ctx.set_option("sssp_delta", -1)


@pytest.mark.parametrize("lh", [0, 1])
@pytest.mark.parametrize("wr", [(1, 255), (1, 10 ** 8), (200, 300)])
@pytest.mark.parametrize("key", ["rmat12", "lattice64", "directed"])
def test_sssp_light_heavy_vs_oracle(ctx, key, wr, lh):
    """Near-far rounds with heavy entries deferred to once per emptied
    frontier (sssp_light_heavy = 1) and without: the oracle's distances for
    widths that split the weights (some rows all light, some all heavy),
    with three widths in turn on one graph (a light-first copy each), on
    a directed graph and with weights far above the width."""
    rng = np.random.default_rng(7)
    if key == "directed":
        n = 3000
        u = rng.integers(0, n, 20000)
        v = rng.integers(0, n, 20000)
        k = np.unique((u * n + v)[u != v])
        rowptr, col, _ = O.coo_to_csr(n, (k // n).astype(np.int32), (k % n).astype(np.int32))
        m = n
    else:
        g = T.rmat(12) if key == "rmat12" else T.lattice(64)
        m, rowptr, col = g.m, g.rowptr, g.col
    wt = rng.integers(wr[0], wr[1] + 1, len(col)).astype(np.float32)
    G = ctx.upload(rowptr.astype(np.int32), col, wt)
    deg = np.diff(rowptr)
    srcs = [int(s) for s in np.nonzero(deg > 0)[0][[0, -1]]] + [int(np.argmax(deg))]
    want = {s: to_ref_sentinel(O.sssp(m, rowptr, col, wt, s)) for s in srcs}
    ctx.set_option("sssp_light_heavy", lh)
    try:
        for delta in (int(np.median(wt)), 64, 16):
            ctx.set_option("sssp_delta", delta)
            for src in srcs:
                got = ctx.sssp(G, src)
                assert np.array_equal(got, want[src]), (key, wr, lh, delta, src)
    finally:
        ctx.set_option("sssp_light_heavy", 1)
        ctx.set_option("sssp_delta", -1)
    G.close()


def test_sssp_widths_on_two_contexts_share_a_graph():
    """Two contexts on two threads run near-far SSSP with light / heavy
    rounds at different widths on one graph at the same time: each width
    has its own light-first copy, kept while the graph lives, so neither
    call's rounds read a copy the other call replaced; every result equals
    the oracle's."""
    import threading

    g = T.rmat(12)
    wt = np.random.default_rng(11).integers(1, 256, len(g.col)).astype(np.float32)
    deg = np.diff(g.rowptr)
    srcs = [int(s) for s in np.nonzero(deg > 0)[0][:: max(1, g.m // 12)]]
    want = {s: to_ref_sentinel(O.sssp(g.m, g.rowptr, g.col, wt, s)) for s in srcs}
    c1, c2 = A.Context(0), A.Context(0)
    try:
        G = c1.upload(g.rowptr32, g.col, wt)
        bad, errors = [], []

        def run(cx, delta):
            try:
                cx.set_option("sssp_light_heavy", 1)
                cx.set_option("sssp_delta", delta)
                for _ in range(3):
                    for s in srcs:
                        if not np.array_equal(cx.sssp(G, s), want[s]):
                            bad.append((delta, s))
            except Exception as e:  # surfaced by the asserts below
                errors.append(repr(e))

        th = [threading.Thread(target=run, args=a) for a in ((c1, 16), (c2, 64))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        assert not bad, bad
        G.close()
    finally:
        c1.close()
        c2.close()


def test_sssp_near_far_edge_cases(ctx):
    ctx.set_option("sssp_delta", 16)
    try:
        g = T.path_graph(200)  # one vertex per round, many splits
        G = ctx.upload(g.rowptr32, g.col, np.full(len(g.col), 3, np.float32))
        assert ctx.sssp(G, 0).tolist() == [3.0 * i for i in range(200)]
        G.close()
        g = T.empty_graph(5)  # isolated source: the first round is already a split
        G = ctx.upload(g.rowptr32, g.col, np.zeros(0, np.float32))
        got = ctx.sssp(G, 2)
        assert got[2] == 0 and np.all(np.delete(got, 2) == A.SSSP_INF)
        G.close()
    finally:
        ctx.set_option("sssp_delta", -1)


def test_sssp_edge_cases(ctx):
    g = T.path_graph(200)
    w = np.full(len(g.col), 3, np.float32)
    G = ctx.upload(g.rowptr32, g.col, w)
    assert ctx.sssp(G, 0).tolist() == [3.0 * i for i in range(200)]
    G.close()
    g = T.empty_graph(5)
    G = ctx.upload(g.rowptr32, g.col, np.zeros(0, np.float32))
    got = ctx.sssp(G, 2)
    assert got[2] == 0 and np.all(np.delete(got, 2) == A.SSSP_INF)
    G.close()
    g = T.fixture("test_bc")
    G = ctx.upload(g.rowptr32, g.col)  # no weights
    with pytest.raises(A.ApspieError):
        ctx.sssp(G, 0)
    G.close()


@pytest.mark.parametrize("bad", [-1.0, np.nan, -np.nan, np.inf, -np.inf])
def test_sssp_rejects_unorderable_weights(ctx, bad):
    """Distances are ordered by their bit patterns, valid for finite
    non-negative floats only: one negative, NaN or infinite weight makes both
    SSSP entry points refuse the graph (status -6) instead of answering."""
    g = T.rmat(10)
    w = g.w.copy()
    w[len(w) // 3] = bad
    G = ctx.upload(g.rowptr32, g.col, w)
    for call in (lambda: ctx.sssp(G, 1), lambda: ctx.sssp_batch(G, [1, 2, 3])):
        with pytest.raises(A.ApspieError) as e:
            call()
        assert e.value.status == -6
    assert ctx.bfs(G, 1)[1] == 0  # the graph itself stays usable
    G.close()
    w = g.w.copy()
    w[::7] = 0.0  # zero weights are fine
    G = ctx.upload(g.rowptr32, g.col, w)
    src = T.sources(g, 1)[0]
    assert np.array_equal(ctx.sssp(G, src), to_ref_sentinel(O.sssp(g.m, g.rowptr, g.col, w, src)))
    G.close()


def big_weights(g, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, len(g.col)).astype(np.float32)


@pytest.mark.parametrize("delta", [-1, 0, 16])
@pytest.mark.parametrize("lo,hi", [(1, 10 ** 8), (10 ** 9, 10 ** 10)])
@pytest.mark.parametrize("key", ["rmat12", "lattice64"])
def test_sssp_large_weights_vs_oracle(ctx, key, lo, hi, delta):
    """Weights up to 1e10 (sums far past 2^24, so FP32 rounds; the result is
    the smallest left-fold FP32 path sum, order-independent): the default
    width (chosen from the weights, no cap), Bellman-Ford rounds, and a width
    of 16, where the split threshold must jump to the pending distances."""
    g = T.rmat(int(key[4:])) if key.startswith("rmat") else T.lattice(int(key[7:]))
    w = big_weights(g, lo, hi, seed=lo % 101)
    ctx.set_option("sssp_delta", delta)
    try:
        G = ctx.upload(g.rowptr32, g.col, w)
        for src in T.sources(g, 3):
            want = to_ref_sentinel(O.sssp(g.m, g.rowptr, g.col, w, src))
            assert np.array_equal(ctx.sssp(G, src), want), (key, lo, delta, src)
            assert ctx.stats()["iterations"] <= 4 * g.m + 8
        G.close()
    finally:
        ctx.set_option("sssp_delta", -1)


@pytest.mark.parametrize("bdelta", [-1, 16])
@pytest.mark.parametrize("nsrc", [32, 128])
@pytest.mark.parametrize("lo,hi", [(1, 10 ** 8), (10 ** 9, 10 ** 10)])
def test_sssp_batch_large_weights_vs_oracle(ctx, lo, hi, nsrc, bdelta):
    """The batched near-far rounds with the same weights: default width and
    a width of 16 (threshold jumps), 32 and 128 sources (one and four source
    words per vertex)."""
    g = T.lattice(48)
    w = big_weights(g, lo, hi, seed=7)
    ctx.set_option("batch_sssp_delta", bdelta)
    try:
        G = ctx.upload(g.rowptr32, g.col, w)
        src = T.sources(g, nsrc, seed=11)
        got = ctx.sssp_batch(G, src)
        for i in range(0, nsrc, 7):
            want = to_ref_sentinel(O.sssp(g.m, g.rowptr, g.col, w, src[i]))
            assert np.array_equal(got[i], want), (lo, nsrc, bdelta, i)
        G.close()
    finally:
        ctx.set_option("batch_sssp_delta", -1)


# ----------------------------------------------------------------------- MIS
@pytest.mark.parametrize("name", T.FIXTURES)
def test_mis_fixtures(ctx, name):
    g = T.fixture(name)
    G = ctx.upload(g.rowptr32, g.col)
    got = ctx.mis(G)  # default priorities: mt19937(5489)
    assert got.tolist() == T.GOLD["fixtures"][name]["mis"]
    G.close()


@pytest.mark.parametrize("key", ["rmat10", "rmat14", "lattice64"])
def test_mis_synthetic_vs_oracle(ctx, key):
    g = T.rmat(int(key[4:])) if key.startswith("rmat") else T.lattice(int(key[7:]))
    G = ctx.upload(g.rowptr32, g.col)
    prio = O.mis_priorities(g.m)
    want = O.mis_greedy(g.m, g.rowptr, g.col, prio)
    assert np.array_equal(ctx.mis(G), want)
    assert np.array_equal(ctx.mis(G, prio), want)
    # other priorities, including heavy ties (tie -> lower id wins)
    for seed in (7, 11):
        p = O.mt_stream(seed, g.m)
        assert np.array_equal(ctx.mis(G, p), O.mis_greedy(g.m, g.rowptr, g.col, p))
    p = (O.mt_stream(3, g.m) % 4).astype(np.uint32)
    assert np.array_equal(ctx.mis(G, p), O.mis_greedy(g.m, g.rowptr, g.col, p))
    G.close()


def test_mis_properties_and_isolated(ctx):
    g = T.random_graph(3000, 0.001, 5)  # has isolated vertices
    G = ctx.upload(g.rowptr32, g.col)
    s = ctx.mis(G)
    rows = np.repeat(np.arange(g.m), np.diff(g.rowptr))
    assert set(np.unique(s).tolist()) <= {0, 1}
    assert not np.any((s[rows] == 1) & (s[g.col] == 1))
    cov = np.zeros(g.m, bool)
    cov[g.col[s[rows] == 1]] = True
    assert np.all((s == 1) | cov)
    assert np.all(s[g.deg == 0] == 1)
    G.close()


# ------------------------------------------------------------------------ TC
@pytest.mark.parametrize("name", T.FIXTURES)
def test_tc_fixtures(ctx, name):
    g = T.fixture(name)
    G = ctx.upload(g.rowptr32, g.col)
    total, rows = ctx.tc(G)
    fx = T.GOLD["fixtures"][name]
    assert total == fx["triangles"] == fx["nx_triangles"]
    assert rows.tolist() == fx["tc_rows"]
    G.close()


@pytest.mark.parametrize("key", ["rmat10", "rmat14", "lattice64"])
def test_tc_synthetic_vs_oracle(ctx, key):
    g = T.rmat(int(key[4:])) if key.startswith("rmat") else T.lattice(int(key[7:]))
    G = ctx.upload(g.rowptr32, g.col)
    want_total, want_rows = O.tc(g.m, g.rowptr, g.col)
    total, rows = ctx.tc(G)
    assert total == want_total == T.GOLD[key]["triangles"]
    assert np.array_equal(rows, want_rows)
    t2, _ = ctx.tc(G, per_row=False)
    assert t2 == total
    G.close()


def test_tc_all_bins():
    """Dense blocks push rows into the block and global-spill tiers."""
    from itertools import combinations
    for n in (40, 200, 2500, 5000):
        u, v = zip(*combinations(range(n), 2)) if n <= 200 else (None, None)
        if n <= 200:
            rp, c = O.edges_to_csr(n, np.array(u, np.int32), np.array(v, np.int32))
        else:  # the last 300 vertices form a clique adjacent to every other vertex
            uu = np.repeat(np.arange(n - 300, n), n - 300)
            vv = np.tile(np.arange(n - 300), 300)
            hu, hv = zip(*combinations(range(n - 300, n), 2))
            uu = np.concatenate([uu, hu])
            vv = np.concatenate([vv, hv])
            rp, c = O.edges_to_csr(n, uu.astype(np.int32), vv.astype(np.int32))
        with A.Context(0) as ctx:
            G = ctx.upload(rp.astype(np.int32), c)
            total, rows = ctx.tc(G)
            wt, wr = O.tc(n, rp, c)
            assert total == wt and np.array_equal(rows, wr), n


def test_tc_longest_rows_in_global_tables(ctx):
    """Rows with more than 16384 entries below the diagonal keep their table in
    global memory: a wheel (hub = the highest id, rim path) has one triangle
    per rim edge, all counted at the hub; a second hub in the middle adds
    upper pairs."""
    n = 20000
    rim = np.arange(n - 2)
    u = np.concatenate([rim[:-1], rim, np.full(9000, 9000)])
    v = np.concatenate([rim[1:], np.full(n - 2, n - 1), np.arange(9001, 18001)])
    rp, c = O.edges_to_csr(n, u.astype(np.int32), v.astype(np.int32))
    G = ctx.upload(rp.astype(np.int32), c)
    total, rows = ctx.tc(G)
    wt, wr = O.tc(n, rp, c)
    assert total == wt and np.array_equal(rows, wr)
    assert rows[n - 1] >= n - 3  # the rim triangles at the hub
    G.close()


def test_tc_rejects_unsorted_rows(ctx):
    rp = np.array([0, 2, 3, 4], np.int32)
    col = np.array([2, 1, 0, 0], np.int32)  # row 0 unsorted
    G = ctx.upload(rp, col)
    with pytest.raises(A.ApspieError):
        ctx.tc(G)
    G.close()
