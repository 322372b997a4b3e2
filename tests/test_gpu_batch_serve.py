"""Batched multi-source traversal, the autotuner and the query server on the
GPU.  Every batched / served result must equal the single-source answer of
the oracle for that source, alone and in the mix."""
import numpy as np
import pytest

import apspie_amd as A
from apspie_amd import server as S
from oracle import oracle as O
from tests import graphs as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with A.Context(0) as c:
        yield c


def oracle_sssp(g, s):
    d = O.sssp(g.m, g.rowptr, g.col, g.w, s)
    return np.where(d == -1, A.SSSP_INF, d).astype(np.float32)


@pytest.mark.parametrize("nsrc", [1, 5, 32, 33, 100, 128, 200])
@pytest.mark.parametrize("key", ["rmat12", "lattice64"])
def test_bfs_batch_vs_oracle(ctx, key, nsrc):
    g = T.rmat(12) if key == "rmat12" else T.lattice(64)
    G = ctx.upload(g.rowptr32, g.col, g.w)
    src = T.sources(g, nsrc)
    if nsrc > 3:
        src[3] = src[1]  # a repeated source
    got = ctx.bfs_batch(G, src)
    for i, s in enumerate(src):
        want, _ = O.bfs(g.m, g.rowptr, g.col, s)
        assert np.array_equal(got[i], want), (key, nsrc, i, s)
    G.close()


@pytest.mark.parametrize("pull_div", [0, 1, 4, 1 << 20])
@pytest.mark.parametrize("nsrc", [5, 33, 200])
@pytest.mark.parametrize("key", ["rmat12", "lattice64", "star"])
def test_bfs_batch_directions(ctx, key, nsrc, pull_div):
    """Push only (0), pull on big levels (1, 4) and pull on every level after
    the first (2^20) give the single-source depths."""
    g = {"rmat12": lambda: T.rmat(12), "lattice64": lambda: T.lattice(64),
         "star": lambda: T.star_graph(3000)}[key]()
    ctx.set_option("batch_pull_div", pull_div)
    try:
        G = ctx.upload(g.rowptr32, g.col)
        src = T.sources(g, nsrc, seed=23)
        got = ctx.bfs_batch(G, src)
        for i, s in enumerate(src):
            assert np.array_equal(got[i], O.bfs(g.m, g.rowptr, g.col, s)[0]), (key, i, s)
        G.close()
    finally:
        ctx.set_option("batch_pull_div", 4)


@pytest.mark.parametrize("snap_mb", [0, 1, 24576])
@pytest.mark.parametrize("key", ["rmat12", "lattice64"])
def test_bfs_batch_level_snapshots(ctx, key, snap_mb):
    """Pulled levels record their discoveries as bit-row snapshots within
    batch_snapshot_mb: none (depths written directly), a few (1 MB: the
    later pulled levels fall back to direct writes) or all."""
    g = T.rmat(12) if key == "rmat12" else T.lattice(64)
    ctx.set_option("batch_snapshot_mb", snap_mb)
    ctx.set_option("batch_pull_div", 1 << 20)
    try:
        G = ctx.upload(g.rowptr32, g.col)
        for nsrc in (7, 100, 300):
            src = T.sources(g, nsrc, seed=41)
            got = ctx.bfs_batch(G, src)
            for i, s in enumerate(src):
                assert np.array_equal(got[i], O.bfs(g.m, g.rowptr, g.col, s)[0]), (key, i, s)
        G.close()
    finally:
        ctx.set_option("batch_snapshot_mb", 24576)
        ctx.set_option("batch_pull_div", 4)


@pytest.mark.parametrize("nsrc", [3, 40, 140])
def test_batch_isolated_and_unreached_sources(ctx, nsrc):
    """Sources that are isolated vertices (marked visited up front in a
    symmetric graph) reach only themselves; vertices of other components stay
    -1 / unreached; a source repeated in the batch gets the same row."""
    g = T.random_graph(3000, 0.0008, 13)  # several components, isolated vertices
    iso = np.nonzero(g.deg == 0)[0]
    assert len(iso) >= 2
    conn = T.sources(g, nsrc, seed=19)
    src = (list(iso[:2]) + conn)[:nsrc]
    src[-1] = src[0]  # a repeated (isolated) source
    G = ctx.upload(g.rowptr32, g.col, g.w)
    got = ctx.bfs_batch(G, src)
    dist = ctx.sssp_batch(G, src)
    for i, s in enumerate(src):
        assert np.array_equal(got[i], O.bfs(g.m, g.rowptr, g.col, s)[0]), (i, s)
        assert np.array_equal(dist[i], oracle_sssp(g, s)), (i, s)
    G.close()


def test_bfs_batch_directed_graph_pushes(ctx):
    """A non-symmetric CSR is never pulled (rows are out-edges only)."""
    rng = np.random.default_rng(9)
    n = 2000
    u = rng.integers(0, n, 12000)
    v = rng.integers(0, n, 12000)
    key = np.unique((u * n + v)[(u != v) & (u < 1700)])
    rp, c, _ = O.coo_to_csr(n, (key // n).astype(np.int32), (key % n).astype(np.int32))
    ctx.set_option("batch_pull_div", 1 << 20)
    try:
        G = ctx.upload(rp.astype(np.int32), c)
        src = [0, 5, 1800, 77] * 10
        got = ctx.bfs_batch(G, src)
        for i, s in enumerate(src):
            assert np.array_equal(got[i], O.bfs(n, rp, c, s)[0]), (i, s)
        G.close()
    finally:
        ctx.set_option("batch_pull_div", 4)


@pytest.mark.parametrize("nsrc", [1, 7, 32, 64, 150])
@pytest.mark.parametrize("key", ["rmat11", "lattice32"])
def test_sssp_batch_vs_oracle(ctx, key, nsrc):
    g = T.rmat(11) if key == "rmat11" else T.lattice(32)
    G = ctx.upload(g.rowptr32, g.col, g.w)
    src = T.sources(g, nsrc)
    got = ctx.sssp_batch(G, src)
    for i, s in enumerate(src):
        assert np.array_equal(got[i], oracle_sssp(g, s)), (key, nsrc, i, s)
    G.close()


@pytest.mark.parametrize("pull_div", [0, 4, 1 << 20])
@pytest.mark.parametrize("nsrc", [20, 100, 400])
@pytest.mark.parametrize("key", ["lattice64", "random", "asym"])
def test_sssp_batch_pulled_rounds(ctx, key, nsrc, pull_div):
    """Pulled near-far rounds (rows of at most 64 entries, symmetric weights):
    never (0), on wide rounds (4) and on every round with work (2^20), at
    one, four and sixteen source words per vertex; a graph with isolated
    vertices and small components; and asymmetric weights, which must keep
    pushing -- every row equals the oracle."""
    if key == "random":
        g = T.random_graph(4000, 0.0008, 13)
        w = g.w
    else:
        g = T.lattice(64)
        w = g.w.copy()
        if key == "asym":  # w(u, v) != w(v, u) on some edges
            w[::7] += 3
    ctx.set_option("batch_sssp_pull_div", pull_div)
    ctx.set_option("batch_sssp_delta", 64)
    try:
        G = ctx.upload(g.rowptr32, g.col, w)
        src = T.sources(g, nsrc, seed=31)
        got = ctx.sssp_batch(G, src)
        for i in range(0, nsrc, 3):
            d = O.sssp(g.m, g.rowptr, g.col, w, src[i])
            want = np.where(d == -1, A.SSSP_INF, d).astype(np.float32)
            assert np.array_equal(got[i], want), (key, nsrc, pull_div, i)
        G.close()
    finally:
        ctx.set_option("batch_sssp_pull_div", 4)
        ctx.set_option("batch_sssp_delta", -1)


@pytest.mark.parametrize("delta", [-1, 0, 1, 16, 256, 1 << 20])
@pytest.mark.parametrize("nsrc", [7, 64, 150])
@pytest.mark.parametrize("key", ["rmat11", "lattice32", "path"])
def test_sssp_batch_near_far(ctx, key, nsrc, delta):
    """Batched near-far rounds (batch_sssp_delta > 0, one threshold for all
    sources) and the Bellman-Ford frontier (0) give the single-source
    distances; the path graph needs many split rounds in a row."""
    if key == "path":
        g = T.path_graph(300)
        w = np.full(len(g.col), 7, np.float32)
    else:
        g = T.rmat(11) if key == "rmat11" else T.lattice(32)
        w = g.w
    ctx.set_option("batch_sssp_delta", delta)
    try:
        G = ctx.upload(g.rowptr32, g.col, w)
        src = T.sources(g, nsrc, seed=29)
        got = ctx.sssp_batch(G, src)
        for i, s in enumerate(src):
            d = O.sssp(g.m, g.rowptr, g.col, w, s)
            want = np.where(d == -1, A.SSSP_INF, d).astype(np.float32)
            assert np.array_equal(got[i], want), (key, nsrc, delta, i, s)
        G.close()
    finally:
        ctx.set_option("batch_sssp_delta", -1)


@pytest.mark.parametrize("graph_on,lps", [(0, 6), (1, 6), (1, 48)])
def test_batch_graph_replay_matches(ctx, graph_on, lps):
    g = T.lattice(64)  # ~100 levels: many chunk replays
    ctx.set_option("batch_graph", graph_on)
    ctx.set_option("bfs_levels_per_sync", lps)
    try:
        G = ctx.upload(g.rowptr32, g.col, g.w)
        src = T.sources(g, 40)
        got = ctx.bfs_batch(G, src)
        gs = ctx.sssp_batch(G, src[:9])
        for i, s in enumerate(src):
            assert np.array_equal(got[i], O.bfs(g.m, g.rowptr, g.col, s)[0])
        for i, s in enumerate(src[:9]):
            assert np.array_equal(gs[i], oracle_sssp(g, s))
        G.close()
    finally:
        ctx.set_option("batch_graph", 1)
        ctx.set_option("bfs_levels_per_sync", 8)


def test_batch_many_sources_split(ctx):
    g = T.rmat(10)
    G = ctx.upload(g.rowptr32, g.col)
    src = (T.sources(g, 300) * 2)[:600]  # > 512: two passes
    got = ctx.bfs_batch(G, src)
    for i in (0, 299, 511, 512, 599):
        assert np.array_equal(got[i], O.bfs(g.m, g.rowptr, g.col, src[i])[0])
    G.close()


def test_autotune_sets_valid_options():
    g = T.rmat(12)
    with A.Context(0) as c:
        G = c.upload(g.rowptr32, g.col, g.w)
        assert c.get_option("sssp_delta") == -1  # chosen per graph until tuned
        c.autotune(G, 0)
        assert c.get_option("bfs_tile") in (1024, 2048, 4096)
        assert c.get_option("bfs_blocks_per_cu") in (2, 4, 8)
        assert c.get_option("bfs_alpha") in (7, 14, 28)
        # Bellman-Ford, or half / once / twice the width chosen from the weights
        # (10 * 255 / mean row length, to the nearest power of two)
        want = 10 * 255 / (len(g.col) / g.m)
        d = 16
        while d < 4096 and d * 2 ** 0.5 < want:
            d *= 2
        d = max(d, 64)
        assert c.get_option("sssp_delta") in (0, d // 2, d, 2 * d)
        for s in T.sources(g, 2):
            assert np.array_equal(c.sssp(G, s), oracle_sssp(g, s))
        c.autotune(G, 32)
        assert c.get_option("batch_blocks_per_cu") in (2, 4, 8)
        assert c.get_option("batch_pull_div") in (2, 4, 8)
        assert c.get_option("batch_sssp_delta") in (0, 256, 1024)
        src = T.sources(g, 4)
        for s in src:
            assert np.array_equal(c.bfs(G, s), O.bfs(g.m, g.rowptr, g.col, s)[0])
        got = c.sssp_batch(G, src)
        for i, s in enumerate(src):
            assert np.array_equal(got[i], oracle_sssp(g, s))
        G.close()


def test_server_mixed_stream_equals_alone():
    graphs = [T.rmat(12), T.lattice(64), T.fixture("chesapeake")]
    with S.Server(slots=3, max_batch=16, hbm_budget=1 << 30) as srv:
        ids = [srv.add_graph(g.rowptr32, g.col, g.w) for g in graphs]
        rng = np.random.default_rng(0)
        reqs = []
        for _ in range(200):
            gi = int(rng.integers(0, len(graphs)))
            r = rng.random()
            kind = A.Q_BFS if r < 0.6 else (A.Q_SSSP if r < 0.9 else A.Q_TC)
            src = int(rng.choice(np.nonzero(graphs[gi].deg)[0]))
            reqs.append((gi, kind, src, srv.submit(ids[gi], kind, src)))
        lat = []
        for gi, kind, src, t in reqs:
            res, ms = srv.wait(t, kind, graphs[gi].m)
            lat.append(ms)
            g = graphs[gi]
            if kind == A.Q_BFS:
                assert np.array_equal(res, O.bfs(g.m, g.rowptr, g.col, src)[0])
            elif kind == A.Q_SSSP:
                assert np.array_equal(res, oracle_sssp(g, src))
            else:
                assert res == O.tc(g.m, g.rowptr, g.col)[0]
        assert min(lat) > 0
        srv.drain()


def test_server_results_into_caller_memory():
    """submit_into: results written straight into the caller's arrays, device
    and host, batched and single; each equals the oracle, and the same stream
    served with the pinned-buffer results gives identical bytes.  Also the
    one-query-per-launch form (max_batch 1) and a TC query among them."""
    graphs = [T.rmat(12), T.lattice(64)]
    rng = np.random.default_rng(3)
    stream = []
    for _ in range(150):
        gi = int(rng.integers(0, len(graphs)))
        r = rng.random()
        kind = A.Q_BFS if r < 0.55 else (A.Q_SSSP if r < 0.95 else A.Q_TC)
        stream.append((gi, kind, int(rng.choice(np.nonzero(graphs[gi].deg)[0]))))
    with A.Context(0) as ctx:
        for max_batch in (32, 1):
            with S.Server(slots=3, max_batch=max_batch, hbm_budget=1 << 30) as srv:
                ids = [srv.add_graph(g.rowptr32, g.col, g.w) for g in graphs]
                bufs, tickets = [], []
                for i, (gi, kind, src) in enumerate(stream):
                    m = graphs[gi].m
                    if kind == A.Q_TC:
                        bufs.append(None)
                        tickets.append(srv.submit_into(ids[gi], kind, src, None, A.DEVICE))
                    elif i % 3 == 0:  # host memory of the caller's own
                        h = np.full(m, 99, np.int32 if kind == A.Q_BFS else np.float32)
                        bufs.append(h)
                        tickets.append(srv.submit_into(ids[gi], kind, src, h.ctypes.data,
                                                       A.HOST))
                    else:
                        d = ctx.dev_alloc(4 * m)
                        bufs.append(d)
                        tickets.append(srv.submit_into(ids[gi], kind, src, d.ptr, A.DEVICE))
                for i, (gi, kind, src) in enumerate(stream):
                    g = graphs[gi]
                    if kind == A.Q_TC:
                        got, _ = srv.wait(tickets[i], kind, g.m)
                        assert got == O.tc(g.m, g.rowptr, g.col)[0]
                        continue
                    assert srv.wait_done(tickets[i]) > 0
                    b = bufs[i]
                    dt = np.int32 if kind == A.Q_BFS else np.float32
                    got = b if isinstance(b, np.ndarray) else b.to_host(dt, g.m)
                    want = (O.bfs(g.m, g.rowptr, g.col, src)[0] if kind == A.Q_BFS
                            else oracle_sssp(g, src))
                    assert np.array_equal(got, want), (max_batch, i)
                    # the pinned-buffer form of the same query
                    res, _ = srv.wait(srv.submit(ids[gi], kind, src), kind, g.m)
                    assert res.tobytes() == got.tobytes()
                    if not isinstance(b, np.ndarray):
                        b.free()
                # a BFS / SSSP query needs an output
                with pytest.raises(A.ApspieError):
                    srv.submit_into(ids[0], A.Q_BFS, 0, None, A.DEVICE)


def test_server_tight_budget_sizes_batches():
    """A worker budget that holds only part of what a batch of 128 would
    need: batches are sized to fit (or split after a budget failure), and
    every result still equals the oracle's."""
    g = T.lattice(64)
    with S.Server(slots=1, max_batch=128, hbm_budget=4 << 20) as srv:
        gi = srv.add_graph(g.rowptr32, g.col, g.w)
        src = T.sources(g, 90, seed=9)
        tickets = [(s, A.Q_BFS, srv.submit(gi, A.Q_BFS, s)) for s in src]
        tickets += [(s, A.Q_SSSP, srv.submit(gi, A.Q_SSSP, s)) for s in src[:30]]
        for s, kind, t in tickets:
            res, _ = srv.wait(t, kind, g.m)
            want = O.bfs(g.m, g.rowptr, g.col, s)[0] if kind == A.Q_BFS else oracle_sssp(g, s)
            assert np.array_equal(res, want), (kind, s)
        srv.drain()


def test_server_budget_not_granule_multiple():
    """A budget that is not a multiple of 256 bytes still gives every worker
    an arena (rounded down) and every result equals the oracle's, over a
    stream that switches graphs and kinds so that batches reuse the arenas'
    ranges."""
    g1, g2 = T.lattice(64), T.rmat(11)
    with S.Server(slots=2, max_batch=64, hbm_budget=(256 << 20) + 100) as srv:
        ids = [srv.add_graph(g.rowptr32, g.col, g.w) for g in (g1, g2)]
        tickets = []
        for rep in range(3):
            for gi, g in zip(ids, (g1, g2)):
                for s in T.sources(g, 20, seed=rep):
                    tickets.append((g, s, A.Q_BFS, srv.submit(gi, A.Q_BFS, s)))
                for s in T.sources(g, 6, seed=rep + 7):
                    tickets.append((g, s, A.Q_SSSP, srv.submit(gi, A.Q_SSSP, s)))
        for g, s, kind, t in tickets:
            res, _ = srv.wait(t, kind, g.m)
            want = O.bfs(g.m, g.rowptr, g.col, s)[0] if kind == A.Q_BFS else oracle_sssp(g, s)
            assert np.array_equal(res, want), (kind, s)
        srv.drain()


def test_server_zero_copy_views():
    """wait(copy=False) hands out read-only views of pinned result buffers:
    equal to the copying wait, intact while later requests run, returned to
    the pool when collected (a released ticket cannot be released again)."""
    import gc
    g = T.lattice(64)
    with S.Server(slots=2, max_batch=8, hbm_budget=1 << 30) as srv:
        gi = srv.add_graph(g.rowptr32, g.col, g.w)
        src = T.sources(g, 12, seed=5)
        tickets = [srv.submit(gi, A.Q_BFS, s) for s in src]
        views = [srv.wait(t, A.Q_BFS, g.m, copy=False)[0] for t in tickets]
        more = [srv.submit(gi, A.Q_SSSP, s) for s in src]  # reuse pressure on the pool
        for t, s in zip(more, src):
            d, _ = srv.wait(t, A.Q_SSSP, g.m, copy=False)
            assert np.array_equal(d, oracle_sssp(g, s))
        for v, s in zip(views, src):
            assert not v.flags.writeable
            assert np.array_equal(v, O.bfs(g.m, g.rowptr, g.col, s)[0])
        t = srv.submit(gi, A.Q_TC)
        assert srv.wait(t, A.Q_TC, g.m, copy=False)[0] == O.tc(g.m, g.rowptr, g.col)[0]
        del views, d
        gc.collect()
        with pytest.raises(A.ApspieError):  # already released by the collector
            A._check(srv._L.apspie_server_release(srv.h, tickets[0]))


def test_server_budget_and_bad_requests():
    g = T.rmat(12)
    with S.Server(slots=1, max_batch=4, hbm_budget=100_000) as srv:
        with pytest.raises(A.ApspieError) as e:
            srv.add_graph(g.rowptr32, g.col)
        assert e.value.status == -4
        gi = srv.add_graph(T.fixture("test_bc").rowptr32, T.fixture("test_bc").col)
        with pytest.raises(A.ApspieError):
            srv.submit(gi, A.Q_BFS, 99)
        with pytest.raises(A.ApspieError):
            srv.submit(gi + 5, A.Q_BFS, 0)


@pytest.mark.parametrize("group", [32, 128])
def test_bfs_batch_sources_in_small_components(ctx, group):
    """Sources spread over a giant component and many small ones (and
    isolated vertices): pairs in different components start settled and come
    out unreached; every row equals the plain BFS.  Also with the long-row
    graph's batch grouping widened to 128 sources."""
    big = T.rmat(12)
    small = T.random_graph(3000, 0.0004, 3)  # many small components
    m = big.m + small.m
    rows = np.concatenate([np.repeat(np.arange(big.m), np.diff(big.rowptr)),
                           big.m + np.repeat(np.arange(small.m), np.diff(small.rowptr))])
    cols = np.concatenate([big.col, big.m + small.col])
    G = ctx.from_coo(m, rows.astype(np.int32), cols.astype(np.int32), symmetrize=True)
    rng = np.random.default_rng(9)
    srcs = [int(x) for x in rng.choice(m, 100, replace=False)]
    ctx.set_option("batch_long_row_group", group)
    try:
        got = ctx.bfs_batch(G, srcs)
        for i, s in enumerate(srcs):
            assert np.array_equal(got[i], ctx.bfs(G, s)), (group, i, s)
    finally:
        ctx.set_option("batch_long_row_group", 256)
        G.close()


def test_context_arena_matches_pool():
    """Scratch carved from an arena gives the results of per-buffer
    allocations, over repeated calls that reuse its ranges; the arena counts
    against the HBM budget; a buffer beyond its largest free range is a
    budget error; 0 releases it and calls allocate per buffer again."""
    g = T.rmat(12)
    src = T.sources(g, 40, seed=5)
    want_bfs = [O.bfs(g.m, g.rowptr, g.col, s)[0] for s in src]
    want_sssp = [oracle_sssp(g, s) for s in src[:20]]
    want_tri = O.tc(g.m, g.rowptr, g.col)[0]
    with A.Context(0) as c:
        G = c.upload(g.rowptr32, g.col, g.w)
        used0 = c.hbm_used()
        c.reserve_arena((64 << 20) + 100)  # rounded down to whole 256-byte granules
        assert c.arena_bytes() == 64 << 20
        assert c.hbm_used() == used0 + (64 << 20)
        for _ in range(2):
            for i, s in enumerate(src[:4]):
                assert np.array_equal(c.bfs(G, s), want_bfs[i]), s
                assert np.array_equal(c.sssp(G, s), want_sssp[i]), s
            got = c.bfs_batch(G, src)
            for i in range(len(src)):
                assert np.array_equal(got[i], want_bfs[i]), i
            got = c.sssp_batch(G, src[:20])
            for i in range(20):
                assert np.array_equal(got[i], want_sssp[i]), i
            assert c.tc(G, per_row=False)[0] == want_tri
        # 40 sources of SSSP hold a 4096 x 64 distance matrix (1 MiB) alone
        c.reserve_arena(1 << 20)
        with pytest.raises(A.ApspieError) as e:
            c.sssp_batch(G, src)
        assert e.value.status == -4  # APSPIE_ERR_BUDGET
        c.trim()
        c.reserve_arena(0)
        assert c.arena_bytes() == 0
        assert c.hbm_used() == used0
        got = c.sssp_batch(G, src)
        for i in range(20):
            assert np.array_equal(got[i], want_sssp[i]), i
        G.close()
