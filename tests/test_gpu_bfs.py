"""BFS on the GPU through the C ABI, bit-exact against the oracle
(SimpleReferenceBfs restated) and the golden fixtures."""
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


def run_bfs(ctx, g, src, **opts):
    for k, v in opts.items():
        ctx.set_option(k, v)
    G = ctx.upload(g.rowptr32, g.col)
    try:
        return ctx.bfs(G, src)
    finally:
        G.close()


@pytest.mark.parametrize("name", T.FIXTURES)
def test_bfs_fixtures(ctx, name):
    g = T.fixture(name)
    got = run_bfs(ctx, g, 0)
    assert got.tolist() == T.GOLD["fixtures"][name]["bfs"]


@pytest.mark.parametrize("tile", [1024, 2048, 4096])
@pytest.mark.parametrize("key", ["rmat10", "rmat14", "lattice64"])
def test_bfs_synthetic_vs_oracle(ctx, key, tile):
    g = T.rmat(int(key[4:])) if key.startswith("rmat") else T.lattice(int(key[7:]))
    ctx.set_option("bfs_tile", tile)
    G = ctx.upload(g.rowptr32, g.col)
    try:
        for src in [T.GOLD[key]["source"]] + T.sources(g, 5):
            want, _ = O.bfs(g.m, g.rowptr, g.col, src)
            got = ctx.bfs(G, src)
            assert np.array_equal(got, want), (key, tile, src)
            st = ctx.stats()
            reached = want >= 0
            assert st["edges_traversed"] == int(g.deg[reached].sum())
    finally:
        G.close()
        ctx.set_option("bfs_tile", 2048)


@pytest.mark.parametrize("direction", [0, 1, 2])
@pytest.mark.parametrize("key", ["rmat10", "rmat14", "lattice64", "test_cc", "star"])
def test_bfs_directions_vs_oracle(ctx, key, direction):
    """push only (0), direction-optimising (1) and pull whenever possible (2)
    give the same depths."""
    if key == "star":
        g = T.star_graph(3000)
    elif key == "test_cc":
        g = T.fixture("test_cc")
    else:
        g = T.rmat(int(key[4:])) if key.startswith("rmat") else T.lattice(int(key[7:]))
    ctx.set_option("bfs_direction", direction)
    G = ctx.upload(g.rowptr32, g.col)
    try:
        for src in T.sources(g, 4, seed=17):
            want, _ = O.bfs(g.m, g.rowptr, g.col, src)
            assert np.array_equal(ctx.bfs(G, src), want), (key, direction, src)
            reached = want >= 0
            assert ctx.stats()["edges_traversed"] == int(g.deg[reached].sum())
    finally:
        G.close()
        ctx.set_option("bfs_direction", 1)


def test_bfs_directed_graph_falls_back_to_push(ctx):
    """Pulling reads rows as in-edges, so a non-symmetric CSR must be
    traversed by pushing only; sinks (out-degree 0) must still be reached."""
    rng = np.random.default_rng(5)
    n = 3000
    u = rng.integers(0, n, 20000).astype(np.int32)
    v = rng.integers(0, n, 20000).astype(np.int32)
    keep = (u != v) & (u < 2500)  # vertices >= 2500 are sinks
    rp, c, _ = O.coo_to_csr(n, u[keep], v[keep])
    for r in range(n):  # sorted rows, no duplicates
        row = np.unique(c[rp[r]:rp[r + 1]])
        c[rp[r]:rp[r] + len(row)] = row
    uu = np.repeat(np.arange(n), np.diff(rp))
    key = np.unique(uu.astype(np.int64) * n + c)
    rp, c, _ = O.coo_to_csr(n, (key // n).astype(np.int32), (key % n).astype(np.int32))
    G = ctx.upload(rp.astype(np.int32), c)
    for d in (0, 1, 2):
        ctx.set_option("bfs_direction", d)
        for src in (0, 7, 2600):
            want, _ = O.bfs(n, rp, c, src)
            assert np.array_equal(ctx.bfs(G, src), want), (d, src)
    ctx.set_option("bfs_direction", 1)
    G.close()


def test_bfs_edge_cases(ctx):
    # star: one hub expands everything in one level; path: one vertex per level
    for g, src in [(T.star_graph(5000), 0), (T.star_graph(5000), 17), (T.path_graph(300), 0),
                   (T.path_graph(300), 150), (T.random_graph(2000, 0.002, 1), 3)]:
        want, _ = O.bfs(g.m, g.rowptr, g.col, src)
        assert np.array_equal(run_bfs(ctx, g, src), want), g
    # isolated source and an edgeless graph
    g = T.empty_graph(10)
    got = run_bfs(ctx, g, 4)
    assert got.tolist() == [-1, -1, -1, -1, 0, -1, -1, -1, -1, -1]
    g = T.fixture("test_cc")
    assert run_bfs(ctx, g, 8).tolist() == O.bfs(g.m, g.rowptr, g.col, 8)[0].tolist()


def test_bfs_rejects_bad_source(ctx):
    g = T.fixture("test_bc")
    G = ctx.upload(g.rowptr32, g.col)
    with pytest.raises(A.ApspieError) as e:
        ctx.bfs(G, g.m)
    assert e.value.status == -1
    with pytest.raises(A.ApspieError):
        ctx.bfs(G, -1)
    G.close()


def test_bfs_repeat_is_deterministic(ctx):
    g = T.rmat(12)
    G = ctx.upload(g.rowptr32, g.col)
    a = ctx.bfs(G, 0)
    for _ in range(3):
        assert np.array_equal(ctx.bfs(G, 0), a)
    G.close()


@pytest.mark.parametrize("levels_per_sync", [1, 3, 64])
def test_bfs_sync_chunking(ctx, levels_per_sync):
    g = T.lattice(64)
    ctx.set_option("bfs_levels_per_sync", levels_per_sync)
    try:
        want, _ = O.bfs(g.m, g.rowptr, g.col, 0)
        assert np.array_equal(run_bfs(ctx, g, 0), want)
    finally:
        ctx.set_option("bfs_levels_per_sync", 8)
