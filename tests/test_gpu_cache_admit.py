"""Graph ingestion on the device and admission into the server: .mtx -> COO ->
CSR built on the GPU (apspie_graph_from_coo), saved to and loaded from the
binary CSR cache, and admitted into a running server from COO or from the
cache while queries are in flight; every graph equals the oracle's CSR and
every answer the oracle's."""
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


def oracle_sssp(m, rp, col, w, s):
    d = O.sssp(m, rp, col, w, s)
    return np.where(d == -1, A.SSSP_INF, d).astype(np.float32)


@pytest.mark.parametrize("name", T.FIXTURES)
def test_mtx_coo_cache_roundtrip(ctx, tmp_path, name):
    """read_mtx + device build (symmetrize, dedup, loops dropped) -> save ->
    load gives the oracle's CSR (mtx_graph) exactly, values included."""
    path = T.GOLDEN_DIR / "mtx" / f"{name}.mtx"
    m, n, row, col, val = A.read_mtx(path)
    G = ctx.from_coo(max(m, n), row, col, val, symmetrize=True)
    om, orp, oc, ov = O.mtx_graph(path.read_text())
    rp, c, v = G.download(values=True)
    assert np.array_equal(rp, orp.astype(np.int32)) and np.array_equal(c, oc)
    assert np.array_equal(v, ov)
    G.save(tmp_path / "g.csr")
    H = ctx.load(tmp_path / "g.csr")
    rp2, c2, v2 = H.download(values=True)
    assert np.array_equal(rp2, rp) and np.array_equal(c2, c) and np.array_equal(v2, v)
    src = int(np.argmax(np.diff(orp)))
    assert np.array_equal(ctx.bfs(H, src), O.bfs(om, orp, oc, src)[0])
    G.close()
    H.close()


def test_rmat14_cache_roundtrip(ctx, tmp_path):
    """R-MAT-14 from COO on the device equals the host generator's CSR; the
    cache keeps the weights; BFS / SSSP on the loaded graph equal the oracle."""
    g = T.rmat(14)
    rows = np.repeat(np.arange(g.m, dtype=np.int32), np.diff(g.rowptr))
    perm = np.random.default_rng(1).permutation(len(rows))  # any input order
    G = ctx.from_coo(g.m, rows[perm], g.col[perm], symmetrize=True)
    rp, c, _ = G.download()
    assert np.array_equal(rp, g.rowptr32) and np.array_equal(c, g.col)
    W = ctx.upload(g.rowptr32, g.col, g.w)
    W.save(tmp_path / "w.csr")
    H = ctx.load(tmp_path / "w.csr")
    assert H.nnz == len(g.col) and H.m == g.m
    for s in T.sources(g, 3):
        assert np.array_equal(ctx.bfs(H, s), O.bfs(g.m, g.rowptr, g.col, s)[0])
        assert np.array_equal(ctx.sssp(H, s), oracle_sssp(g.m, g.rowptr, g.col, g.w, s))
    bad = tmp_path / "bad.csr"
    raw = bytearray((tmp_path / "w.csr").read_bytes())
    raw[1000] ^= 0x10
    bad.write_bytes(bytes(raw))
    with pytest.raises(A.ApspieError):
        ctx.load(bad)
    for X in (G, W, H):
        X.close()


def test_server_admits_from_coo_and_cache_while_busy(tmp_path):
    """A server busy with queries on one graph admits a second graph built on
    the device from COO and a third loaded from the cache; the queries in
    flight and those on the new graphs all equal the oracle."""
    g0, g1, g2 = T.lattice(64), T.rmat(12), T.fixture("chesapeake")
    A.csr_write(tmp_path / "g2.csr", g2.rowptr32, g2.col, g2.w)
    with S.Server(slots=3, max_batch=16, hbm_budget=1 << 30) as srv:
        i0 = srv.add_graph(g0.rowptr32, g0.col, g0.w)
        src0 = T.sources(g0, 60, seed=2)
        first = [(s, k, srv.submit(i0, k, s)) for s in src0 for k in (A.Q_BFS, A.Q_SSSP)]
        rows1 = np.repeat(np.arange(g1.m, dtype=np.int32), np.diff(g1.rowptr))
        low = g1.col < rows1  # one triangle of the pattern; the build mirrors it
        i1 = srv.add_graph_coo(g1.m, rows1[low], g1.col[low], symmetrize=True)
        i2 = srv.add_graph_file(tmp_path / "g2.csr")
        later = [(i1, g1, s, srv.submit(i1, A.Q_BFS, s)) for s in T.sources(g1, 20)]
        later += [(i2, g2, s, srv.submit(i2, A.Q_SSSP, s)) for s in T.sources(g2, 10)]
        tc1 = srv.submit(i1, A.Q_TC)
        for s, k, t in first:
            res, _ = srv.wait(t, k, g0.m)
            want = (O.bfs(g0.m, g0.rowptr, g0.col, s)[0] if k == A.Q_BFS
                    else oracle_sssp(g0.m, g0.rowptr, g0.col, g0.w, s))
            assert np.array_equal(res, want), (k, s)
        for gi, g, s, t in later:
            k = A.Q_BFS if gi == i1 else A.Q_SSSP
            res, _ = srv.wait(t, k, g.m)
            want = (O.bfs(g.m, g.rowptr, g.col, s)[0] if k == A.Q_BFS
                    else oracle_sssp(g.m, g.rowptr, g.col, g.w, s))
            assert np.array_equal(res, want), (gi, s)
        assert srv.wait(tc1, A.Q_TC, g1.m)[0] == O.tc(g1.m, g1.rowptr, g1.col)[0]
        with pytest.raises(A.ApspieError) as e:
            srv.add_graph_file(tmp_path / "missing.csr")
        assert e.value.status == -1
        srv.drain()


def test_server_admission_budget_uses_built_size():
    """The budget is checked against the graph as built: a COO whose
    symmetrized CSR does not fit is refused (status -4) and leaves the pool
    usable."""
    g = T.rmat(12)
    rows = np.repeat(np.arange(g.m, dtype=np.int32), np.diff(g.rowptr))
    with S.Server(slots=1, max_batch=4, hbm_budget=400_000) as srv:
        with pytest.raises(A.ApspieError) as e:
            srv.add_graph_coo(g.m, rows, g.col, symmetrize=True)
        assert e.value.status == -4
        f = T.fixture("test_bc")
        frows = np.repeat(np.arange(f.m, dtype=np.int32), np.diff(f.rowptr))
        gi = srv.add_graph_coo(f.m, frows, f.col, symmetrize=False)
        res, _ = srv.wait(srv.submit(gi, A.Q_BFS, 0), A.Q_BFS, f.m)
        assert np.array_equal(res, O.bfs(f.m, f.rowptr, f.col, 0)[0])
