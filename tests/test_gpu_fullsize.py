"""Parity at the workload sizes: R-MAT scale 22 (BFS, SSSP), R-MAT scale 20
(MIS, TC) and the 2048^2 lattice (plain and batched traversals).

Known answers come from tests/golden/golden_full.json (oracle outputs, the
oracle being pinned at smaller sizes); the rest is checked through
size-independent properties: every edge spans at most one BFS level and
every reached vertex has a parent one level up; SSSP distances satisfy the
triangle inequality on every edge and are tight on some in-edge; batched
results equal the single-source results; per-row triangle counts add up to
the total."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import apspie_amd as A

pytestmark = pytest.mark.gpu

GF = json.loads((Path(__file__).parent / "golden" / "golden_full.json").read_text())


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ctx():
    with A.Context(0) as c:
        yield c


def build(kind):
    if kind == "rmat22":
        rp, col = A.gen_rmat(22)
    elif kind == "rmat20":
        rp, col = A.gen_rmat(20)
    else:
        rp, col = A.gen_lattice(2048)
    g = GF[kind]
    assert sha(rp) == g["rowptr_sha"] and sha(col) == g["col_sha"], kind
    return rp, col


@pytest.fixture(scope="module")
def rmat22(ctx):
    rp, col = build("rmat22")
    w = A.gen_weights(rp, col)
    assert sha(w) == GF["rmat22"]["weights_sha"]
    G = ctx.upload(rp, col, w)
    yield rp, col, w, G
    G.close()


def bfs_properties(rp, col, d, src):
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.int32), np.diff(rp))
    du, dv = d[rows], d[col]
    assert d[src] == 0
    both = (du >= 0) & (dv >= 0)
    assert np.all(np.abs(du[both] - dv[both]) <= 1)
    assert not np.any((du >= 0) ^ (dv >= 0))  # reached set is closed
    parent = np.zeros(len(d), bool)
    parent[col[(du >= 0) & (dv == du + 1)]] = True
    assert np.all(parent[(d > 0)])


def sssp_properties(rp, col, w, d, src):
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.int32), np.diff(rp))
    inf = A.SSSP_INF
    du, dv = d[rows], d[col]
    reach = du < inf
    assert d[src] == 0
    assert np.all(dv[reach] <= du[reach] + w[reach])
    tight = np.zeros(len(d), bool)
    tight[col[reach & (dv == du + w)]] = True
    assert np.all(tight[(d > 0) & (d < inf)])


def test_rmat22_bfs(ctx, rmat22):
    rp, col, w, G = rmat22
    g = GF["rmat22"]
    src = g["sources"]
    assert len(src) == 64
    for s in src:
        d = ctx.bfs(G, s)
        if str(s) in g["bfs_sha"]:
            assert sha(d) == g["bfs_sha"][str(s)], s
        st = ctx.stats()
        assert st["edges_traversed"] > 0
    for s in src[4:8]:
        bfs_properties(rp, col, ctx.bfs(G, s), s)


def test_rmat22_bfs_batch_equals_plain(ctx, rmat22):
    rp, col, w, G = rmat22
    src = GF["rmat22"]["sources"][:32]
    got = ctx.bfs_batch(G, src)
    for i in (0, 1, 17, 31):
        assert np.array_equal(got[i], ctx.bfs(G, src[i]))


def test_rmat22_sssp(ctx, rmat22):
    rp, col, w, G = rmat22
    g = GF["rmat22"]
    for s_str, h in g["sssp_sha"].items():
        d = ctx.sssp(G, int(s_str))
        want = d.copy()
        want[want == A.SSSP_INF] = -1  # oracle convention for unreached
        assert sha(want) == h
    s = g["sources"][5]
    sssp_properties(rp, col, w, ctx.sssp(G, s), s)


def test_rmat20_mis_tc(ctx):
    rp, col = build("rmat20")
    g = GF["rmat20"]
    G = ctx.upload(rp, col)
    s = ctx.mis(G)
    assert sha(s) == g["mis_sha"] and int((s == 1).sum()) == g["mis_size"]
    total, rows = ctx.tc(G)
    assert total == g["triangles"]
    assert sha(rows) == g["tc_rows_sha"]
    assert int(rows.astype(np.int64).sum()) == total
    G.close()


def test_lattice2048_plain_and_batched(ctx):
    rp, col = build("lattice2048")
    w = A.gen_weights(rp, col)
    g = GF["lattice2048"]
    assert sha(w) == g["weights_sha"]
    G = ctx.upload(rp, col, w)
    for s_str, h in g["bfs_sha"].items():
        assert sha(ctx.bfs(G, int(s_str))) == h
    for s_str, h in g["sssp_sha"].items():
        d = ctx.sssp(G, int(s_str))
        d[d == A.SSSP_INF] = -1
        assert sha(d) == h
    src = g["sources"]
    got = ctx.bfs_batch(G, src)
    for i, s in enumerate(src):
        assert np.array_equal(got[i], ctx.bfs(G, s))
    gs = ctx.sssp_batch(G, src[:2])
    for i, s in enumerate(src[:2]):
        assert np.array_equal(gs[i], ctx.sssp(G, s))
    G.close()
