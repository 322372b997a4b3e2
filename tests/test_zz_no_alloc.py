"""apspie_bfs_noalloc (the BFS entry point that allocates nothing) against
apspie_bfs and the oracle, through the C ABI: two calls in a row with the
second on a smaller graph, host and device outputs, pushed levels (symmetry
not yet known) and pulled ones (known), and the size limit."""
import ctypes as C

import numpy as np
import pytest

import apspie_amd as A
from oracle import oracle as O
from tests import graphs as T

pytestmark = pytest.mark.gpu

L = A.lib()


@pytest.fixture(scope="module")
def ctx():
    with A.Context(0) as c:
        yield c


def noalloc_host(ctx, G, src):
    out = np.full(G.m, -7, np.int32)
    st = L.apspie_bfs_noalloc(ctx.h, G.h, int(src), out.ctypes.data, A.HOST)
    A._check(st, ctx.h)
    return out


def noalloc_dev(ctx, G, src):
    buf = ctx.dev_alloc(4 * G.m)
    try:
        A._check(L.apspie_bfs_noalloc(ctx.h, G.h, int(src), buf.ptr, A.DEVICE), ctx.h)
        return buf.to_host(np.int32, G.m)
    finally:
        buf.free()


def test_two_calls_second_smaller(ctx):
    """Larger graph first, then a smaller one (the array keeps the first
    call's contents): both equal apspie_bfs and the oracle, host and device
    outputs, before and after the graphs' symmetry is known."""
    big, small = T.rmat(14), T.rmat(10)
    Gb = ctx.upload(big.rowptr32, big.col)
    Gs = ctx.upload(small.rowptr32, small.col)
    try:
        for g, G in ((big, Gb), (small, Gs)):  # symmetry unknown: every level pushes
            for s in T.sources(g, 2):
                want, _ = O.bfs(g.m, g.rowptr, g.col, s)
                assert np.array_equal(noalloc_host(ctx, G, s), want), (g, s)
                assert np.array_equal(noalloc_dev(ctx, G, s), want), (g, s)
        for g, G in ((big, Gb), (small, Gs)):  # apspie_bfs checks symmetry: pulls from now on
            for s in T.sources(g, 3, seed=9):
                ref = ctx.bfs(G, s)
                st_ref = ctx.stats()
                got = noalloc_host(ctx, G, s)
                st = ctx.stats()
                assert np.array_equal(got, ref), (g, s)
                assert np.array_equal(noalloc_dev(ctx, G, s), ref), (g, s)
                assert st["edges_traversed"] == st_ref["edges_traversed"], (g, s)
                assert st["iterations"] == st_ref["iterations"], (g, s)
    finally:
        Gb.close()
        Gs.close()


def test_symmetrized_graph_and_lattice(ctx):
    """A graph built symmetric on the device pulls from the first call; a
    lattice (long paths) and a graph with isolated vertices."""
    for g in (T.lattice(64), T.random_graph(3000, 0.001, 5)):
        rows = np.repeat(np.arange(g.m, dtype=np.int32), np.diff(g.rowptr32))
        G = ctx.from_coo(g.m, rows, g.col, symmetrize=True)
        try:
            for s in T.sources(g, 3):
                want, _ = O.bfs(g.m, g.rowptr, g.col, s)
                assert np.array_equal(noalloc_host(ctx, G, s), want), (g, s)
                assert np.array_equal(noalloc_dev(ctx, G, s), ctx.bfs(G, s)), (g, s)
        finally:
            G.close()


def test_too_large_is_a_clear_error(ctx):
    """2^23 vertices need more than the static array: APSPIE_ERR_OVERFLOW with
    the sizes in the message, checked before anything runs; the library stays
    usable."""
    m = 1 << 23
    rp = ctx.dev_alloc(4 * (m + 1))
    col = ctx.dev_alloc(4)
    out = ctx.dev_alloc(4 * m)
    try:
        rp.from_host(np.zeros(m + 1, np.int32))
        csr = A._Csr(m, m, 0, rp.ptr, col.ptr, None)
        h = C.c_void_p()
        A._check(L.apspie_graph_wrap(ctx.h, C.byref(csr), C.byref(h)), ctx.h)
        G = A.Graph(ctx, h)
        st = L.apspie_bfs_noalloc(ctx.h, G.h, 0, out.ptr, A.DEVICE)
        assert st == -5
        assert "too large for the no-alloc BFS" in L.apspie_last_error(ctx.h).decode()
        G.close()
    finally:
        rp.free()
        col.free()
        out.free()
    g = T.rmat(8)
    G = ctx.upload(g.rowptr32, g.col)
    s = T.sources(g, 1)[0]
    assert np.array_equal(noalloc_host(ctx, G, s), O.bfs(g.m, g.rowptr, g.col, s)[0])
    G.close()


def test_rejects_bad_arguments(ctx):
    g = T.rmat(8)
    G = ctx.upload(g.rowptr32, g.col)
    out = np.empty(g.m, np.int32)
    assert L.apspie_bfs_noalloc(ctx.h, G.h, g.m, out.ctypes.data, A.HOST) == -1
    assert L.apspie_bfs_noalloc(ctx.h, G.h, 0, None, A.HOST) == -1
    assert L.apspie_bfs_noalloc(ctx.h, G.h, 0, out.ctypes.data, 5) == -1
    G.close()
