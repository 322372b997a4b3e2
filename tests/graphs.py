"""Test inputs: the reference's MatrixMarket fixtures and the synthetic graphs,
built by the oracle (checker side) as int64/int32 CSR plus int32 copies for
upload."""
from __future__ import annotations

import json
from functools import lru_cache
from pathlib import Path

import numpy as np

from oracle import oracle as O

GOLDEN_DIR = Path(__file__).resolve().parent / "golden"
GOLD = json.loads((GOLDEN_DIR / "golden.json").read_text())
FIXTURES = sorted(GOLD["fixtures"])


class G:
    """A symmetric graph: oracle view (int64 rowptr) and upload view (int32)."""

    def __init__(self, name, m, rowptr, col, weights=True):
        self.name = name
        self.m = m
        self.rowptr = np.ascontiguousarray(rowptr, np.int64)
        self.col = np.ascontiguousarray(col, np.int32)
        self.rowptr32 = self.rowptr.astype(np.int32)
        self.w = O.assign_weights(m, self.rowptr, self.col) if weights else None

    @property
    def deg(self):
        return np.diff(self.rowptr)

    def __repr__(self):
        return f"G({self.name}, m={self.m}, nnz={len(self.col)})"


@lru_cache(maxsize=None)
def fixture(name) -> G:
    m, rp, c, _ = O.mtx_graph((GOLDEN_DIR / "mtx" / f"{name}.mtx").read_text())
    return G(name, m, rp, c)


@lru_cache(maxsize=None)
def rmat(scale, ef=16) -> G:
    m, rp, c = O.rmat_graph(scale, ef)
    return G(f"rmat{scale}", m, rp, c)


@lru_cache(maxsize=None)
def lattice(W) -> G:
    m, rp, c = O.lattice_graph(W)
    return G(f"lattice{W}", m, rp, c)


def path_graph(n) -> G:
    rp, c = O.edges_to_csr(n, np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32))
    return G(f"path{n}", n, rp, c)


def star_graph(n) -> G:
    rp, c = O.edges_to_csr(n, np.zeros(n - 1, np.int32), np.arange(1, n, dtype=np.int32))
    return G(f"star{n}", n, rp, c)


def empty_graph(n) -> G:
    return G(f"empty{n}", n, np.zeros(n + 1, np.int64), np.zeros(0, np.int32))


def random_graph(n, p, seed) -> G:
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, int(p * n * n / 2)).astype(np.int32)
    v = rng.integers(0, n, len(u)).astype(np.int32)
    rp, c = O.edges_to_csr(n, u, v)
    return G(f"rand{n}_{seed}", n, rp, c)


def sources(g: G, k, seed=3):
    """k distinct vertices with degree > 0 from mt19937(seed) draws mod m."""
    draws = O.mt_stream(seed, 64 * k + 1000)
    out, seen = [], set()
    deg = g.deg
    for d in draws.tolist():
        v = d % g.m
        if deg[v] > 0 and v not in seen:
            seen.add(v)
            out.append(v)
            if len(out) == k:
                break
    return out
