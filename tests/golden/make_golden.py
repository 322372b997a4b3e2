"""Writes tests/golden/golden.json: known answers for the oracle and the HIP path.

Every expected value here is computed WITHOUT the oracle under test, from
independent sources, so that agreement pins the oracle:

* MT19937 draws: numpy's legacy MT19937 (init_genrand seeding, genrand_res53
  doubles) -- a separate implementation of the generator that the reference's
  src/mt19937ar.cpp implements.
* Graph construction: numpy restatements of the synthetic-input definitions
  (R-MAT quadrant walk, lattice with seeded removals and long-range edges,
  symmetric integer weights), independent of oracle/oracle.c.
* BFS depths and SSSP distances: scipy.sparse.csgraph (breadth-first
  shortest_path, Dijkstra in float64 -- exact for integer weights).
* Triangles: sparse (L @ L) .* L row sums in int64 (scipy.sparse), and
  networkx.triangles for the fixtures.
* MIS: the greedy-by-(priority desc, id asc) closed form, in plain Python.
* Components: scipy.sparse.csgraph.connected_components.

The MatrixMarket inputs under tests/golden/mtx are the reference's own
fixtures (dataset/small/*.mtx), copied as data.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import networkx as nx
import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

HERE = Path(__file__).resolve().parent


# ------------------------------------------------------------------ helpers
def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def legacy_mt(seed: int):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    return bg


def mt_u32(seed: int, n: int) -> np.ndarray:
    return legacy_mt(seed).random_raw(n).astype(np.uint32)


def mt_res53(seed: int, n: int) -> np.ndarray:
    return np.random.RandomState(seed).random_sample(n)


def read_mtx(path: Path):
    """MatrixMarket coordinate reader: first index -> col, second -> row
    (the reference's readMtx column swap), 1-based -> 0-based."""
    lines = [ln for ln in path.read_text().splitlines() if ln.strip()]
    body = [ln for ln in lines if not ln.startswith("%")]
    m, n, z = (int(t) for t in body[0].split()[:3])
    rows, cols = [], []
    for ln in body[1:1 + z]:
        t = ln.split()
        cols.append(int(t[0]) - 1)
        rows.append(int(t[1]) - 1)
    return max(m, n), np.array(rows, np.int64), np.array(cols, np.int64)


def sym_csr(n: int, u: np.ndarray, v: np.ndarray) -> sp.csr_matrix:
    """symmetrize + dedup + drop self-loops, as a sorted CSR pattern"""
    keep = u != v
    u, v = u[keep], v[keep]
    r = np.concatenate([u, v])
    c = np.concatenate([v, u])
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = key // n, key % n
    A = sp.csr_matrix((np.ones(len(r), np.float64), (r, c)), shape=(n, n))
    A.sort_indices()
    return A


def rmat_edges(scale: int, ef: int = 16, a=0.57, b=0.19, c=0.19, seed=1):
    ne = ef << scale
    r = mt_res53(seed, ne * scale).reshape(ne, scale)
    ub = (r >= a + b).astype(np.int64)                   # quadrants (1,0),(1,1)
    vb = (((r >= a) & (r < a + b)) | (r >= a + b + c)).astype(np.int64)
    w = 1 << np.arange(scale - 1, -1, -1, dtype=np.int64)  # first draw = MSB
    return (ub * w).sum(1), (vb * w).sum(1)


def lattice_edges(W: int, seed: int = 1):
    n = W * W
    eu, ev = [], []
    for i in range(n):
        r, c = divmod(i, W)
        if c < W - 1:
            eu.append(i); ev.append(i + 1)
        if r < W - 1:
            eu.append(i); ev.append(i + W)
    E = len(eu)
    bg = legacy_mt(seed)
    draws = iter(bg.random_raw(4 * E + 100000).astype(np.uint64).tolist())
    perm = list(range(E))
    for i in range(E - 1, 0, -1):
        j = next(draws) % (i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    gone = set(perm[: E // 10])
    out_u = [eu[i] for i in range(E) if i not in gone]
    out_v = [ev[i] for i in range(E) if i not in gone]
    have = set(zip(out_u, out_v))
    added = 0
    while added < E // 1000:
        a = next(draws) % n
        b = next(draws) % n
        if a == b:
            continue
        p = (min(a, b), max(a, b))
        if p in have:
            continue
        have.add(p)
        out_u.append(p[0]); out_v.append(p[1])
        added += 1
    return np.array(out_u, np.int64), np.array(out_v, np.int64)


def weights(A: sp.csr_matrix, seed: int = 2) -> np.ndarray:
    """1 + (mt19937(seed) draw % 255) per pair u<v in CSR order, mirrored."""
    n = A.shape[0]
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    cols = A.indices.astype(np.int64)
    up = cols > rows
    w = np.zeros(len(cols), np.float32)
    d = mt_u32(seed, int(up.sum())).astype(np.int64)
    w[up] = (1 + d % 255).astype(np.float32)
    keys = rows * n + cols
    mirror = np.searchsorted(keys, cols[up] * n + rows[up])
    w[mirror] = w[up]
    return w


def bfs_depth(A: sp.csr_matrix, src: int) -> np.ndarray:
    d = csg.shortest_path(A, method="D", unweighted=True, indices=src)
    out = np.where(np.isinf(d), -1, d).astype(np.int32)
    return out


def sssp_dist(A: sp.csr_matrix, w: np.ndarray, src: int) -> np.ndarray:
    W = sp.csr_matrix((w.astype(np.float64), A.indices, A.indptr), shape=A.shape)
    d = csg.dijkstra(W, indices=src)
    return np.where(np.isinf(d), -1.0, d).astype(np.float32)


def tc_rows(A: sp.csr_matrix) -> np.ndarray:
    L = sp.tril(A, -1, format="csr").astype(np.int64)
    return np.asarray((L @ L).multiply(L).sum(axis=1)).ravel().astype(np.int32)


def mis_greedy(A: sp.csr_matrix, prio: np.ndarray) -> np.ndarray:
    n = A.shape[0]
    order = np.lexsort((np.arange(n), -prio.astype(np.int64)))
    st = np.full(n, -1, np.int32)
    ip, ix = A.indptr, A.indices
    for v in order.tolist():
        if st[v] != -1:
            continue
        st[v] = 1
        st[ix[ip[v]:ip[v + 1]]] = 0
        st[v] = 1
    return st


def summary(A: sp.csr_matrix, src: int, full: bool, w=None) -> dict:
    n = A.shape[0]
    depth = bfs_depth(A, src)
    prio = mt_u32(5489, n)
    mis = mis_greedy(A, prio)
    rows = tc_rows(A)
    ncomp, lab = csg.connected_components(A, directed=False)
    rec = {
        "n": int(n),
        "nnz": int(A.nnz),
        "rowptr_sha": sha(A.indptr.astype(np.int32)),
        "col_sha": sha(A.indices.astype(np.int32)),
        "components": int(ncomp),
        "source": int(src),
        "bfs_sha": sha(depth),
        "bfs_max_depth": int(depth.max()),
        "bfs_reached": int((depth >= 0).sum()),
        "triangles": int(rows.astype(np.int64).sum()),
        "tc_rows_sha": sha(rows),
        "mis_sha": sha(mis),
        "mis_size": int((mis == 1).sum()),
    }
    if w is not None:
        dist = sssp_dist(A, w, src)
        rec["weights_sha"] = sha(w)
        rec["sssp_sha"] = sha(dist)
        rec["sssp_max"] = float(dist.max())
    if full:
        rec.update(bfs=depth.tolist(), tc_rows=rows.tolist(), mis=mis.tolist(),
                   rowptr=A.indptr.tolist(), col=A.indices.tolist())
        if w is not None:
            rec["sssp"] = sssp_dist(A, w, src).tolist()
    return rec


def main():
    out = {"mt19937": {
        "seed5489_first": int(mt_u32(5489, 1)[0]),
        "seed5489_10000th": int(mt_u32(5489, 10000)[-1]),
        "seed1_first8": mt_u32(1, 8).tolist(),
        "seed2_first8": mt_u32(2, 8).tolist(),
        "res53_seed1_first4": mt_res53(1, 4).tolist(),
    }}

    fx = {}
    for p in sorted((HERE / "mtx").glob("*.mtx")):
        n, r, c = read_mtx(p)
        A = sym_csr(n, r, c)
        G = nx.Graph()
        G.add_nodes_from(range(n))
        rr = np.repeat(np.arange(n), np.diff(A.indptr))
        G.add_edges_from(zip(rr.tolist(), A.indices.tolist()))
        rec = summary(A, 0, True, w=weights(A))
        rec["nx_triangles"] = int(sum(nx.triangles(G).values()) // 3)
        fx[p.stem] = rec
    out["fixtures"] = fx

    # R-MAT scale 10: full edge-list hashes (generator known answer)
    u, v = rmat_edges(10)
    A10 = sym_csr(1 << 10, u, v)
    out["rmat10"] = dict(edges_u_sha=sha(u.astype(np.int32)),
                         edges_v_sha=sha(v.astype(np.int32)),
                         first_edges=[[int(a), int(b)] for a, b in zip(u[:8], v[:8])],
                         **summary(A10, int(np.argmax(np.diff(A10.indptr))), False,
                                   w=weights(A10)))

    # R-MAT scale 14 (the CPU-path BFS workload): highest-degree source
    u, v = rmat_edges(14)
    A14 = sym_csr(1 << 14, u, v)
    deg = np.diff(A14.indptr)
    out["rmat14"] = summary(A14, int(np.argmax(deg)), False, w=weights(A14))

    # lattices 32x32 and 64x64
    for W in (32, 64):
        eu, ev = lattice_edges(W)
        A = sym_csr(W * W, eu, ev)
        rec = summary(A, 0, False, w=weights(A))
        rec.update(edges_u_sha=sha(eu.astype(np.int32)), edges_v_sha=sha(ev.astype(np.int32)),
                   n_edges=int(len(eu)))
        out[f"lattice{W}"] = rec

    (HERE / "golden.json").write_text(json.dumps(out, indent=1, sort_keys=True))
    print("wrote", HERE / "golden.json")


if __name__ == "__main__":
    main()
