"""Writes tests/golden/golden_full.json: known answers at the workload sizes
(R-MAT scale 20 / 22, lattice 2048^2), computed by the oracle, which is pinned
at smaller sizes against independent sources (tests/test_oracle.py).  The
GPU tests use these so the box does not rerun the slow CPU parts.

Run:  python tests/golden/make_golden_full.py   (several minutes, ~10 GB RAM)
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
from oracle import oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sources(rowptr, k, seed=3):
    """k distinct vertices with degree > 0: v = draw % m over mt19937(seed)."""
    m = len(rowptr) - 1
    deg = np.diff(rowptr)
    out, seen = [], set()
    for d in O.mt_stream(seed, 4096 * k).tolist():
        v = d % m
        if deg[v] > 0 and v not in seen:
            seen.add(v)
            out.append(v)
            if len(out) == k:
                return out
    raise RuntimeError("not enough sources")


def main():
    out = {}
    t = time.time()
    m, rp, c = O.rmat_graph(20)
    total, rows = O.tc(m, rp, c)
    mis = O.mis_greedy(m, rp, c, O.mis_priorities(m))
    out["rmat20"] = dict(n=m, nnz=int(len(c)), rowptr_sha=sha(rp.astype(np.int32)),
                         col_sha=sha(c), triangles=int(total), tc_rows_sha=sha(rows),
                         mis_sha=sha(mis), mis_size=int((mis == 1).sum()))
    print("rmat20", time.time() - t, flush=True)

    t = time.time()
    m, rp, c = O.rmat_graph(22)
    src = sources(rp, 64)
    w = O.assign_weights(m, rp, c)
    rec = dict(n=m, nnz=int(len(c)), rowptr_sha=sha(rp.astype(np.int32)), col_sha=sha(c),
               weights_sha=sha(w), sources=src, bfs_sha={}, sssp_sha={})
    for s in src[:4]:
        rec["bfs_sha"][str(s)] = sha(O.bfs(m, rp, c, s)[0])
    for s in src[:2]:
        rec["sssp_sha"][str(s)] = sha(O.sssp(m, rp, c, w, s))
    out["rmat22"] = rec
    print("rmat22", time.time() - t, flush=True)
    del rp, c, w

    t = time.time()
    m, rp, c = O.lattice_graph(2048)
    w = O.assign_weights(m, rp, c)
    lab = O.components(m, rp, c)
    big = np.bincount(lab).argmax()
    comp = np.nonzero(lab == big)[0]
    pick = O.mt_stream(4, 64) % len(comp)
    src = [int(comp[i]) for i in pick[:4]]
    rec = dict(n=m, nnz=int(len(c)), rowptr_sha=sha(rp.astype(np.int32)), col_sha=sha(c),
               weights_sha=sha(w), largest_component=int(len(comp)), sources=src,
               bfs_sha={}, sssp_sha={})
    for s in src[:2]:
        rec["bfs_sha"][str(s)] = sha(O.bfs(m, rp, c, s)[0])
        rec["sssp_sha"][str(s)] = sha(O.sssp(m, rp, c, w, s))
    out["lattice2048"] = rec
    print("lattice2048", time.time() - t, flush=True)

    (HERE / "golden_full.json").write_text(json.dumps(out, indent=1, sort_keys=True))
    print("wrote golden_full.json")


if __name__ == "__main__":
    main()
