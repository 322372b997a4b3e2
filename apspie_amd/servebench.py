"""Served query stream over a resident graph pool (the serving workload):
Poisson arrivals of mixed queries (60% BFS, 30% SSSP, 10% TC) with a
uniformly random graph and source, submitted to the continuous-batching
server; reports queries/s and p50/p99 latency (submit -> result in pinned host
memory, handed out without a copy).

The same stream is replayed against the plain form (one worker, one query per
launch, no coalescing) for comparison.
"""
from __future__ import annotations

import threading
import time

import numpy as np

from . import Q_BFS, Q_SSSP, Q_TC, gen_lattice, gen_rmat, gen_weights
from .server import Server

POOL = [("rmat", 16), ("rmat", 18), ("rmat", 20), ("rmat", 22),
        ("lattice", 512), ("lattice", 1024), ("lattice", 2048), ("lattice", 4096)]


def build_pool(pool=POOL, log=None):
    graphs = []
    for kind, p in pool:
        t0 = time.time()
        rp, col = gen_rmat(p) if kind == "rmat" else gen_lattice(p)
        w = gen_weights(rp, col)
        graphs.append((f"{kind}{p}", rp, col, w))
        if log:
            log(f"  {kind}{p}: m={len(rp) - 1} nnz={len(col)} ({time.time() - t0:.1f}s)")
    return graphs


def make_stream(graphs, n, seed=7):
    """n queries: (graph index, kind, source), 60/30/10 BFS/SSSP/TC."""
    rng = np.random.default_rng(seed)
    out = []
    nz = [np.nonzero(np.diff(g[1]))[0] for g in graphs]
    for _ in range(n):
        gi = int(rng.integers(len(graphs)))
        r = rng.random()
        kind = Q_BFS if r < 0.6 else (Q_SSSP if r < 0.9 else Q_TC)
        src = int(nz[gi][rng.integers(len(nz[gi]))])
        out.append((gi, kind, src))
    return out


def run_stream(server, ids, graphs, stream, rate, seed=11):
    """Submit `stream` with exponential inter-arrival times at `rate` q/s
    from a separate thread; wait for all; returns qps and latency stats."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(stream)) if rate > 0 else np.zeros(len(stream))
    tickets = [None] * len(stream)
    t0 = time.perf_counter()

    def submitter():
        t = t0
        for i, (gi, kind, src) in enumerate(stream):
            t += gaps[i]
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = server.submit(ids[gi], kind, src)

    th = threading.Thread(target=submitter)
    th.start()
    lat = []
    for i, (gi, kind, src) in enumerate(stream):
        while tickets[i] is None:
            time.sleep(0.0002)
        # results as views of the server's pinned buffers (no host copy)
        res, ms = server.wait(tickets[i], kind, len(graphs[gi][1]) - 1, copy=False)
        del res
        lat.append(ms)
    th.join()
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    return {"offered_qps": rate, "queries": len(stream), "qps": len(stream) / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "wall_s": wall}


def serve_benchmark(graphs, n_queries=300, rates=None, slots=4, max_batch=128,
                    hbm_budget=64 << 30, log=None):
    """Plain capacity first (the whole stream, one query per launch), then the
    served form on the same stream at three offered rates (x1, x4, x16 the
    plain capacity) and saturated."""
    res = {}
    stream = make_stream(graphs, n_queries)
    with Server(slots=1, max_batch=1, hbm_budget=hbm_budget) as plain:
        ids = [plain.add_graph(g[1], g[2], g[3]) for g in graphs]
        r = run_stream(plain, ids, graphs, stream, rate=0)
        res["plain_saturated"] = r
        if log:
            log(f"  plain: {r['qps']:.1f} q/s p50 {r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms")
    cap = res["plain_saturated"]["qps"]
    rates = rates or [cap, 4 * cap, 16 * cap]
    with Server(slots=slots, max_batch=max_batch, hbm_budget=hbm_budget) as srv:
        ids = [srv.add_graph(g[1], g[2], g[3]) for g in graphs]
        res["served"] = []
        for rate in rates:
            r = run_stream(srv, ids, graphs, stream, rate=rate)
            res["served"].append(r)
            if log:
                log(f"  served @ {rate:.1f} q/s offered: {r['qps']:.1f} q/s p50 {r['p50_ms']:.1f} "
                    f"ms p99 {r['p99_ms']:.1f} ms")
        r = run_stream(srv, ids, graphs, stream, rate=0)
        res["served_saturated"] = r
        if log:
            log(f"  served saturated: {r['qps']:.1f} q/s p50 {r['p50_ms']:.1f} ms "
                f"p99 {r['p99_ms']:.1f} ms")
    res["speedup_saturated"] = res["served_saturated"]["qps"] / cap
    res["config"] = {"slots": slots, "max_batch": max_batch, "hbm_budget": hbm_budget,
                     "mix": "60% BFS / 30% SSSP / 10% TC", "pool": [g[0] for g in graphs]}
    return res
