"""Served query stream over a resident graph pool (the serving workload):
Poisson arrivals of mixed queries (60% BFS, 30% SSSP, 10% TC) with a
uniformly random graph and source, submitted to the continuous-batching
server; reports queries/s and p50/p99 latency (submit -> result in pinned host
memory, handed out without a copy).

The same stream is replayed against two plain forms (one query per launch,
no coalescing): with one worker (the drop-in loop) and with as many workers
as the served form, so that the gain from concurrency and the gain from
batching show separately.  Each form's busy time is broken down per (graph,
kind).  An untimed pass afterwards checks the served results against the
plain ones query by query (xxh3 digests of the result bytes).
"""
from __future__ import annotations

import threading
import time

import numpy as np
import xxhash

from . import DEVICE, Q_BFS, Q_SSSP, Q_TC, Context, _check, gen_lattice, gen_rmat, gen_weights, lib
from .server import Server

POOL = [("rmat", 16), ("rmat", 18), ("rmat", 20), ("rmat", 22),
        ("lattice", 512), ("lattice", 1024), ("lattice", 2048), ("lattice", 4096)]
KIND_NAMES = {Q_BFS: "bfs", Q_SSSP: "sssp", Q_TC: "tc"}


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


def digest(res):
    return int(res) if isinstance(res, int) else xxhash.xxh3_64_intdigest(res)


class ResultArena:
    """One device allocation holding every query's result row of a stream
    (4 m bytes per BFS / SSSP query): the caller-owned result arrays the
    reference's entry points write (d_bfsResult, bfs.cuh:32-33)."""

    def __init__(self, ctx, graphs, stream):
        self.ctx = ctx
        sizes = [0 if kind == Q_TC else 4 * (len(graphs[gi][1]) - 1) for gi, kind, _ in stream]
        self.off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.buf = ctx.dev_alloc(max(int(self.off[-1]), 256))

    def ptr(self, i):
        return self.buf.ptr + int(self.off[i])

    def row(self, i, dtype):
        """Query i's result, copied to the host (outside any timing)."""
        n = int(self.off[i + 1] - self.off[i]) // 4
        out = np.empty(n, dtype)
        _check(lib().apspie_memcpy(self.ctx.h, out.ctypes.data, self.ptr(i), out.nbytes, 1),
               self.ctx.h)
        return out

    def free(self):
        self.buf.free()


def run_stream(server, ids, graphs, stream, rate, seed=11, digests=False, arena=None):
    """Submit `stream` with exponential inter-arrival times at `rate` q/s
    (0: all at once) from a separate thread; wait for all; returns qps and
    latency stats (and the results' digests when asked, outside timing).
    With an arena every BFS / SSSP result goes to the query's row of it in
    device memory; without, to the server's pinned host buffers."""
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
            if arena is not None:
                tickets[i] = server.submit_into(ids[gi], kind, src, arena.ptr(i), DEVICE)
            else:
                tickets[i] = server.submit(ids[gi], kind, src)

    th = threading.Thread(target=submitter)
    th.start()
    lat = []
    dig = []
    tc_counts = {}
    for i, (gi, kind, src) in enumerate(stream):
        while tickets[i] is None:
            time.sleep(0.0002)
        if arena is not None and kind != Q_TC:
            ms = server.wait_done(tickets[i])
        else:
            # results as views of the server's pinned buffers (no host copy)
            res, ms = server.wait(tickets[i], kind, len(graphs[gi][1]) - 1, copy=False)
            if kind == Q_TC:
                tc_counts[i] = res
            elif digests:
                dig.append((i, digest(res)))
            del res
        lat.append(ms)
    th.join()
    wall = time.perf_counter() - t0
    if digests:
        if arena is not None:
            dig = [(i, digest(arena.row(i, np.int32 if kind == Q_BFS else np.float32)))
                   for i, (gi, kind, src) in enumerate(stream) if kind != Q_TC]
        dig = dict(dig)
        dig.update({i: digest(c) for i, c in tc_counts.items()})
        dig = [dig[i] for i in range(len(stream))]
    lat = np.array(lat)
    out = {"offered_qps": rate, "queries": len(stream), "qps": len(stream) / wall,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "wall_s": wall, "results": "device" if arena is not None else "host"}
    return (out, dig) if digests else out


def busy_breakdown(server, ids, graphs):
    """Busy time per (graph, kind) since the last reset, largest first, with
    its phases (traversal / count calls, result copies, pinned buffers
    obtained, attempts that ran out of budget and the trims after them)."""
    rows = []
    for gi, g in enumerate(graphs):
        for kind in (Q_BFS, Q_SSSP, Q_TC):
            st = server.busy_stats(ids[gi], kind, reset=True)
            q = st["queries"]
            if q:
                rows.append({"graph": g[0], "kind": KIND_NAMES[kind], "batches": st["batches"],
                             "queries": q, "busy_ms": st["busy_ms"],
                             "ms_per_query": st["busy_ms"] / q,
                             "traverse_ms": st["traverse_ms"], "copy_ms": st["copy_ms"],
                             "alloc_ms": st["alloc_ms"], "retries": st["retries"],
                             "retry_ms": st["retry_ms"]})
    total = sum(r["busy_ms"] for r in rows) or 1.0
    for r in rows:
        r["share"] = r["busy_ms"] / total
    return sorted(rows, key=lambda r: -r["busy_ms"])


def serve_benchmark(graphs, n_queries=2000, rates=None, slots=4, max_batch=128,
                    hbm_budget=64 << 30, n_verify=400, log=None):
    """Plain capacity first (one worker, one query per launch), then the plain
    form with `slots` workers, then the served form on the same stream at
    three offered rates (x1, x4, x16 the one-worker plain capacity) and
    saturated.  Every form writes each result into a device array the caller
    owns (the reference's result contract); one saturated run of each form
    with the results in host memory gives the PCIe-inclusive rates.  Finally
    an untimed check of served against plain results."""
    res = {}
    stream = make_stream(graphs, n_queries)
    actx = Context(0)  # owns the result arena
    arena = ResultArena(actx, graphs, stream)

    def say(tag, r):
        if log:
            log(f"  {tag}: {r['qps']:.1f} q/s p50 {r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms "
                f"(results in {r['results']} memory)")

    with Server(slots=1, max_batch=1, hbm_budget=hbm_budget) as plain:
        ids = [plain.add_graph(g[1], g[2], g[3]) for g in graphs]
        r = run_stream(plain, ids, graphs, stream, rate=0, arena=arena)
        res["plain_saturated"] = r
        res["plain_busy"] = busy_breakdown(plain, ids, graphs)
        say("plain, 1 worker", r)
        r = run_stream(plain, ids, graphs, stream, rate=0)
        res["plain_saturated_host"] = r
        busy_breakdown(plain, ids, graphs)
        say("plain, 1 worker", r)
    with Server(slots=slots, max_batch=1, hbm_budget=hbm_budget) as plainn:
        ids = [plainn.add_graph(g[1], g[2], g[3]) for g in graphs]
        r = run_stream(plainn, ids, graphs, stream, rate=0, arena=arena)
        res["plain_slots_saturated"] = r
        busy_breakdown(plainn, ids, graphs)
        say(f"plain, {slots} workers", r)
    cap = res["plain_saturated"]["qps"]
    rates = rates or [cap, 4 * cap, 16 * cap]
    with Server(slots=slots, max_batch=max_batch, hbm_budget=hbm_budget) as srv:
        ids = [srv.add_graph(g[1], g[2], g[3]) for g in graphs]
        res["served"] = []
        for rate in rates:
            r = run_stream(srv, ids, graphs, stream, rate=rate, arena=arena)
            res["served"].append(r)
            say(f"served @ {rate:.1f} q/s offered", r)
        busy_breakdown(srv, ids, graphs)
        r = run_stream(srv, ids, graphs, stream, rate=0, arena=arena)
        res["served_saturated"] = r
        res["served_busy"] = busy_breakdown(srv, ids, graphs)
        say("served saturated", r)
        r = run_stream(srv, ids, graphs, stream, rate=0)
        res["served_saturated_host"] = r
        busy_breakdown(srv, ids, graphs)
        say("served saturated", r)
        # untimed: the served results of the stream's first n_verify queries
        part = stream[:n_verify]
        _, served_dig = run_stream(srv, ids, graphs, part, rate=0, digests=True, arena=arena)
    with Server(slots=1, max_batch=1, hbm_budget=hbm_budget) as plain:
        ids = [plain.add_graph(g[1], g[2], g[3]) for g in graphs]
        _, plain_dig = run_stream(plain, ids, graphs, part, rate=0, digests=True)
    arena.free()
    actx.close()
    same = sum(int(a == b) for a, b in zip(served_dig, plain_dig))
    res["verified"] = {"queries": len(part), "equal_to_plain": same,
                       "how": "xxh3-64 of each result's bytes: served into device memory vs "
                              "one query per launch into host memory"}
    if log:
        log(f"  verified: {same}/{len(part)} served results equal the plain ones")
    res["speedup_saturated"] = res["served_saturated"]["qps"] / cap
    res["speedup_vs_plain_same_slots"] = (res["served_saturated"]["qps"]
                                          / res["plain_slots_saturated"]["qps"])
    res["speedup_saturated_host"] = (res["served_saturated_host"]["qps"]
                                     / res["plain_saturated_host"]["qps"])
    res["config"] = {"slots": slots, "max_batch": max_batch, "hbm_budget": hbm_budget,
                     "queries_per_run": n_queries, "results": "caller-owned device arrays "
                     "(one per query) unless marked host",
                     "mix": "60% BFS / 30% SSSP / 10% TC", "pool": [g[0] for g in graphs]}
    return res
