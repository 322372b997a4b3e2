"""Benchmark of the hot path: single-source BFS (Boolean SpMSpV + visited
mask) on R-MAT scale 22 (4,194,304 vertices, edge factor 16, a/b/c/d =
.57/.19/.19/.05, seed 1, symmetric, int32 CSR), 64 non-isolated sources from
mt19937(3), each run separately with the graph resident in HBM.

A step is one BFS from one source (the 64 sources are cycled).  The line
printed is the whole-job GTEPS: sum over steps of the edges traversed
(sum of out-degrees of the reached vertices) / timed wall time.

    python bench.py [--gpus 1] [--steps 64] [--warmup 4] [--no-extra]

Extra fields (SSSP GTEPS, batched/served throughput) are measured after the
main line's timed region unless --no-extra is given.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import apspie_amd as A  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sources(rowptr, k, seed=3):
    """k distinct vertices with degree > 0: v = draw % m over std::mt19937(seed)."""
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    draws = bg.random_raw(4096 * k).astype(np.uint64)
    m = len(rowptr) - 1
    deg = np.diff(rowptr)
    out, seen = [], set()
    for d in draws.tolist():
        v = d % m
        if deg[v] > 0 and v not in seen:
            seen.add(v)
            out.append(int(v))
            if len(out) == k:
                return out
    raise RuntimeError("not enough sources")




def read_traffic(path: Path):
    """HBM bytes per launch of the level kernel, from the committed
    rocprofv3 --pmc summary (line 'traffic_bytes_per_launch: N')."""
    if not path.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", path.read_text())
    return float(m.group(1)) if m else None


def host_cpu():
    """The host CPU's model name and the threads this process may use."""
    model = "unknown"
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0)),
            "online": os.cpu_count()}


class one_core:
    """Pins the calling thread to one core (the last one this process may
    use, away from the runtime's helper threads on the first) for the
    single-threaded CPU timings, and restores the affinity afterwards."""

    def __enter__(self):
        self.saved = os.sched_getaffinity(0)
        self.core = max(self.saved)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *a):
        os.sched_setaffinity(0, self.saved)


def cpu_baseline(rowptr, col, srcs, budget_s=15.0):
    """The oracle's restatement of SimpleReferenceBfs (testBfs.cpp:6-72), one
    pinned core, on a bounded sample of the same sources."""
    from oracle import oracle as O
    rp64 = rowptr.astype(np.int64)
    m = len(rowptr) - 1
    deg = np.diff(rowptr)
    edges, t_tot, n = 0, 0.0, 0
    with one_core() as pin:
        for s in srcs:
            t0 = time.perf_counter()
            d, _ = O.bfs(m, rp64, col, s)
            t_tot += time.perf_counter() - t0
            edges += int(deg[d >= 0].sum())
            n += 1
            if t_tot > budget_s:
                break
    return {"value": edges / t_tot / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
            "pinned_core": pin.core, "host_cpu": host_cpu(),
            "sample": f"oracle deque BFS (SimpleReferenceBfs restated) from the first {n} of "
                      f"the 64 sources on the same RMAT-22 graph, {t_tot:.1f} s"}


def copy_calibration(nbytes, reps=3):
    """A plain device-to-device copy of a known size through the library."""
    with A.Context(0) as ctx:
        src = ctx.dev_alloc(nbytes)
        dst = ctx.dev_alloc(nbytes)
        for _ in range(reps):
            A._check(A.lib().apspie_memcpy(ctx.h, dst.ptr, src.ptr, nbytes, 2), ctx.h)
        ctx.sync()
        src.free()
        dst.free()
    log(f"copied {nbytes} bytes x{reps} (device to device)")


def cpu_config1(ctx):
    """The CPU reference path on its own case: undirected R-MAT scale 14
    (16,384 vertices, edge factor 16, seed 1), source = the highest-degree
    vertex (lowest id on ties), SimpleReferenceBfs (testBfs.cpp:6-72)
    restated in the oracle, one pinned core; the GPU BFS on the same graph
    and source beside it."""
    from oracle import oracle as O
    rowptr, col = A.gen_rmat(14)
    m = len(rowptr) - 1
    deg = np.diff(rowptr)
    src = int(np.argmax(deg))
    rp64 = rowptr.astype(np.int64)
    reps, t = 0, 0.0
    with one_core() as pin:
        while t < 2.0 or reps < 5:
            t0 = time.perf_counter()
            d, _ = O.bfs(m, rp64, col, src)
            t += time.perf_counter() - t0
            reps += 1
    edges = int(deg[d >= 0].sum())
    G = ctx.upload(rowptr, col)
    g = ctx.bfs(G, src)
    assert np.array_equal(g, d), "GPU BFS differs from the CPU reference path on RMAT-14"
    gpu = []
    for _ in range(20):
        ctx.bfs(G, src)
        gpu.append(ctx.stats()["elapsed_ms"])
    G.close()
    ms = 1e3 * t / reps
    return {"workload": "bfs_rmat14_maxdeg_source", "vertices": m, "nnz": len(col),
            "source": src, "edges_traversed": edges, "cpu_ms": ms,
            "cpu_gteps": edges / (ms / 1e3) / 1e9, "cpu_runs": reps, "cores": 1,
            "pinned_core": pin.core, "host_cpu": host_cpu(), "kind": "port",
            "gpu_ms_median": float(np.median(gpu)),
            "gpu_gteps": edges / (float(np.median(gpu)) / 1e3) / 1e9,
            "depths_equal": True}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--extras", default="sssp,mis,tc,spgemm,batched,served",
                    help="comma list of the extra workloads to run after the main line")
    ap.add_argument("--served-queries", type=int, default=2000,
                    help="length of the served query stream (same stream for every form)")
    ap.add_argument("--batch-graph", type=int, default=1,
                    help="replay batched level chunks from a captured hipGraph (1) or launch "
                         "them one by one (0)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES bytes device to device three times and exit: the "
                         "reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE "
                         "counters to bytes (each copy dispatch reads and writes BYTES)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the traversals "
              "are measured on one GPU's share; sharding by source or by rows is a design, "
              "not a run", file=sys.stderr)
        sys.exit(2)

    t0 = time.time()
    rowptr, col = A.gen_rmat(args.scale)
    m, nnz = len(rowptr) - 1, len(col)
    log(f"rmat{args.scale}: m={m} nnz={nnz} built in {time.time() - t0:.1f}s")
    srcs = sources(rowptr, 64)

    ctx = A.Context(0)
    G = ctx.upload(rowptr, col)
    out = ctx.dev_alloc(4 * m)
    optr = out.ptr
    deg = np.diff(rowptr)

    # the timed steps stride over all 64 sources (K < 64: every 64/K-th one;
    # K >= 64: the 64 in turn); the warm-up runs on sources that are not timed
    # when there are any
    timed = [srcs[(i * 64) // args.steps] if args.steps <= 64 else srcs[i % 64]
             for i in range(args.steps)]
    spare = [s for s in srcs if s not in set(timed)] or srcs
    for i in range(args.warmup):
        ctx.bfs(G, spare[i % len(spare)], out=optr, mem=A.DEVICE)
    ctx.sync()

    # ---- timed region: K single-source BFS runs.  The dominant kernel
    # (k_bfs_level: one level, push or pull) is timed live with HIP events on
    # the context stream: one interval around each call's back-to-back level
    # launches (the trailing launches that find the frontier empty included,
    # as rocprofv3 counts them); two events per call, a few microseconds.
    ctx.set_option("profiling", 1)
    edges, k_ms, k_launch, alg_bytes = 0, 0.0, 0, 0.0
    per_ms, per_gteps = [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        ctx.bfs(G, timed[i], out=optr, mem=A.DEVICE)
        st = ctx.stats()
        edges += st["edges_traversed"]
        per_ms.append(st["elapsed_ms"])  # hipEvent pair around the entry point
        per_gteps.append(st["edges_traversed"] / (st["elapsed_ms"] * 1e6))
        k_ms += st["kernel_ms"]
        k_launch += st["kernel_launches"]
        # algorithmic bytes of a traversal: 8 B per traversed edge (column
        # index + visited probe) and 20 B per reached vertex (row-offset pair,
        # depth write, frontier write + read); the 4 B per vertex of the depth
        # initialisation belong to the init kernel, not to the levels
        alg_bytes += 8.0 * st["edges_traversed"] + 20.0 * st["vertices_reached"]
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    ctx.set_option("profiling", 0)
    gteps = edges / elapsed / 1e9
    achieved = alg_bytes / (k_ms / 1e3)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": read_traffic(ROOT / "profiles" / "k_bfs_level_rmat22_pmc.md"),
                "kernel": "k_bfs_level<2048> (one BFS level, push or pull)",
                "avg_launch_us": 1e3 * k_ms / max(k_launch, 1),
                "alg_bytes_per_launch": alg_bytes / max(k_launch, 1),
                "launches": k_launch}

    rec = {"metric": "GTEPS (BFS, RMAT-22, single source, 64 sources)", "value": gteps,
           "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic (R-MAT generator, seed 1)",
           "config": {"workload": f"bfs_rmat{args.scale}_64src", "scale": args.scale,
                      "edge_factor": 16, "abcd": [0.57, 0.19, 0.19, 0.05], "vertices": m,
                      "nnz": nnz, "sources": 64, "source_rng": "mt19937(3) % m, degree > 0",
                      "timed_sources": len(set(timed)),
                      "timed_order": "every 64/K-th of the 64 sources for K < 64 steps"},
           "roofline": roofline,
           "per_source": {"median_ms": float(np.median(per_ms)),
                          "p99_ms": float(np.percentile(per_ms, 99)),
                          "harmonic_gteps": len(per_gteps) / sum(1.0 / g for g in per_gteps),
                          "timing": "device time per call (HIP events around the entry point)"},
           "graph500_gteps": gteps / 2}

    if not args.no_extra:
        ctx.set_option("batch_graph", args.batch_graph)
        rec["extra"] = extras(ctx, G, rowptr, col, srcs, optr, cpu=not args.no_cpu,
                              which=set(args.extras.split(",")),
                              served_queries=args.served_queries)
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(rowptr, col, srcs)
        rec["cpu_config1"] = cpu_config1(ctx)
        log(f"  config 1 (CPU path, RMAT-14): {rec['cpu_config1']}")
    else:
        rec["cpu_baseline"] = None
    G.close()
    ctx.close()
    print(json.dumps(rec), flush=True)


def extras(ctx, G, rowptr, col, srcs, optr, cpu=True, which=None, served_queries=240):
    """The other workloads of the path, each with its time, the roofline
    fraction where the per-unit bytes are defined, and a bounded CPU sample of
    the oracle on the same graph; then the batched and served forms."""
    which = which or {"sssp", "mis", "tc", "spgemm", "batched", "served"}
    ex = {}
    if "sssp" in which:
        ex["sssp_rmat22"] = sssp_rmat22(ctx, G, rowptr, col, srcs, optr, cpu)
    ex.update(other_workloads(ctx, cpu, which))
    if "batched" in which:
        ex["batched_lattice2048"] = batched(ctx)
    if "served" in which:
        from apspie_amd import servebench as SB
        log("serving pool:")
        pool = SB.build_pool(log=log)
        ex["served"] = SB.serve_benchmark(pool, n_queries=served_queries, log=log)
    return ex


def sssp_rmat22(ctx, G, rowptr, col, srcs, optr, cpu):
    """SSSP on the same RMAT-22 graph with integer weights 1..255, all 64 sources:
    GTEPS over the edges of the reached vertices, the roofline fraction at
    12 B per traversed edge (column, weight, distance probe) + 20 B per
    reached vertex + 4 B per vertex (init) per source, and one source of the
    oracle's Dijkstra on the host."""
    from oracle import oracle as O
    m = len(rowptr) - 1
    rp64 = rowptr.astype(np.int64)
    w = A.gen_weights(rowptr, col)
    Gw = ctx.upload(rowptr, col, w)
    fout = ctx.dev_alloc(4 * m)
    edges, relaxed, t, alg = 0, 0, 0.0, 0.0
    for s in srcs:  # the same reached set as SSSP's -> E_r, n_r per source
        ctx.bfs(G, s, out=optr, mem=A.DEVICE)
        st = ctx.stats()
        edges += st["edges_traversed"]
        alg += 12.0 * st["edges_traversed"] + 20.0 * st["vertices_reached"] + 4.0 * m
    ctx.sssp(Gw, srcs[0], out=fout.ptr, mem=A.DEVICE)  # warm-up (the light-first copy)
    for s in srcs:  # one query after another on the weighted graph
        ctx.sssp(Gw, s, out=fout.ptr, mem=A.DEVICE)
        st = ctx.stats()
        relaxed += st["edges_traversed"]
        t += st["elapsed_ms"]
    res = {"gteps": edges / (t / 1e3) / 1e9, "ms_per_source": t / len(srcs),
           "sources": len(srcs), "expanded_edges_per_traversed": relaxed / edges,
           "roofline_frac": alg / (t / 1e3) / HBM_PEAK}
    if cpu:
        deg = np.diff(rowptr)
        with one_core():
            t0 = time.perf_counter()
            d = O.sssp(m, rp64, col, w, srcs[0])
            tc_ = time.perf_counter() - t0
        res["cpu_baseline"] = {
            "value": int(deg[d >= 0].sum()) / tc_ / 1e9, "unit": "GTEPS", "cores": 1,
            "kind": "port", "sample": f"oracle binary-heap SSSP (SimpleReferenceSssp restated), "
                                      f"1 source of the same graph, {tc_:.1f} s"}
    Gw.close()
    fout.free()
    return res


def other_workloads(ctx, cpu, which):
    """MIS and TC on R-MAT, the general SpGEMM on R-MAT-14 (A * A)."""
    from oracle import oracle as O
    res = {}
    if which & {"mis", "tc"}:
        rp20, col20 = A.gen_rmat(20)
        m20 = len(rp20) - 1
        G20 = ctx.upload(rp20, col20)
        if "mis" in which:
            mout = ctx.dev_alloc(4 * m20)  # result stays in HBM, as for the traversals
            ctx.mis(G20, out=mout.ptr, mem=A.DEVICE)
            t0 = time.perf_counter()
            ctx.mis(G20, out=mout.ptr, mem=A.DEVICE)
            ms = (time.perf_counter() - t0) * 1e3
            mout.free()
            st = ctx.stats()
            # per round: 8 B per candidate + 8 B per candidate edge, and 8 B per
            # vertex for the state
            alg = (8.0 * st["vertices_reached"] + 8.0 * st["edges_traversed"]
                   + 8.0 * m20 * st["iterations"])
            res["mis_rmat20"] = {"ms": ms, "rounds": st["iterations"],
                                 "roofline_frac": alg / (ms / 1e3) / HBM_PEAK}
            if cpu:
                prio = O.mis_priorities(m20)
                t0 = time.perf_counter()
                O.mis_greedy(m20, rp20.astype(np.int64), col20, prio)
                res["mis_rmat20"]["cpu_baseline"] = {
                    "value": (time.perf_counter() - t0) * 1e3, "unit": "ms", "cores": 1, "kind": "port",
                    "sample": "oracle greedy MIS by descending priority (the set Luby rounds converge "
                              "to), same graph"}
        if "tc" in which:
            # TC: 4 B per row offset and per L entry, 4 B per wedge of L.L (the
            # reference formulation's sum over (i,k) in L of |L_k|) and 4 B per row
            rows = np.repeat(np.arange(m20), np.diff(rp20))
            low = col20 < rows
            lcnt = np.bincount(rows[low], minlength=m20).astype(np.int64)
            wedges = int(lcnt[col20[low]].sum())
            alg = 4.0 * (m20 + 1) + 4.0 * int(low.sum()) + 4.0 * wedges + 4.0 * m20
            ctx.tc(G20, per_row=True)
            t0 = time.perf_counter()
            total, _ = ctx.tc(G20, per_row=True)
            ms = (time.perf_counter() - t0) * 1e3
            res["tc_rmat20"] = {"ms": ms, "triangles": int(total), "reference_wedges": wedges,
                                "probes": ctx.stats()["edges_traversed"],
                                "roofline_frac": alg / (ms / 1e3) / HBM_PEAK}
            res["tc_rmat20"]["autotuned"] = tuned_vs_default(
                ctx, lambda: ctx.tc(G20, per_row=True), lambda: ctx.autotune_hash(G20),
                ("tc_medium_row",))
        G20.close()
    if "tc" in which:
        rp18, col18 = A.gen_rmat(18)
        G18 = ctx.upload(rp18, col18)
        ctx.tc(G18)
        t0 = time.perf_counter()
        ctx.tc(G18)
        res["tc_rmat18"] = {"ms": (time.perf_counter() - t0) * 1e3}
        G18.close()
        if cpu:
            t0 = time.perf_counter()
            O.tc(len(rp18) - 1, rp18.astype(np.int64), col18)
            res["tc_rmat18"]["cpu_baseline"] = {
                "value": (time.perf_counter() - t0) * 1e3, "unit": "ms", "cores": 1, "kind": "port",
                "sample": "oracle merge-intersection TC (ewiseMultTc semantics), same graph"}
    if "spgemm" in which:
        # SpGEMM C = A * A on R-MAT-14 (unit values)
        rp14, col14 = A.gen_rmat(14)
        G14 = ctx.upload(rp14, col14, np.ones(len(col14), np.float32))
        ctx.spgemm(G14, G14).close()
        t0 = time.perf_counter()
        C = ctx.spgemm(G14, G14)
        ms = (time.perf_counter() - t0) * 1e3
        # products sum_k deg(k)^2 (A symmetric): 4 B each in the symbolic pass
        # (B's column), 8 B in the numeric pass (column + value); 8 B per
        # output entry, 8 B per entry of A per pass
        deg14 = np.diff(rp14).astype(np.int64)
        prods = int((deg14 * deg14).sum())
        alg = 12.0 * prods + 8.0 * C.nnz + 16.0 * len(col14)
        res["spgemm_rmat14_AxA"] = {"ms": ms, "nnz_c": C.nnz, "products": prods,
                                    "roofline_frac": alg / (ms / 1e3) / HBM_PEAK}
        C.close()
        res["spgemm_rmat14_AxA"]["autotuned"] = tuned_vs_default(
            ctx, lambda: ctx.spgemm(G14, G14).close(), lambda: ctx.autotune_hash(G14),
            ("spgemm_wave_ub", "spgemm_block_ub", "spgemm_dense"))
        G14.close()
        if cpu:
            m14 = len(rp14) - 1
            t0 = time.perf_counter()
            O.spgemm(m14, m14, m14, rp14.astype(np.int64), col14, None, rp14.astype(np.int64),
                     col14, None)
            res["spgemm_rmat14_AxA"]["cpu_baseline"] = {
                "value": (time.perf_counter() - t0) * 1e3, "unit": "ms", "cores": 1, "kind": "port",
                "sample": "oracle Gustavson SpGEMM, same matrices"}
    for k, v in res.items():
        log(f"  {k}: {v}")
    return res


TUNABLE_DEFAULTS = {"tc_medium_row": 512, "spgemm_wave_ub": 256, "spgemm_block_ub": 4096,
                    "spgemm_dense": 1}


def tuned_vs_default(ctx, run, tune, names, reps=5):
    """Median wall time of run() at the default options, then after tune()
    (the autotuner's picks for `names` reported), then the defaults again."""
    def med():
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))
    run()
    base = med()
    t0 = time.perf_counter()
    tune()
    tune_s = time.perf_counter() - t0
    picked = {k: ctx.get_option(k) for k in names}
    tuned = med()
    for k in names:
        ctx.set_option(k, TUNABLE_DEFAULTS[k])
    return {"ms_default": base, "ms_tuned": tuned, "speedup": base / tuned, "picked": picked,
            "tuning_s": tune_s}


def batched(ctx):
    """Batched multi-source BFS / SSSP on the 2048^2 lattice (10% of edges
    removed, 0.1% long-range edges, integer weights 1..255), sources uniform
    over the largest component, against the plain per-source loop."""
    rp, col = A.gen_lattice(2048)
    w = A.gen_weights(rp, col)
    G = ctx.upload(rp, col, w)
    m = len(rp) - 1
    # largest component = vertices reached from a high-degree vertex of it
    depth = ctx.bfs(G, int(np.argmax(np.diff(rp))))
    comp = np.nonzero(depth >= 0)[0]
    rng = np.random.default_rng(4)
    srcs = [int(x) for x in rng.choice(comp, 512, replace=False)]
    out = ctx.dev_alloc(4 * m)
    res = {"largest_component": int(len(comp))}
    for kind, fn, fb in (("bfs", ctx.bfs, ctx.bfs_batch), ("sssp", ctx.sssp, ctx.sssp_batch)):
        n_plain = 8
        fn(G, srcs[-1], out=out.ptr, mem=A.DEVICE)  # warm
        t0 = time.perf_counter()
        for s in srcs[:n_plain]:
            fn(G, s, out=out.ptr, mem=A.DEVICE)
        plain_ms = (time.perf_counter() - t0) * 1e3 / n_plain
        res[f"{kind}_plain_ms_per_query"] = plain_ms
        for B in (32, 128, 512):
            buf = ctx.dev_alloc(4 * B * m)
            fb(G, srcs[:B], out=buf.ptr, mem=A.DEVICE)  # scratch sizing, graph capture
            t0 = time.perf_counter()
            fb(G, srcs[:B], out=buf.ptr, mem=A.DEVICE)
            ms = (time.perf_counter() - t0) * 1e3
            buf.free()
            res[f"{kind}_batch{B}_ms_per_query"] = ms / B
            res[f"{kind}_batch{B}_speedup"] = plain_ms / (ms / B)
            log(f"  lattice2048 {kind} batch {B}: {ms / B:.3f} ms/query vs plain "
                f"{plain_ms:.3f} ({plain_ms / (ms / B):.1f}x)")
    out.free()
    G.close()
    return res


if __name__ == "__main__":
    main()
