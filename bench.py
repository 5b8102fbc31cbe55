#!/usr/bin/env python3
"""Benchmark of the graph hot path on one MI355X.

Workload (R-MAT scale 22, edge factor 16, a/b/c/d = 0.57/0.19/0.19/0.05,
seed 1, symmetrised and deduplicated, uint32 CSR, weights uniform in [1, 64]
from seed 2; 8 sources drawn with seed 3 among vertices with degree > 0):
one step = BFS from the 8 sources + SSSP from the 8 sources + CC + PageRank
(d = 0.85, 20 iterations). The graph and every output stay in HBM; no host
copy runs inside the timed region.

value = GTEPS of the traversal queries: sum over the 16 BFS/SSSP queries of
m_r (sum of out-degrees of the reached vertices, directed edges) divided by
their wall time. PageRank iterations/s and CC time are reported beside it.

roofline: the query kind with the largest summed main-kernel time in the
warmup steps is the dominant kernel, and only its launches are timed (HIP
events around each launch on the graph's stream) in the timed steps; for
BFS or SSSP those of half the sources of a step, the other half in the
next step (the events cost a query about 10% of its rate, now paid by half
the queries). The timed queries' algorithmic bytes (BFS 12*V_r + 8*m_r, SSSP
12*V_r + 12*m_r, CC 16*V + 8*E, PageRank 8*E + 20*V per iteration) divided by
their HIP-event-timed launch durations give `achieved`.

cpu_baseline: the oracle's restatement of the reference's frontier loop
(oracle/oracle.c, orc_ref_loop) on the host's stated cores and on one
thread, BFS and SSSP from the same 8 sources (16 queries), as GTEPS over the
same m_r; the reference's host loops (degree pass, overload gather) timed
on the partition graph ride along as host_loops.

After the timed steps (none of this is in `value`):
  batched  64 sources (seed 3) per launch vs the same 64 single queries,
           median of 3 passes each: queries/s, on the bench graph and at the
           stated size (R-MAT scale 24 and Erdos-Renyi scale 24), batch knobs
           picked by the autotuner.
  partition  R-MAT scale 26 with its edges in pinned host memory, streamed
           in 256 MB chunks into a device cache under a 4 GB budget: BFS and
           PageRank(10), host-link GB/s against a pinned copy probe; then the
           reference's host loops on the same graph (degree pass, and the
           overload gather of every BFS level under the reference's static
           partition for that budget).
  serving  6 resident R-MAT graphs (scales 18..24), 3000 requests, mix 50%
           BFS / 30% SSSP / 15% CC / 5% PageRank(10), Poisson arrivals at
           700/s (seed 4, near the plain path's saturation) under a 24 GiB
           query-state budget (the graphs' batched state alone would take
           about 27 GB, so the budget forces evictions), each graph autotuned
           first for single queries and batches (PageRank by the in-list pull
           and SSSP without sweep rounds, whose binned layout would crowd the
           budget): the plain per-query path (one at a time) vs the server
           (continuous batching); p50/p99, mean batch and evictions at that
           load, at the same arrivals compressed to 1600/s (above the plain
           path's saturation), and with every request queued at once; every
           request's checksum must be the same in every phase of both.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "test-world_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_rmat22.md")


def measured_traffic(kernel, scale):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    summary (two rocprofv3 --pmc passes scaled by a copy kernel), or None when
    that summary does not cover this kernel and workload."""
    if scale != 22 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            if line.startswith("traffic_per_launch:"):
                name, _, val = line.split(":", 1)[1].rpartition("=")
                if name.strip() == kernel:
                    return int(val.strip())
    return None


def host_info():
    """The host the CPU baseline runs on: logical CPUs, the ones this process
    may use, the CPU model and the thread count the machine's environment
    states (OMP_NUM_THREADS)."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = info["nproc"]
    info["model"] = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    info["omp_num_threads"] = int(omp) if omp.isdigit() else None
    return info


def host_cores(info):
    """Threads for host work: the CPUs this process may use, or fewer when the
    environment states a smaller share (OMP_NUM_THREADS)."""
    n = info["affinity"] or 1
    if info["omp_num_threads"]:
        n = min(n, info["omp_num_threads"])
    return max(1, n)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def batched_leg(g, srcs, V, pg, log, label):
    """64 sources per launch vs the same 64 queries one at a time (outputs in
    HBM, equal results): queries/s of each, the median of 3 passes for both
    forms, and their ratio."""
    import numpy as np

    n = len(srcs)
    buf = pg.DeviceBuffer(4 * n * V)
    one = pg.DeviceBuffer(4 * V)
    res = {"sources": n, "graph": label, "passes": 3, "statistic": "median"}
    try:
        for kind in ("bfs", "sssp"):
            fb = g.bfs_batch if kind == "bfs" else g.sssp_batch
            f1 = g.bfs if kind == "bfs" else g.sssp
            fb(srcs, out=buf)
            f1(int(srcs[0]), out=one)
            tb, t1 = [], []
            for _ in range(3):
                t = time.perf_counter()
                fb(srcs, out=buf)
                tb.append(time.perf_counter() - t)
                t = time.perf_counter()
                for s in srcs:
                    f1(int(s), out=one)
                t1.append(time.perf_counter() - t)
            mb, m1 = sorted(tb)[1], sorted(t1)[1]
            res[kind + "_batched_qps"] = round(n / mb, 1)
            res[kind + "_plain_qps"] = round(n / m1, 1)
            res[kind + "_ratio"] = round(m1 / mb, 2)
        # spot check: one row of each batch against the single query
        g.bfs_batch(srcs, out=buf)
        k = n // 2
        row = buf.view(4 * k * V, 4 * V).to_host(np.uint32, V)
        res["bfs_row_equal"] = bool((row == g.bfs(int(srcs[k]))[0]).all())
        g.sssp_batch(srcs, out=buf)
        row = buf.view(4 * k * V, 4 * V).to_host(np.uint32, V)
        res["sssp_row_equal"] = bool((row == g.sssp(int(srcs[k]))[0]).all())
    finally:
        buf.free()
        one.free()
    log("batched %s: %s" % (label, res))
    return res


def batched_scale24_leg(args, pg, np, cores, log):
    """The batched form at its stated size: R-MAT scale 24 and an
    Erdos-Renyi graph with the same V and edge factor, 64 BFS + 64 SSSP
    sources (seed 3, degree > 0), batch knobs picked by the autotuner."""
    out = {}
    for name in args.batch_graphs.split(","):
        if not name:
            continue
        t = time.time()
        sc = int(name[-2:])
        if name.startswith("rmat"):
            csr = pg.gen_rmat(sc, 16, 0.57, 0.19, 0.19, seed=1, threads=cores)
        else:
            csr = pg.gen_er(sc, 16, seed=1, threads=cores)
        csr.add_weights(1, 64, seed=2, threads=cores)
        deg = np.diff(csr.offsets.astype(np.int64))
        srcs = np.array(np.random.default_rng(3).choice(np.nonzero(deg > 0)[0], 64, replace=False), dtype=np.uint32)
        g = pg.Graph.from_csr(csr)
        V, E = csr.V, csr.E
        del csr
        try:
            log("batched %s: V=%d E=%d ready in %.1fs" % (name, V, E, time.time() - t))
            tuned = None
            if not args.no_tune:
                c = g.autotune(srcs, what=pg.TUNE_BFS_BATCH | pg.TUNE_SSSP_BATCH)
                tuned = {f: getattr(c, f) for f in ("batch_td_to_bu", "batch_grid", "batch_sssp_grid", "batch_delta", "batch_chunk", "batch_pull")}
            r = batched_leg(g, srcs, V, pg, log, name)
            r.update({"V": V, "E": E, "tuning": tuned})
            out[name] = r
        finally:
            g.close()
    return out


def serving_leg(args, pg, np, cores, log):
    """Poisson stream (seed 4) over resident R-MAT graphs, mix 50% BFS / 30%
    SSSP / 15% CC / 5% PageRank (10 iterations), uniform graphs and sources;
    the plain per-query path (one query at a time, no batching) against the
    server (continuous batching, coalescing, one worker per graph) at the same
    offered load, then both saturated with every request at once. Each graph
    is autotuned first (single-query and batch knobs), and both servers use
    the tuned graphs."""
    scales = [int(x) for x in args.serve_scales.split(",") if x]
    graphs = []
    t = time.time()
    tune_rng = np.random.default_rng(5)
    for sc in scales:
        c = pg.gen_rmat(sc, 16, 0.57, 0.19, 0.19, seed=1, threads=cores).add_weights(1, 64, seed=2, threads=cores)
        # PageRank by the in-list pull and SSSP without sweep rounds: the
        # binned layout both would use is about fifteen times PageRank's pull
        # state, and under the state budget it would evict the batch state of
        # the other graphs
        gr = pg.Graph.from_csr(c, config=pg.default_config(pr_layout=pg.PR_PULL, sssp_sweep=-1.0))
        if not args.no_tune:
            # the autotuner picks each graph's knobs for both paths (single
            # queries and batches) on 64 of its own sources
            deg = np.diff(c.offsets.astype(np.int64))
            ts = np.array(tune_rng.choice(np.nonzero(deg > 0)[0], 64, replace=False), dtype=np.uint32)
            gr.autotune(ts, what=pg.TUNE_BFS | pg.TUNE_SSSP | pg.TUNE_BFS_BATCH | pg.TUNE_SSSP_BATCH)
        graphs.append(gr)
        del c
    log("serving: %d graphs (scales %s) ready and tuned in %.1fs" % (len(graphs), scales, time.time() - t))
    rng = np.random.default_rng(4)
    N = args.serve_requests
    kinds = rng.choice([pg.Q_BFS, pg.Q_SSSP, pg.Q_CC, pg.Q_PAGERANK], N, p=[0.5, 0.3, 0.15, 0.05])
    gids = rng.integers(len(graphs), size=N)
    reqs = [pg.Server.request(int(gids[i]), int(kinds[i]), source=int(rng.integers(graphs[gids[i]].V)),
                              iterations=10, damping=0.85, tag=i) for i in range(N)]
    gaps = rng.exponential(1.0, N)
    arrivals = np.cumsum(gaps * (1e9 / args.serve_qps)).astype(np.uint64)
    # the same arrival pattern compressed to the overload rate: above what the
    # plain path can serve, so its queue grows while the server batches
    overload = np.cumsum(gaps * (1e9 / args.serve_overload_qps)).astype(np.uint64)
    out = {"graphs": ["rmat%d" % s for s in scales], "requests": N, "offered_qps": args.serve_qps,
           "overload_qps": args.serve_overload_qps, "mix": "50% BFS, 30% SSSP, 15% CC, 5% PageRank(10)"}
    budget = int(args.serve_budget_gb * (1 << 30))
    out["state_budget_bytes"] = budget
    sums = {}
    for mode, kw in (("plain", dict(max_batch=1, max_inflight=1, coalesce=False, state_budget_bytes=budget)),
                     ("served", dict(state_budget_bytes=budget))):
        with pg.Server(**kw) as srv:
            for gr in graphs:
                srv.add_graph(gr)
            srv.replay(reqs[:300], np.zeros(300, np.uint64))  # warm: state, captures
            srv.drain()
            srv.completions()
        res = {}
        for phase, arr in (("offered", arrivals), ("overloaded", overload), ("saturated", np.zeros(N, np.uint64))):
            with pg.Server(**kw) as srv:
                for gr in graphs:
                    srv.add_graph(gr)
                t = time.perf_counter()
                srv.replay(reqs, arr)
                srv.drain()
                el = time.perf_counter() - t
                s = srv.summary()
                comps = srv.completions()
                sums[(mode, phase)] = {c.tag: c.checksum for c in comps if c.status == 0}
                res[phase] = {"qps": round(s.completed / el, 1), "p50_ms": round(s.p50_ns / 1e6, 3),
                              "p99_ms": round(s.p99_ns / 1e6, 3), "mean_batch": round(s.mean_batch, 2),
                              "ok": int(sum(c.status == 0 for c in comps)), "rejected": int(s.rejected),
                              "evictions": int(s.evictions), "state_bytes_peak": int(s.state_bytes_peak)}
        out[mode] = res
        log("serving %s: %s" % (mode, res))
    out["saturated_qps_ratio"] = round(out["served"]["saturated"]["qps"] / out["plain"]["saturated"]["qps"], 2)
    out["overloaded_qps_ratio"] = round(out["served"]["overloaded"]["qps"] / out["plain"]["overloaded"]["qps"], 2)
    # every phase of both servers returns the same result for each request
    ref = sums[("plain", "offered")]
    out["results_identical"] = all(v == ref for v in sums.values())
    for gr in graphs:
        gr.close()
    return out


def reference_host_loops(ptr, col, levels, budget, np, log):
    """The reference's host loops on the partition graph, timed as written
    (oracle/oracle.c): the degree pass of ptGraph/main.cu:1703-1710 (median
    of 3), and the overload gather of ptGraph/main.cu:1126-1156 with its
    20-thread fan-out (:1915-1938) for every BFS level from vertex 0. The
    overload vertices of a level are its vertices whose edges are not all in
    the static prefix of P edges, P = 0.6 * (budget - 40 * V) / 4 rounded
    down to whole 4096-edge fragments (main.cu:1693-1700, :1759-1765)."""
    import oracle

    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()
    V, E = ptr.size, col.size
    runs = []
    for _ in range(3):
        deg, ns = oracle.degree_pass(ptr, E)
        runs.append(ns)
    t_deg = sorted(runs)[1]
    P = int(0.6 * (budget - 40 * V) / 4)
    P -= P % 4096
    p64 = ptr.astype(np.int64)
    static = (p64 < P) & (p64 + deg.astype(np.int64) - 1 < P)
    t_gather, edges, nodes, nlev = 0, 0, 0, 0
    top = int(levels[levels != 0xFFFFFFFF].max())
    for lev in range(1, top + 1):
        active = np.nonzero(levels == lev)[0]
        over = active[~static[active]].astype(np.uint32)
        if over.size == 0:
            continue
        got, ns = oracle.overload_gather(ptr, col, deg, over, threads=20)
        t_gather += ns
        edges += got.size
        nodes += over.size
        nlev += 1
        del got
    res = {
        "degree_pass": {"ms": round(t_deg / 1e6, 2), "vertices": V,
                        "gbs": round(8.0 * V / max(t_deg, 1), 2)},  # reads ptr[i], ptr[i+1]; writes degree[i]
        "overload_gather": {"ms": round(t_gather / 1e6, 1), "levels": nlev, "vertices": nodes, "edges": edges,
                            "static_prefix_edges": P, "threads": 20,
                            "gbs": round(8.0 * edges / max(t_gather, 1), 2)},  # 4 B read + 4 B written per edge
    }
    log("reference host loops: %s" % res)
    return res


def partition_leg(args, pg, np, cores, log):
    """Edges kept in pinned host memory and streamed in 256 MB chunks into a
    device cache under a device budget (R-MAT scale 26 under 4 GB by
    default): BFS from vertex 0 (3 runs) and PageRank (10 iterations), with
    the host-link rate against a pinned copy probe and the bytes moved
    against the lower bound (4 bytes per column of every reached vertex for
    BFS; the non-resident columns, every iteration, for PageRank). Then the
    reference's host loops on the same host CSR."""
    sc = args.partition_scale
    t = time.time()
    csr = pg.gen_rmat(sc, 16, 0.57, 0.19, 0.19, seed=1, threads=cores)
    V, E = csr.V, csr.E
    budget = int(args.partition_budget_gb * (1 << 30))
    cfg = pg.default_config(flags=pg.FLAG_STREAM, chunk_bytes=256 << 20, device_budget_bytes=budget)
    g = pg.Graph(csr.node_pointers, csr.cols, config=cfg)
    chunks, slots, slot_bytes = g.stream_info()
    log("partition: rmat%d V=%d E=%d, %d chunks, %d slots, device %.2f GB, ready in %.1fs" % (
        sc, V, E, chunks, slots, g.device_bytes / 1e9, time.time() - t))
    probe = pg.host_link_probe(256 << 20, 8)
    out = pg.DeviceBuffer(4 * V)
    res = {"graph": "rmat%d" % sc, "V": V, "E": E, "device_budget_bytes": budget, "device_bytes": g.device_bytes,
           "chunks": chunks, "slots": slots, "chunk_bytes": slot_bytes, "host_link_probe_gbs": round(probe, 2)}
    levels = None
    try:
        runs = []
        for _ in range(3):
            t = time.perf_counter()
            _, st = g.bfs(0, out=out)
            dt = time.perf_counter() - t
            runs.append((dt, st.edges_traversed, st.host_link_bytes, st.host_link_ns, st.chunk_loads, st.chunk_hits))
        dt, m_r, lb, lns, loads, hits = min(runs)
        res["bfs"] = {"ms": round(dt * 1e3, 1), "gteps": round(m_r / dt / 1e9, 3),
                      "link_bytes": lb, "link_gbs": round(lb / max(lns, 1), 2),
                      "link_vs_lower_bound": round(lb / (4.0 * m_r), 3), "chunk_loads": loads, "chunk_hits": hits}
        if not args.no_cpu:
            levels = out.to_host(np.uint32, V)
        t = time.perf_counter()
        _, st = g.pagerank(0.85, 10, out=out)
        dt = time.perf_counter() - t
        lower = 10 * max(0, 4 * E - slots * slot_bytes)
        res["pagerank10"] = {"ms": round(dt * 1e3, 1), "iters_per_s": round(10 / dt, 2),
                             "link_bytes": st.host_link_bytes,
                             "link_gbs": round(st.host_link_bytes / max(st.host_link_ns, 1), 2),
                             "link_vs_lower_bound": round(st.host_link_bytes / max(lower, 1), 3),
                             "chunk_loads": st.chunk_loads, "chunk_hits": st.chunk_hits}
        t = time.perf_counter()
        _, st = g.cc(out=out)
        dt = time.perf_counter() - t
        res["cc"] = {"ms": round(dt * 1e3, 1), "components": st.reached, "link_bytes": st.host_link_bytes,
                     "link_gbs": round(st.host_link_bytes / max(st.host_link_ns, 1), 2)}

        def bfs_pr(tag):
            t = time.perf_counter()
            _, sb = g.bfs(0, out=out)
            tb = time.perf_counter() - t
            t = time.perf_counter()
            _, sp = g.pagerank(0.85, 10, out=out)
            tp = time.perf_counter() - t
            r = {"bfs_ms": round(tb * 1e3, 1), "bfs_link_bytes": sb.host_link_bytes,
                 "pagerank10_ms": round(tp * 1e3, 1), "pagerank10_link_bytes": sp.host_link_bytes}
            log("partition %s: %s" % (tag, r))
            return r

        # degree-aware residency: the leading chunks (the highest-degree
        # vertices) pinned in their slots, against least-recently-used only
        g.tune(stream_pinned=max(0, slots - 2))
        bfs_pr("warm")
        res["pinned"] = dict(stream_pinned=max(0, slots - 2), **bfs_pr("pinned"))
        g.tune(stream_pinned=0)
        if not args.no_tune:
            t = time.time()
            c = g.autotune(np.array([0], np.uint32), what=pg.TUNE_STREAM)
            res["tuned"] = dict(chunk_bytes=c.chunk_bytes, stream_pinned=c.stream_pinned,
                                chunks=g.stream_info()[0], slots=g.stream_info()[1],
                                tune_s=round(time.time() - t, 1), **bfs_pr("tuned"))
    finally:
        out.free()
        g.close()
    log("partition: %s" % res)
    host_loops = None
    if levels is not None:
        host_loops = reference_host_loops(csr.node_pointers, csr.cols, levels, budget, np, log)
        host_loops["graph"] = "rmat%d" % sc
    del csr
    return res, host_loops


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--sources", type=int, default=8)
    ap.add_argument("--pr-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-timing", action="store_true",
                    help="do not time main-kernel launches with HIP events (no roofline)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-source leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tuning (no autotuner run)")
    ap.add_argument("--no-partition", action="store_true", help="skip the streamed (partition) leg")
    ap.add_argument("--partition-scale", type=int, default=26)
    ap.add_argument("--partition-budget-gb", type=float, default=4.0)
    ap.add_argument("--serve-scales", default="18,20,21,22,23,24")
    ap.add_argument("--serve-qps", type=float, default=700.0)
    ap.add_argument("--serve-overload-qps", type=float, default=1600.0,
                    help="second offered load, above the plain path's saturation")
    ap.add_argument("--serve-budget-gb", type=float, default=24.0,
                    help="query-state budget of both servers (0 = unlimited)")
    ap.add_argument("--batch-graphs", default="rmat24,er24",
                    help="graphs of the batched leg at the stated size (empty: skip)")
    ap.add_argument("--serve-requests", type=int, default=3000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB 3 times (hardware-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import ptgraph as pg

    pg.init(0)
    if args.copy_probe:
        # the library's streaming copy kernel (k_copy_bytes) reads and writes
        # exactly MIB MiB each time, so its byte counters calibrate those of
        # the graph kernels in the same profiled run
        nb = args.copy_probe << 20
        a, b = pg.DeviceBuffer(nb), pg.DeviceBuffer(nb)
        for _ in range(3):
            b.copy_from(a)
        a.free()
        b.free()
        log("copy probe: 3 x %d MiB" % args.copy_probe)
    hinfo = host_info()
    cores = host_cores(hinfo)
    t0 = time.time()
    csr = pg.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=1, threads=cores)
    csr.add_weights(1, 64, seed=2, threads=cores)
    deg = np.diff(csr.offsets.astype(np.int64))
    rng = np.random.default_rng(3)
    srcs = [int(s) for s in rng.choice(np.nonzero(deg > 0)[0], args.sources, replace=False)]
    log("graph V=%d E=%d generated in %.1fs; sources %s" % (csr.V, csr.E, time.time() - t0, srcs))

    cfg = pg.default_config(flags=0)
    g = pg.Graph.from_csr(csr, config=cfg)
    V, E = csr.V, csr.E
    out = pg.DeviceBuffer(4 * V)

    kinds = ("bfs", "sssp", "cc", "pagerank")
    srcs64 = np.array(rng.choice(np.nonzero(deg > 0)[0], 64, replace=False), dtype=np.uint32)

    def set_timing(kind, timed):
        on = not args.no_timing and (timed == "all" or kind == timed)
        g.set_flags(pg.FLAG_TIMING if on else 0)

    def step(acc, timed, half=None):
        # half: time the traversal queries of the sources with index parity
        # `half` only (None: all of them)
        for kind, fn in (("bfs", g.bfs), ("sssp", g.sssp)):
            for i, s in enumerate(srcs):
                set_timing(kind, timed if half is None or i % 2 == half else None)
                t = time.perf_counter()
                _, st = fn(s, out=out)
                acc[kind].append((time.perf_counter() - t, st.edges_traversed, st.reached, st.main_kernel_ns,
                                  st.main_launches, st.iterations))
        set_timing("cc", timed)
        t = time.perf_counter()
        _, st = g.cc(out=out)
        acc["cc"].append((time.perf_counter() - t, E, V, st.main_kernel_ns, st.main_launches, st.iterations))
        set_timing("pagerank", timed)
        t = time.perf_counter()
        _, st = g.pagerank(0.85, args.pr_iters, out=out)
        acc["pagerank"].append((time.perf_counter() - t, E, V, st.main_kernel_ns, st.main_launches,
                                st.iterations))

    # Warmup steps time every kind's main kernel to find the dominant one; the
    # timed steps then time only that kernel, so the per-launch events do not
    # slow the other queries.
    wacc = {k: [] for k in kinds}
    for i in range(args.warmup):
        step(wacc, "all")
        log("warmup step %d done" % (i + 1))
    dom = max(kinds, key=lambda k: sum(r[3] for r in wacc[k])) if wacc["bfs"] else "sssp"
    log("dominant kernel kind: %s" % dom)
    acc = {k: [] for k in kinds}
    t_start = time.perf_counter()
    for i in range(args.steps):
        # the dominant kernel's launches are timed for half the traversal
        # sources of a step, the other half in the next step, so that every
        # source is timed equally often while the events slow half as many
        # queries
        step(acc, dom, half=i % 2 if dom in ("bfs", "sssp") else None)
    t_total = time.perf_counter() - t_start
    log("timed %d steps in %.3fs" % (args.steps, t_total))

    # One more pass of the traversal queries with host-memory outputs: the
    # rate a caller sees when the result crosses PCIe (reported beside value).
    g.set_flags(0)
    pcie = {}
    for kind, fn in (("bfs", g.bfs), ("sssp", g.sssp)):
        t = time.perf_counter()
        m = 0
        for s in srcs:
            _, st = fn(s)
            m += st.edges_traversed
        pcie[kind + "_gteps"] = round(m / (time.perf_counter() - t) / 1e9, 4)

    # The single-query path keeps its default tuning (the profiles under
    # profiles/ describe it); the autotuner picks the batched kernels' knobs.
    tuned = None
    if not args.no_tune and not args.no_batch:
        t = time.time()
        c = g.autotune(srcs64, what=pg.TUNE_BFS_BATCH | pg.TUNE_SSSP_BATCH)
        tuned = {f: getattr(c, f) for f in ("batch_td_to_bu", "batch_grid", "batch_sssp_grid", "batch_delta", "batch_chunk", "batch_pull")}
        log("autotuner %.1fs: %s" % (time.time() - t, tuned))
    batched = None
    if not args.no_batch:
        batched = {"rmat%d" % args.scale: batched_leg(g, srcs64, V, pg, log, "rmat%d (the bench graph)" % args.scale)}
        batched.update(batched_scale24_leg(args, pg, np, cores, log))
    serving = None if args.no_serve else serving_leg(args, pg, np, cores, log)
    partition, host_loops = (None, None) if args.no_partition else partition_leg(args, pg, np, cores, log)

    def tot(k, idx):
        return sum(r[idx] for r in acc[k])

    trav_edges = tot("bfs", 1) + tot("sssp", 1)
    trav_time = tot("bfs", 0) + tot("sssp", 0)
    gteps = trav_edges / trav_time / 1e9
    bfs_gteps = tot("bfs", 1) / tot("bfs", 0) / 1e9
    sssp_gteps = tot("sssp", 1) / tot("sssp", 0) / 1e9
    pr_iters_s = args.pr_iters * len(acc["pagerank"]) / tot("pagerank", 0)
    cc_ms = 1e3 * tot("cc", 0) / len(acc["cc"])

    # algorithmic bytes of each kind's main kernel, summed over the timed region
    def algo_bytes(k, r):
        _, m_r, v_r = r[0], r[1], r[2]
        if k == "bfs":
            return 12 * v_r + 8 * m_r
        if k == "sssp":
            return 12 * v_r + 12 * m_r
        if k == "cc":
            return 16 * V + 8 * E
        return args.pr_iters * (8 * E + 20 * V)

    def kind_stats(accd):
        res = {}
        for k in kinds:
            ns = sum(r[3] for r in accd[k])
            launches = sum(r[4] for r in accd[k])
            b = sum(algo_bytes(k, r) for r in accd[k] if r[4])  # the timed queries' bytes
            res[k] = {
                "main_kernel_ms": ns / 1e6,
                "launches": launches,
                "algo_bytes": b,
                "achieved_gbs": b / ns if ns else 0.0,
                "avg_launch_us": ns / launches / 1e3 if launches else 0.0,
            }
        return res

    per_kind = kind_stats(acc)
    per_kind_warm = kind_stats(wacc)
    # SSSP: a round is the push kernel plus, behind it, the pull kernel and
    # the two sweep kernels (whichever of them expands the round)
    # PageRank: an iteration is the binned accumulate kernel plus, behind it,
    # the kernel that finishes the vertices (the layout the graph's automatic
    # choice takes at R-MAT scale 22)
    kernel_names = {"bfs": "k_bfs_round", "sssp": "k_relax_round<true, false> + k_pull_round + k_sw_acc + k_sw_fin",
                    "cc": "k_expand<HookRest>", "pagerank": "k_prb_acc + k_prb_fin"}
    achieved = per_kind[dom]["achieved_gbs"]
    roofline = None if args.no_timing else {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(kernel_names[dom], args.scale),
        "kernel": kernel_names[dom],
        "bytes_per_launch": int(per_kind[dom]["algo_bytes"] / max(1, per_kind[dom]["launches"])),
        "avg_launch_us": round(per_kind[dom]["avg_launch_us"], 2),
    }

    cpu = None
    if not args.no_cpu:
        import oracle

        if not os.path.exists(oracle._LIB_PATH):
            oracle.build()
        # all 8 sources, BFS and SSSP each: 1 warm-up query of each kind, then
        # the median of 5 passes over the 16 queries on the stated cores, and
        # one pass on one thread; m_r of each query is the GPU's (the same
        # fixed point)
        m_b = sum(acc["bfs"][i][1] for i in range(len(srcs)))
        m_s = sum(acc["sssp"][i][1] for i in range(len(srcs)))

        def cpu_pass(threads):
            tb = ts = 0
            for s in srcs:
                tb += oracle.ref_loop_timed(oracle.BFS, csr.node_pointers, csr.cols, src=s, threads=threads)[1]
                ts += oracle.ref_loop_timed(oracle.SSSP, csr.node_pointers, csr.cols, csr.weights, src=s,
                                            threads=threads)[1]
            return tb, ts

        log("cpu baseline: reference loop, %d threads and 1 thread, BFS + SSSP from %d sources" % (cores, len(srcs)))
        oracle.ref_loop_timed(oracle.BFS, csr.node_pointers, csr.cols, src=srcs[0], threads=cores)
        oracle.ref_loop_timed(oracle.SSSP, csr.node_pointers, csr.cols, csr.weights, src=srcs[0], threads=cores)
        passes = sorted((cpu_pass(cores) for _ in range(5)), key=lambda x: x[0] + x[1])
        t_b, t_s = passes[2]
        t1_b, t1_s = cpu_pass(1)
        cpu = {
            "value": round((m_b + m_s) / ((t_b + t_s) / 1e9) / 1e9, 4),
            "unit": "GTEPS",
            "cores": cores,
            "kind": "port",
            "sample": "oracle orc_ref_loop (the reference's label/scan/compact/relax loop), BFS + SSSP from all %d "
                      "bench sources on the same graph, median of 5 passes after 1 warm-up: %.3fs BFS + %.3fs SSSP "
                      "on %d threads (the host's stated share: OMP_NUM_THREADS %s, %s of %s CPUs usable)"
                      % (len(srcs), t_b / 1e9, t_s / 1e9, cores, hinfo["omp_num_threads"], hinfo["affinity"],
                         hinfo["nproc"]),
            "single_thread_value": round((m_b + m_s) / ((t1_b + t1_s) / 1e9) / 1e9, 4),
            "bfs_gteps": round(m_b / (t_b / 1e9) / 1e9, 4),
            "sssp_gteps": round(m_s / (t_s / 1e9) / 1e9, 4),
            "host": hinfo,
            "host_loops": host_loops,
        }
        log("cpu baseline %.4f GTEPS (1 thread %.4f)" % (cpu["value"], cpu["single_thread_value"]))

    line = {
        "metric": "GTEPS (BFS/SSSP) & PR iters/s at RMAT-22; batched queries/s + p99 vs plain",
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT (seed 1), weights uniform [1,64] (seed 2), sources seed 3",
        "config": {
            "workload": "rmat%d-ef16: BFS x%d + SSSP x%d + CC + PageRank x%d per step" % (
                args.scale, len(srcs), len(srcs), args.pr_iters),
            "scale": args.scale,
            "edge_factor": 16,
            "V": V,
            "E": E,
            "sources": srcs,
        },
        "bfs_gteps": round(bfs_gteps, 4),
        "sssp_gteps": round(sssp_gteps, 4),
        "rounds_per_query": {k: round(tot(k, 5) / max(1, len(acc[k])), 2) for k in ("bfs", "sssp")},
        "pagerank_iters_per_s": round(pr_iters_s, 2),
        "cc_ms": round(cc_ms, 3),
        # every kind's main kernel, timed in the warmup steps
        "per_kernel_warmup": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in d.items()}
                              for k, d in per_kind_warm.items()},
        "host_output_gteps": pcie,
        "tuning": tuned,
        "batched": batched,
        "serving": serving,
        "partition": partition,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.free()
    g.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
