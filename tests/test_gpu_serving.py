"""GPU tests of the serving loop (ptg_server_*): every result served in a mix
of concurrent, batched and coalesced requests equals the oracle's (BFS queue
levels, Dijkstra distances and reach-min labels bit for bit, PageRank within
1e-6 L1 of the FP64 oracle) and the same query run alone (bit for bit,
PageRank included: its sums run in a fixed order), with memory admission
evicting, waiting and rejecting as configured."""
import os
import sys

import numpy as np
import pytest

import oracle
import ptgraph as pg

pytestmark = pytest.mark.gpu

MIX = ((pg.Q_BFS, 0.50), (pg.Q_SSSP, 0.30), (pg.Q_CC, 0.15), (pg.Q_PAGERANK, 0.05))
PR_L1 = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    pg.init(0)
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()


@pytest.fixture(scope="module")
def csrs():
    return [pg.gen_rmat(10, 16, seed=1).add_weights(1, 64, seed=2),
            pg.gen_rmat(12, 8, seed=5).add_weights(1, 64, seed=2),
            pg.gen_er(11, 8, seed=9).add_weights(1, 64, seed=2)]


def make_requests(csrs, n, rng, out_bufs):
    kinds = rng.choice([k for k, _ in MIX], n, p=[p for _, p in MIX])
    reqs = []
    for i in range(n):
        gid = int(rng.integers(len(csrs)))
        V = csrs[gid].V
        src = int(rng.integers(V))
        out = None
        if out_bufs is not None:
            out_bufs.append(np.zeros(V, np.uint32))
            out = out_bufs[-1].ctypes.data
        reqs.append(pg.Server.request(gid, int(kinds[i]), source=src, iterations=10, damping=0.85, out=out, tag=i))
    return reqs


def alone(graphs, r):
    g = graphs[r.graph]
    if r.kind == pg.Q_BFS:
        return g.bfs(r.source)[0]
    if r.kind == pg.Q_SSSP:
        return g.sssp(r.source)[0]
    if r.kind == pg.Q_CC:
        return g.cc()[0]
    return g.pagerank(r.damping, r.iterations)[0].view(np.uint32)


def against_oracle(csr, r, got):
    """The served (or alone) result of request `r` against the oracle."""
    ptr, col = csr.node_pointers, csr.cols
    if r.kind == pg.Q_BFS:
        np.testing.assert_array_equal(got, oracle.bfs(ptr, col, r.source)[0])
    elif r.kind == pg.Q_SSSP:
        np.testing.assert_array_equal(got, oracle.dijkstra(ptr, col, csr.weights, r.source))
    elif r.kind == pg.Q_CC:
        np.testing.assert_array_equal(got, oracle.cc_reach_min(ptr, col))
    else:
        want = oracle.pagerank_f64(ptr, col, r.damping, r.iterations)
        assert float(np.abs(got.view(np.float32).astype(np.float64) - want).sum()) <= PR_L1


def check(reqs, comps, outs, ref_graphs, srv=None, csrs=None):
    assert len(comps) == len(reqs)
    if srv is not None:
        assert srv.error() == "", srv.error()
    by_tag = {c.tag: c for c in comps}
    cache = {}
    for r in reqs:
        c = by_tag[r.tag]
        assert c.status == pg.OK, (r.tag, c.status)
        key = (r.graph, r.kind, r.source if r.kind in (pg.Q_BFS, pg.Q_SSSP) else 0)
        if key not in cache:
            cache[key] = alone(ref_graphs, r)
            if csrs is not None:
                against_oracle(csrs[r.graph], r, cache[key])
        want = cache[key]
        assert c.checksum == pg.checksum(want), "checksum tag %d kind %d" % (r.tag, r.kind)
        if outs is not None:
            np.testing.assert_array_equal(outs[r.tag], want, err_msg="tag %d kind %d" % (r.tag, r.kind))
            if csrs is not None and r.kind != pg.Q_PAGERANK:
                against_oracle(csrs[r.graph], r, outs[r.tag])
        assert c.submit_ns <= c.start_ns <= c.done_ns


def test_mix_equals_alone_burst_and_replay(csrs):
    rng = np.random.default_rng(4)
    ref = [pg.Graph.from_csr(c) for c in csrs]
    graphs = [pg.Graph.from_csr(c) for c in csrs]
    try:
        with pg.Server() as srv:
            for g in graphs:
                srv.add_graph(g)
            say("case serving burst 400")
            outs = []
            reqs = make_requests(csrs, 400, rng, outs)
            srv.pause()
            for r in reqs:
                srv.submit(r)
            srv.pause(False)
            srv.drain()
            comps = srv.completions()
            check(reqs, comps, outs, ref, srv, csrs)
            s = srv.summary()
            assert s.completed == 400 and s.rejected == 0
            assert s.mean_batch > 1.5  # a burst is served in batches
            assert max(c.batch_size for c in comps) > 8
            say("case serving replay 300")
            outs2 = []
            reqs2 = make_requests(csrs, 300, rng, outs2)
            for r in reqs2:
                r.tag += 1000
            outs_by_tag = {r.tag: o for r, o in zip(reqs2, outs2)}
            arrivals = np.cumsum(rng.exponential(1e9 / 20000, 300)).astype(np.uint64)
            srv.replay(reqs2, arrivals)
            srv.drain()
            check(reqs2, srv.completions(), outs_by_tag, ref, srv, csrs)
            assert srv.summary().p99_ns >= srv.summary().p50_ns > 0
    finally:
        for g in ref + graphs:
            g.close()


def test_plain_mode_serves_one_at_a_time(csrs):
    rng = np.random.default_rng(6)
    ref = [pg.Graph.from_csr(c) for c in csrs]
    graphs = [pg.Graph.from_csr(c) for c in csrs]
    try:
        with pg.Server(max_batch=1, max_inflight=1, coalesce=False) as srv:
            for g in graphs:
                srv.add_graph(g)
            say("case serving plain 120")
            reqs = make_requests(csrs, 120, rng, None)  # checksum only
            for r in reqs:
                srv.submit(r)
            srv.drain()
            comps = srv.completions()
            assert all(c.batch_size == 1 for c in comps)
            check(reqs, comps, None, ref, srv, csrs)
            # one at a time: launch intervals do not overlap
            iv = sorted((c.start_ns, c.done_ns) for c in comps)
            assert all(b[0] >= a[1] for a, b in zip(iv, iv[1:]))
    finally:
        for g in ref + graphs:
            g.close()


def test_memory_admission_evicts_and_rejects(csrs):
    rng = np.random.default_rng(8)
    ref = [pg.Graph.from_csr(c) for c in csrs]
    graphs = [pg.Graph.from_csr(c) for c in csrs]
    try:
        # room for one graph's batched state at a time: batches on different
        # graphs must release each other's state
        biggest = max(c.V for c in csrs)
        budget = 64 * 4 * biggest * 2 + 32 * 8 * biggest + 40 * biggest + (1 << 20)
        with pg.Server(state_budget_bytes=budget, min_batch=2) as srv:
            for g in graphs:
                srv.add_graph(g)
            say("case serving admission budget %d" % budget)
            reqs = []
            for i in range(240):
                gid = i % 3
                kind = pg.Q_BFS if (i // 3) % 2 == 0 else pg.Q_SSSP
                reqs.append(pg.Server.request(gid, kind, source=int(rng.integers(csrs[gid].V)), tag=i))
            srv.pause()  # queue everything first so launches are batches
            for r in reqs:
                srv.submit(r)
            srv.pause(False)
            srv.drain()
            comps = srv.completions()
            check(reqs, comps, None, ref, srv, csrs)
            s = srv.summary()
            info = "launches %d batch %.1f peak %d evictions %d" % (s.launches, s.mean_batch, s.state_bytes_peak,
                                                                   s.evictions)
            say("case serving admission summary: " + info)
            assert s.evictions > 0, info
            assert s.state_bytes_peak <= budget, info
        # a budget no batch fits in: batched launches are rejected, single
        # queries (no extra state) still run
        with pg.Server(state_budget_bytes=4096, min_batch=2) as srv:
            gid = srv.add_graph(graphs[0])
            say("case serving admission reject")
            reqs = [pg.Server.request(gid, pg.Q_BFS, source=s, tag=s) for s in range(20)]
            srv.pause()
            srv.submit(reqs[0])
            srv.pause(False)
            srv.drain()  # alone: a single query, no batch state
            srv.pause()
            for r in reqs[1:]:
                srv.submit(r)
            srv.pause(False)
            srv.drain()
            comps = srv.completions()
            st = {c.tag: c.status for c in comps}
            assert len(st) == 20
            assert st[0] == pg.OK and all(st[t] == pg.ERR_REJECTED for t in range(1, 20))
            for c in comps:
                if c.status == pg.OK:
                    assert c.batch_size == 1
                    assert c.checksum == pg.checksum(ref[0].bfs(int(c.tag))[0])
                else:
                    assert c.status == pg.ERR_REJECTED and c.batch_size > 1
            assert srv.summary().rejected == sum(v != pg.OK for v in st.values())
    finally:
        for g in ref + graphs:
            g.close()


def test_submit_errors(csrs):
    g = pg.Graph.from_csr(csrs[0])
    unweighted = pg.Graph.from_csr(csrs[1], weighted=False)
    try:
        with pg.Server(queue_cap=2, max_inflight=1) as srv:
            gid = srv.add_graph(g)
            uid = srv.add_graph(unweighted)
            say("case serving submit errors")
            for bad, code in ((pg.Server.request(7, pg.Q_BFS), pg.ERR_INVALID_ARG),
                              (pg.Server.request(gid, 9), pg.ERR_INVALID_ARG),
                              (pg.Server.request(gid, pg.Q_BFS, source=csrs[0].V), pg.ERR_INVALID_ARG),
                              (pg.Server.request(uid, pg.Q_SSSP), pg.ERR_UNSUPPORTED),
                              (pg.Server.request(gid, pg.Q_PAGERANK, damping=2.0), pg.ERR_INVALID_ARG)):
                with pytest.raises(pg.PtgError) as e:
                    srv.submit(bad)
                assert e.value.code == code
            srv.pause()
            srv.submit(pg.Server.request(gid, pg.Q_BFS, source=1, tag=1))
            srv.submit(pg.Server.request(gid, pg.Q_BFS, source=2, tag=2))
            with pytest.raises(pg.PtgError) as e:
                srv.submit(pg.Server.request(gid, pg.Q_BFS, source=3, tag=3))  # queue_cap 2
            assert e.value.code == pg.ERR_REJECTED
            with pytest.raises(pg.PtgError):
                srv.drain()  # paused
            srv.pause(False)
            srv.drain()
            comps = srv.completions()
            assert sorted(c.tag for c in comps) == [1, 2] and all(c.status == pg.OK for c in comps)
    finally:
        g.close()
        unweighted.close()
