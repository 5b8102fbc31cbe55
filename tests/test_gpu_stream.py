"""GPU parity of streamed graphs (PTG_FLAG_STREAM): edges (and weights) in
pinned host memory, a device chunk cache under a budget. BFS levels, SSSP
distances and CC labels must equal the oracle's bit for bit and PageRank must
stay within 1e-6 L1 of the FP64 oracle and equal the resident graph's ranks bit
for bit, whatever the chunk size and however many chunks are evicted."""
import os
import sys

import numpy as np
import pytest

import oracle
import ptgraph as pg

pytestmark = pytest.mark.gpu
PR_L1 = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    pg.init(0)
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()


def star_plus(rng, leaves, extra_v, extra_e):
    """Vertex 0 joined to `leaves` vertices plus a sparse random part, symmetric."""
    V = 1 + leaves + extra_v
    src = [np.zeros(leaves, np.int64), np.arange(1, leaves + 1)]
    dst = [np.arange(1, leaves + 1), np.zeros(leaves, np.int64)]
    a = rng.integers(1, V, extra_e)
    b = rng.integers(1, V, extra_e)
    keep = a != b
    src += [a[keep], b[keep]]
    dst += [b[keep], a[keep]]
    s, d = np.concatenate(src), np.concatenate(dst)
    pairs = np.unique(s * V + d)
    s, d = pairs // V, pairs % V
    ptr = np.searchsorted(s, np.arange(V)).astype(np.uint32)
    return ptr, d.astype(np.uint32)


CASES = {
    # chunk bytes, budget bytes (0 = unlimited)
    "one_chunk": (256 << 20, 0),
    "small_chunks": (16 << 10, 0),
    "tiny_cache": (16 << 10, 1),       # budget: state + two slots
    "split_hub": (8 << 10, 1),         # the hub is larger than a chunk
}


def state_budget(ptr, col, chunk, flags):
    """A device budget that leaves exactly two chunk slots."""
    cfg = pg.default_config(flags=pg.FLAG_STREAM | flags, chunk_bytes=chunk)
    with pg.Graph(ptr, col, config=cfg) as g:
        nch, nslots, slot = g.stream_info()
        assert nslots == nch  # unlimited budget: a slot per chunk
        return g.device_bytes - nslots * slot + 2 * slot


@pytest.mark.parametrize("case", list(CASES))
def test_streamed_rmat_bfs_pagerank(case):
    chunk, small = CASES[case]
    if case == "split_hub":
        rng = np.random.default_rng(2)
        ptr, col = star_plus(rng, 6000, 3000, 20000)
    else:
        csr = pg.gen_rmat(12, 16, seed=1)
        ptr, col = csr.node_pointers, csr.cols
    say("case streamed %s V=%d E=%d" % (case, ptr.size, col.size))
    with pg.Graph(ptr, col) as rg:
        resident_pr = {it: rg.pagerank(0.85, it)[0] for it in (0, 1, 10)}
    cfg = pg.default_config(flags=pg.FLAG_STREAM, chunk_bytes=chunk)
    if small:
        cfg.device_budget_bytes = state_budget(ptr, col, chunk, 0)
    with pg.Graph(ptr, col, config=cfg) as g:
        assert g.is_symmetric()
        nch, nslots, _ = g.stream_info()
        say("case streamed %s: %d chunks, %d slots" % (case, nch, nslots))
        if small:
            assert nslots == 2 and nch > 2
        deg = np.diff(np.append(ptr.astype(np.int64), col.size))
        srcs = [0, int(np.argmax(deg)), int(np.nonzero(deg == 0)[0][0]) if (deg == 0).any() else 1, ptr.size - 1]
        loads = 0
        for s in srcs:
            lv, st = g.bfs(s)
            want = oracle.bfs(ptr, col, s)[0]
            np.testing.assert_array_equal(lv, want, err_msg="bfs src %d" % s)
            reached = want != 0xFFFFFFFF
            assert st.reached == int(reached.sum())
            assert st.iterations == int(want[reached].max())
            assert st.edges_traversed == int(deg[reached].sum())
            loads += st.chunk_loads
        for iters in (0, 1, 10):
            r, st = g.pagerank(0.85, iters)
            l1 = float(np.abs(r.astype(np.float64) - oracle.pagerank_f64(ptr, col, 0.85, iters)).sum())
            assert l1 <= PR_L1, (iters, l1)
            # exact sums: the streamed ranks are the resident ranks, bit for bit
            np.testing.assert_array_equal(r.view(np.uint32), resident_pr[iters].view(np.uint32),
                                          err_msg="streamed vs resident pagerank, %d iterations" % iters)
        r2, st2 = g.pagerank(0.85, 10)
        assert np.array_equal(r.view(np.uint32), r2.view(np.uint32))  # deterministic
        assert loads > 0  # the first query copies its chunks in
        if small:
            # a 2-slot cache copies nearly every chunk again in every iteration
            assert st2.chunk_loads >= 10 * (nch - 2)
            assert st2.host_link_bytes > 0 and st2.host_link_ns > 0
        else:
            assert st2.chunk_loads == 0 and st2.chunk_hits == 10 * nch  # everything stayed resident


def test_streamed_no_relabel_and_directed():
    rng = np.random.default_rng(5)
    ptr, col = star_plus(rng, 5000, 2000, 8000)
    say("case streamed no_relabel")
    cfg = pg.default_config(flags=pg.FLAG_STREAM | pg.FLAG_NO_RELABEL, chunk_bytes=8 << 10)
    with pg.Graph(ptr, col, config=cfg) as g:
        for s in (0, 17, ptr.size - 1):
            np.testing.assert_array_equal(g.bfs(s)[0], oracle.bfs(ptr, col, s)[0])
        r, _ = g.pagerank(0.85, 10)
        assert float(np.abs(r.astype(np.float64) - oracle.pagerank_f64(ptr, col, 0.85, 10)).sum()) <= PR_L1
    say("case streamed directed")
    V, E = 3000, 12000
    a, b = rng.integers(0, V, E), rng.integers(0, V, E)
    o = np.lexsort((b, a))
    a, b = a[o], b[o]
    dptr = np.searchsorted(a, np.arange(V)).astype(np.uint32)
    dcol = b.astype(np.uint32)
    with pg.Graph(dptr, dcol, config=pg.default_config(flags=pg.FLAG_STREAM, chunk_bytes=4 << 10)) as g:
        assert not g.is_symmetric()
        for s in (0, 99):
            np.testing.assert_array_equal(g.bfs(s)[0], oracle.bfs(dptr, dcol, s)[0])
        np.testing.assert_array_equal(g.cc()[0], oracle.cc_reach_min(dptr, dcol))  # label propagation rounds
        for fn in (lambda: g.pagerank(0.85, 5), lambda: g.bfs_batch(np.array([0, 1], np.uint32))):
            with pytest.raises(pg.PtgError) as e:
                fn()
            assert e.value.code == pg.ERR_UNSUPPORTED


def test_streamed_budget_and_probe():
    csr = pg.gen_rmat(10, 16, seed=1)
    say("case streamed budget")
    cfg = pg.default_config(flags=pg.FLAG_STREAM, chunk_bytes=4 << 10, device_budget_bytes=64 << 10)
    with pytest.raises(pg.PtgError) as e:
        pg.Graph(csr.node_pointers, csr.cols, config=cfg)
    assert e.value.code == pg.ERR_BUDGET
    gbs = pg.host_link_probe(64 << 20, 4)
    say("case host link %.1f GB/s" % gbs)
    assert gbs > 1.0


def weighted_case(case):
    if case == "split_hub":
        rng = np.random.default_rng(2)
        ptr, col = star_plus(rng, 6000, 3000, 20000)
        E = col.size
        nxt = np.append(ptr[1:], E).astype(np.int64)
        src = np.repeat(np.arange(ptr.size), nxt - ptr.astype(np.int64))
        lo, hi = np.minimum(src, col), np.maximum(src, col)
        w = ((lo * 2654435761 + hi * 40503) % 64 + 1).astype(np.uint32)  # symmetric weights in [1, 64]
        return ptr, col, w
    csr = pg.gen_rmat(12, 16, seed=1).add_weights(1, 64, seed=2)
    # copies: the arrays are views of memory the HostCsr frees with itself
    return csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()


@pytest.mark.parametrize("case", list(CASES))
def test_streamed_sssp_and_cc(case):
    chunk, small = CASES[case]
    ptr, col, w = weighted_case(case)
    say("case streamed sssp/cc %s V=%d E=%d" % (case, ptr.size, col.size))
    cfg = pg.default_config(flags=pg.FLAG_STREAM, chunk_bytes=chunk)
    if small:
        with pg.Graph(ptr, col, w, config=cfg) as g:
            nch, nslots, slot = g.stream_info()
            cfg.device_budget_bytes = g.device_bytes - nslots * slot + 2 * slot
    with pg.Graph(ptr, col, w, config=cfg) as g:
        nch, nslots, _ = g.stream_info()
        if small:
            assert nslots == 2 and nch > 2
        deg = np.diff(np.append(ptr.astype(np.int64), col.size))
        for s in (0, int(np.argmax(deg)), ptr.size - 1):
            d, st = g.sssp(s)
            want = oracle.dijkstra(ptr, col, w, s)
            np.testing.assert_array_equal(d, want, err_msg="sssp src %d" % s)
            reached = want != oracle.SSSP_SENTINEL_SHARE
            assert st.reached == int(reached.sum())
            assert st.edges_traversed == int(deg[reached].sum())
            np.testing.assert_array_equal(g.bfs(s)[0], oracle.bfs(ptr, col, s)[0])  # columns only
        lab, st = g.cc()
        want = oracle.cc_union_find(ptr, col)
        np.testing.assert_array_equal(lab, want)
        assert st.reached == int(np.unique(want).size)
        r, _ = g.pagerank(0.85, 10)
        assert float(np.abs(r.astype(np.float64) - oracle.pagerank_f64(ptr, col, 0.85, 10)).sum()) <= PR_L1


def test_streamed_aos_opt_sentinel_and_no_relabel():
    ptr, col, w = weighted_case("small_chunks")
    aos = np.stack([col, w], axis=1)
    say("case streamed aos opt sentinel")
    for flags in (pg.FLAG_STREAM, pg.FLAG_STREAM | pg.FLAG_NO_RELABEL):
        cfg = pg.default_config(flags=flags, chunk_bytes=16 << 10, sssp_sentinel=oracle.SSSP_SENTINEL_OPT)
        with pg.Graph.from_aos(ptr, aos, config=cfg) as g:
            for s in (0, 5, 4000):
                np.testing.assert_array_equal(g.sssp(s)[0], oracle.dijkstra(ptr, col, w, s, oracle.SSSP_SENTINEL_OPT))
            np.testing.assert_array_equal(g.cc()[0], oracle.cc_union_find(ptr, col))


def test_streamed_directed_sssp():
    rng = np.random.default_rng(11)
    V, E = 4000, 20000
    a, b = rng.integers(0, V, E), rng.integers(0, V, E)
    o = np.lexsort((b, a))
    a, b = a[o], b[o]
    dptr = np.searchsorted(a, np.arange(V)).astype(np.uint32)
    dcol = b.astype(np.uint32)
    dw = rng.integers(0, 9, E).astype(np.uint32)  # zero weights included
    say("case streamed directed sssp")
    cfg = pg.default_config(flags=pg.FLAG_STREAM, chunk_bytes=4 << 10)
    with pg.Graph(dptr, dcol, dw, config=cfg) as g:
        cfg.device_budget_bytes = g.device_bytes - g.stream_info()[1] * g.stream_info()[2] + 2 * g.stream_info()[2]
    with pg.Graph(dptr, dcol, dw, config=cfg) as g:
        assert g.stream_info()[1] == 2
        for s in (0, 17, V - 1):
            np.testing.assert_array_equal(g.sssp(s)[0], oracle.dijkstra(dptr, dcol, dw, s))
        np.testing.assert_array_equal(g.cc()[0], oracle.cc_reach_min(dptr, dcol))


def test_streamed_replan_pinned_and_autotune():
    """Re-cutting the chunks (set_config chunk_bytes), pinning the leading
    chunks and the streamed autotuner change no result; a chunk size the
    budget cannot hold two slots of is refused and the old plan kept."""
    ptr, col, w = weighted_case("small_chunks")
    say("case streamed replan")
    cfg = pg.default_config(flags=pg.FLAG_STREAM, chunk_bytes=16 << 10)
    with pg.Graph(ptr, col, w, config=cfg) as g:
        nch, nslots, slot = g.stream_info()
        cfg.device_budget_bytes = g.device_bytes - nslots * slot + 4 * slot  # four slots
    want_b = oracle.bfs(ptr, col, 0)[0]
    want_d = oracle.dijkstra(ptr, col, w, 5)
    want_r = oracle.pagerank_f64(ptr, col, 0.85, 10)
    with pg.Graph(ptr, col, w, config=cfg) as g:
        nch0, nsl0, _ = g.stream_info()
        assert nsl0 == 4 and nch0 > 4

        def check():
            np.testing.assert_array_equal(g.bfs(0)[0], want_b)
            np.testing.assert_array_equal(g.sssp(5)[0], want_d)
            r, _ = g.pagerank(0.85, 10)
            assert float(np.abs(r.astype(np.float64) - want_r).sum()) <= PR_L1

        check()
        g.tune(chunk_bytes=32 << 10)  # larger chunks: fewer of them, fewer (larger) slots in the same budget
        nch1, nsl1, _ = g.stream_info()
        assert nch1 < nch0 and 2 <= nsl1 < nch1
        check()
        g.tune(stream_pinned=1)
        _, st = g.pagerank(0.85, 3)
        check()
        with pytest.raises(pg.PtgError) as e:
            g.tune(chunk_bytes=1 << 30)  # one chunk of every edge: a weighted slot pair no longer fits
        assert e.value.code == pg.ERR_BUDGET
        assert g.stream_info()[:2] == (nch1, nsl1) and g.get_config().chunk_bytes == 32 << 10
        check()
        c = g.autotune(np.array([0, 5], np.uint32), what=pg.TUNE_STREAM)
        assert c.chunk_bytes in [x << 20 for x in (64, 128, 256, 512, 1024)] + [32 << 10]
        check()
        with pytest.raises(pg.PtgError) as e:
            g.autotune(np.array([0], np.uint32), what=pg.TUNE_BFS)
        assert e.value.code == pg.ERR_UNSUPPORTED
