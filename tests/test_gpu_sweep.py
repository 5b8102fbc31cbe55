"""SSSP sweep rounds (ptg_config.sssp_sweep) through the HIP path.

Where a round's push and pull would both scan more than E x sssp_sweep edges,
the round relaxes every edge through the binned layout instead: each
vertex's value goes to the bins of its (source, destination partition)
pairs, each destination partition takes the minimum of bin + weight over its
entries in LDS, and a vertex whose minimum is below its value takes it and
joins the next frontier. The distances must equal the oracle's Dijkstra bit
for bit whether sweeps never happen (-1), follow the default rule (0) or
happen in every large round (1e-9), on:
  * R-MAT-16 with one weight per undirected pair, with and without
    relabeling and deferral, and with pulls preferred (no sweeps then);
  * R-MAT-14 with each direction of an edge weighted on its own and some
    edges repeated (a sweep relaxes the out-edges with their own weights);
  * zero weights with the 0xFFFFFFFF sentinel;
  * weights above 255 (no weight byte: sweeps refused).
The layout is shared with binned PageRank, whose ranks must keep their bits
when SSSP sweeps and PageRank iterations take turns on one graph."""
import os
import sys

import numpy as np
import pytest

import oracle
import ptgraph as pg

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    pg.init(0)
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()


def sources(off, n, seed):
    deg = np.diff(np.asarray(off, dtype=np.int64))
    nz = np.nonzero(deg > 0)[0]
    return [int(np.argmax(deg))] + [int(x) for x in np.random.default_rng(seed).choice(nz, n - 1, replace=False)]


def run_settings(g, srcs, want, label, settings, **tune):
    """Every source under every sweep setting; returns swept rounds per setting."""
    swept = {}
    for x in settings:
        g.tune(sssp_sweep=x, **tune)
        total = 0
        for s in srcs:
            say("case %s sssp_sweep %g src %d" % (label, x, s))
            d, st = g.sssp(s)
            np.testing.assert_array_equal(d, want[s], err_msg="%s sssp_sweep %g src %d" % (label, x, s))
            total += st.sweep_rounds
        swept[x] = total
    say("%s swept rounds %s" % (label, swept))
    if -1.0 in swept:
        assert swept[-1.0] == 0
    return swept


@pytest.mark.parametrize("flags", [0, pg.FLAG_NO_RELABEL], ids=["relabel", "no_relabel"])
def test_sweep_rmat16_pair_weights(flags):
    csr = pg.gen_rmat(16, 16, seed=1).add_weights(1, 64, seed=2)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    srcs = sources(csr.offsets, 4, 31)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w, config=pg.default_config(flags=flags)) as g:
        assert g.is_symmetric()
        for ddeg in (0, 0xFFFFFFFF):  # deferral by default / off
            swept = run_settings(g, srcs, want, "rmat16 flags %d defer %d" % (flags, ddeg), (-1.0, 0.0, 1e-9, 0.1),
                                 sssp_defer_degree=ddeg)
            assert swept[1e-9] > 0
        # pulls preferred in every round: a sweep never pays
        swept = run_settings(g, srcs, want, "rmat16 pulls preferred", (1e-9,), sssp_pull=1e-9)
        assert swept[1e-9] == 0
        # pull kernels off: no in-pairs, so no sweeps either
        swept = run_settings(g, srcs, want, "rmat16 pulls off", (1e-9,), sssp_pull=-1.0)
        assert swept[1e-9] == 0


def directed_weights_with_repeats(scale, seed):
    """Symmetric R-MAT with 5% of the undirected pairs repeated (both
    directions again) and an independent weight in [1, 64] per directed edge."""
    csr = pg.gen_rmat(scale, 16, seed=1)
    off = csr.offsets.astype(np.int64)
    V = off.size - 1
    src = np.repeat(np.arange(V), np.diff(off))
    dst = csr.cols.astype(np.int64)
    rng = np.random.default_rng(seed)
    fwd = np.nonzero((src < dst) & (rng.random(src.size) < 0.05))[0]
    a = np.concatenate([src, src[fwd], dst[fwd]])
    b = np.concatenate([dst, dst[fwd], src[fwd]])
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    ptr = np.searchsorted(a, np.arange(V)).astype(np.uint32)
    w = rng.integers(1, 65, a.size).astype(np.uint32)
    return ptr, b.astype(np.uint32), w


def test_sweep_directed_weights_and_repeats():
    ptr, col, w = directed_weights_with_repeats(14, 6)
    srcs = sources(np.append(ptr, col.size), 4, 32)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w) as g:
        assert g.is_symmetric()
        swept = run_settings(g, srcs, want, "rmat14 directed weights", (-1.0, 0.0, 1e-9))
        assert swept[1e-9] > 0


def test_sweep_zero_weights_opt_sentinel():
    csr = pg.gen_rmat(15, 16, seed=3).add_weights(0, 4, seed=4)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    srcs = sources(csr.offsets, 3, 33)
    want = {s: oracle.dijkstra(ptr, col, w, s, oracle.SSSP_SENTINEL_OPT) for s in srcs}
    cfg = pg.default_config(sssp_sentinel=pg.SSSP_SENTINEL_OPT)
    with pg.Graph(ptr, col, w, config=cfg) as g:
        swept = run_settings(g, srcs, want, "rmat15 weights [0, 4] opt sentinel", (-1.0, 0.0, 1e-9))
        assert swept[1e-9] > 0


def test_sweep_refused_above_byte_weights():
    csr = pg.gen_rmat(14, 16, seed=3).add_weights(1, 1000, seed=4)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    assert int(w.max()) > 255
    srcs = sources(csr.offsets, 2, 34)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w) as g:
        swept = run_settings(g, srcs, want, "rmat14 weights up to 1000", (-1.0, 1e-9))
        assert swept[1e-9] == 0


def test_sweeps_and_binned_pagerank_share_the_layout():
    """SSSP sweeps write their bins and rows into the binned layout's
    buffers; PageRank on the same graph (binned, before and after) keeps the
    in-list pull's bits, and the distances stay Dijkstra's."""
    csr = pg.gen_rmat(15, 16, seed=5).add_weights(1, 64, seed=6)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    srcs = sources(csr.offsets, 3, 35)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w, config=pg.default_config(pr_layout=pg.PR_PULL)) as g:
        say("case shared layout: pull PageRank")
        pull_ranks, _ = g.pagerank(0.85, 20)
    with pg.Graph(ptr, col, w, config=pg.default_config(pr_layout=pg.PR_BINNED, sssp_sweep=1e-9)) as g:
        say("case shared layout: binned PageRank first")
        r, _ = g.pagerank(0.85, 20)
        np.testing.assert_array_equal(r.view(np.uint32), pull_ranks.view(np.uint32))
        for s in srcs:
            say("case shared layout: sweep src %d" % s)
            d, st = g.sssp(s)
            np.testing.assert_array_equal(d, want[s])
            say("case shared layout: PageRank after src %d" % s)
            r, _ = g.pagerank(0.85, 20)
            np.testing.assert_array_equal(r.view(np.uint32), pull_ranks.view(np.uint32))
        assert g.sssp(srcs[0])[1].sweep_rounds > 0


def test_sweep_refused_when_pairs_outnumber_half_the_edges():
    """A uniform random graph spreads each list over many partitions: the
    (source, partition) pairs exceed half the edges, the layout is refused
    once (and not rebuilt by later queries), nothing sweeps, and PageRank
    takes the in-list pull; the distances stay Dijkstra's."""
    rng = np.random.default_rng(9)
    V, M = 20 * 16384, 1300000
    a, b = rng.integers(0, V, M), rng.integers(0, V, M)
    keep = a != b
    a, b = np.concatenate([a[keep], b[keep]]), np.concatenate([b[keep], a[keep]])
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    ptr = np.searchsorted(a, np.arange(V)).astype(np.uint32)
    col = b.astype(np.uint32)
    w = np.empty(col.size, np.uint32)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    w[:] = (lo * 2654435761 + hi * 40503) % 64 + 1  # one weight per undirected pair
    srcs = sources(np.append(ptr, col.size), 2, 36)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w, config=pg.default_config(sssp_sweep=1e-9)) as g:
        assert g.is_symmetric()
        for rep in range(2):
            for s in srcs:
                say("case refused layout rep %d src %d" % (rep, s))
                d, st = g.sssp(s)
                np.testing.assert_array_equal(d, want[s])
                assert st.sweep_rounds == 0
        say("case refused layout: PageRank")
        auto, _ = g.pagerank(0.85, 10)
        g.tune(pr_layout=pg.PR_PULL)
        pull, _ = g.pagerank(0.85, 10)
        np.testing.assert_array_equal(auto.view(np.uint32), pull.view(np.uint32))


def test_layout_gives_way_under_a_device_budget():
    """The binned layout is optional: under a device budget with no room for
    it, SSSP runs without sweeps and PageRank (automatic layout) by the
    in-list pull, with the same results; without the budget the same graph
    sweeps."""
    csr = pg.gen_rmat(15, 16, seed=7).add_weights(1, 64, seed=8)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    srcs = sources(csr.offsets, 2, 37)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    want_pr = oracle.pagerank_f64(ptr, col, 0.85, 10)
    # the device bytes of everything the queries need but the binned layout
    with pg.Graph(ptr, col, w, config=pg.default_config(sssp_sweep=-1.0, pr_layout=pg.PR_PULL)) as g:
        for s in srcs:
            g.sssp(s)
        g.pagerank(0.85, 10)
        base = g.device_bytes
    budget = base + (1 << 20)
    with pg.Graph(ptr, col, w, config=pg.default_config(sssp_sweep=1e-9, device_budget_bytes=budget)) as g:
        for s in srcs:
            say("case tight budget sssp src %d" % s)
            d, st = g.sssp(s)
            np.testing.assert_array_equal(d, want[s])
            assert st.sweep_rounds == 0
        say("case tight budget pagerank")
        r, _ = g.pagerank(0.85, 10)
        assert float(np.abs(r.astype(np.float64) - want_pr).sum()) <= 1e-6
        assert g.device_bytes <= budget
    with pg.Graph(ptr, col, w, config=pg.default_config(sssp_sweep=1e-9)) as g:
        say("case no budget sweeps")
        d, st = g.sssp(srcs[0])
        np.testing.assert_array_equal(d, want[srcs[0]])
        assert st.sweep_rounds > 0
