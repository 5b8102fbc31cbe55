"""GPU parity: the HIP path through libptgraph.so against the oracle and the
golden fixtures. BFS levels, SSSP distances and CC labels are compared
bit-exactly; PageRank (FP32) must be within 1e-6 L1 of the FP64 oracle."""
import glob
import hashlib
import os
import sys

import numpy as np
import pytest

import oracle
import ptgraph as pg

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "hand_*.npz"))) + [
    os.path.join(GOLDEN, "rmat8.npz"), os.path.join(GOLDEN, "rmat10.npz")]
PR_L1 = 1e-6  # FP32 PageRank vs FP64 oracle, L1 over all vertices

CONFIGS = {
    "default": {},
    "tiny_bins": {"small_degree": 1, "large_degree": 64},
    "bottom_up": {"td_to_bu": 1e9, "bu_to_td": 1e9},
    "top_down": {"td_to_bu": 1e-9},
    "one_block": {"grid_blocks": 1},
    "odd_grid": {"grid_blocks": 7, "small_degree": 3, "large_degree": 64},
    "no_relabel": {"flags": pg.FLAG_NO_RELABEL},
    "no_relabel_bu": {"flags": pg.FLAG_NO_RELABEL, "td_to_bu": 1e9, "bu_to_td": 1e9, "small_degree": 2},
    "degree_order": {"order_hub_degree": 0},
    "hub_order_big": {"order_hub_degree": 3},
    "delta1": {"sssp_delta": 1},      # light = zero-weight edges only
    "delta1000": {"sssp_delta": 1000},  # every edge light
    # SSSP deferral: every vertex of degree > 1 waits for a window of 1
    "defer_all": {"small_degree": 1, "large_degree": 64, "sssp_defer_degree": 1, "sssp_defer_window": 1},
    "defer_off": {"sssp_defer_degree": 0xFFFFFFFF},
    "defer_split": {"sssp_delta": 24, "sssp_defer_degree": 8, "sssp_defer_window": 1},
    # SSSP pull rounds in every large round / never
    "pull_always": {"sssp_pull": 1e-9},
    "pull_never": {"sssp_pull": -1.0},
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    pg.init(0)
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()


def ptr_degrees(ptr, E):
    """Out-degrees from V reference-layout offsets (the last one runs to E)."""
    p = np.append(np.asarray(ptr, dtype=np.int64), E)
    return np.diff(p)


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def pr_l1(ptr, col, got, iters=20):
    want = oracle.pagerank_f64(ptr, col, 0.85, iters)
    return float(np.abs(got.astype(np.float64) - want).sum())


def sort_rows(f):
    # same graph with every adjacency list sorted (enables the symmetric-graph
    # paths: bottom-up BFS and union-find CC)
    ptr, col, w = f["ptr"], f["col"], f["w"]
    E = col.size
    nxt = np.append(ptr[1:], E)
    col2, w2 = col.copy(), w.copy()
    for u in range(ptr.size):
        o = np.argsort(col[ptr[u]:nxt[u]], kind="stable")
        col2[ptr[u]:nxt[u]] = col[ptr[u]:nxt[u]][o]
        w2[ptr[u]:nxt[u]] = w[ptr[u]:nxt[u]][o]
    g = dict(f)
    g["col"], g["w"] = col2, w2
    return g


@pytest.mark.parametrize("rows", ["as_given", "sorted"])
@pytest.mark.parametrize("cfg", list(CONFIGS), ids=list(CONFIGS))
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_fixture_parity(path, cfg, rows):
    f = load(path)
    if rows == "sorted":
        f = sort_rows(f)
    say("case %s cfg=%s rows=%s V=%d E=%d" % (os.path.basename(path), cfg, rows, f["ptr"].size, f["col"].size))
    with pg.Graph(f["ptr"], f["col"], f["w"], config=pg.default_config(**CONFIGS[cfg])) as g:
        for i, s in enumerate(f["srcs"]):
            lv, st = g.bfs(int(s))
            np.testing.assert_array_equal(lv, f["bfs"][i], err_msg="bfs src %d" % s)
            reached = f["bfs"][i] != 0xFFFFFFFF
            assert st.reached == int(reached.sum())
            assert st.iterations == int(f["bfs"][i][reached].max())
            d, st = g.sssp(int(s))
            np.testing.assert_array_equal(d, f["sssp"][i], err_msg="sssp src %d" % s)
            # reach statistics (a scan over V that need not be a multiple of 64)
            r_ss = f["sssp"][i] != oracle.SSSP_SENTINEL_SHARE
            assert st.reached == int(r_ss.sum())
            assert st.edges_traversed == int(ptr_degrees(f["ptr"], f["col"].size)[r_ss].sum())
        lab, _ = g.cc()
        np.testing.assert_array_equal(lab, f["cc"])
        r, _ = g.pagerank(0.85, 20)
        assert np.abs(r.astype(np.float64) - f["pr"]).sum() <= PR_L1


@pytest.fixture(scope="module")
def rmat16():
    g = pg.gen_rmat(16, 16, seed=1).add_weights(1, 64, seed=2)
    gg = pg.Graph.from_csr(g)
    yield g, gg
    gg.close()


def test_rmat16_bfs_golden(rmat16):
    g, gg = rmat16
    say("case rmat16 bfs from 0")
    meta = {}
    with open(os.path.join(GOLDEN, "rmat16_bfs0.txt")) as fh:
        for line in fh:
            if not line.startswith("#"):
                k, v = line.split()
                meta[k] = v
    lv, st = gg.bfs(0)
    assert hashlib.sha256(lv.tobytes()).hexdigest() == meta["levels_sha256"]
    assert st.iterations == int(meta["max_level"])
    assert st.reached == int(meta["reached"])
    assert st.edges_traversed == int(meta["edges_traversed"])


def test_rmat16_vs_oracle(rmat16):
    g, gg = rmat16
    ptr, col, w = g.node_pointers, g.cols, g.weights
    deg = np.diff(g.offsets.astype(np.int64))
    rng = np.random.default_rng(3)
    srcs = [0] + list(rng.choice(np.nonzero(deg > 0)[0], 6, replace=False)) + [int(np.nonzero(deg == 0)[0][0])]
    for s in srcs:
        say("case rmat16 src %d" % s)
        lv, _ = gg.bfs(int(s))
        want, _, _ = oracle.ref_loop(oracle.BFS, ptr, col, src=int(s), threads=8)
        np.testing.assert_array_equal(lv, want)
        d, st = gg.sssp(int(s))
        np.testing.assert_array_equal(d, oracle.dijkstra(ptr, col, w, int(s)))
    say("case rmat16 cc")
    lab, st = gg.cc()
    np.testing.assert_array_equal(lab, oracle.cc_reach_min(ptr, col))
    assert st.reached == len(np.unique(lab))
    say("case rmat16 pagerank")
    for iters in (0, 1, 20):
        r, _ = gg.pagerank(0.85, iters)
        assert pr_l1(ptr, col, r, iters) <= PR_L1


def test_rmat16_sentinel_opt_and_aos(rmat16):
    g, _ = rmat16
    say("case rmat16 aos + opt sentinel")
    edges = np.stack([g.cols, g.weights], axis=1).astype(np.uint32)
    cfg = pg.default_config(sssp_sentinel=pg.SSSP_SENTINEL_OPT)
    with pg.Graph.from_aos(g.node_pointers, edges, config=cfg) as ga:
        deg = np.diff(g.offsets.astype(np.int64))
        s = int(np.nonzero(deg > 0)[0][100])
        d, _ = ga.sssp(s)
        np.testing.assert_array_equal(d, oracle.dijkstra(g.node_pointers, g.cols, g.weights, s,
                                                         oracle.SSSP_SENTINEL_OPT))
        assert (d == 0xFFFFFFFF).any()


def test_device_outputs_and_determinism(rmat16):
    g, gg = rmat16
    say("case rmat16 device outputs")
    buf = pg.DeviceBuffer(4 * g.V)
    try:
        h, _ = gg.bfs(5)
        gg.bfs(5, out=buf)
        np.testing.assert_array_equal(buf.to_host(np.uint32, g.V), h)
        r1, _ = gg.pagerank(0.85, 20)
        gg.pagerank(0.85, 20, out=buf)
        np.testing.assert_array_equal(buf.to_host(np.float32, g.V), r1)
    finally:
        buf.free()
    for _ in range(3):
        r2, _ = gg.pagerank(0.85, 20)
        np.testing.assert_array_equal(r2, r1)  # bitwise: fixed reduction order
        l2, _ = gg.bfs(5)
        np.testing.assert_array_equal(l2, h)
        c2, _ = gg.cc()
        np.testing.assert_array_equal(c2, oracle.cc_reach_min(g.node_pointers, g.cols))


def test_pagerank_bits_do_not_depend_on_the_work_split():
    """PageRank sums are exact fixed-point sums (63 fraction bits), so the
    ranks are the same bits however the work is divided: the binned layout,
    the in-list pull on one workgroup, an odd grid, every list cut into
    64-edge hub chunks, the default bins, and the allocation-free entry
    point's buffers. And within 1e-6 L1 of FP64."""
    csr = pg.gen_rmat(14, 16, seed=9)
    ptr, col = csr.node_pointers, csr.cols
    want = oracle.pagerank_f64(ptr, col, 0.85, 20)
    ref = None
    pull = pg.PR_PULL
    for cfg in ({}, {"pr_layout": pull}, {"pr_layout": pull, "grid_blocks": 1},
                {"pr_layout": pull, "grid_blocks": 37}, {"pr_layout": pull, "small_degree": 1, "large_degree": 64},
                {"pr_layout": pull, "flags": pg.FLAG_NO_RELABEL}, {"flags": pg.FLAG_NO_RELABEL},
                {"pr_layout": pull, "pr_block": 512}, {"pr_layout": pull, "pr_block": 1024},
                {"pr_layout": pull, "pr_block": 1024, "grid_blocks": 3}, {"grid_blocks": 3}):
        say("case pagerank split %s" % cfg)
        with pg.Graph.from_csr(csr, weighted=False, config=pg.default_config(**cfg)) as g:
            r, _ = g.pagerank(0.85, 20)
            assert float(np.abs(r.astype(np.float64) - want).sum()) <= PR_L1
            if ref is None:
                ref = r
            np.testing.assert_array_equal(r.view(np.uint32), ref.view(np.uint32), err_msg=str(cfg))
            buf = pg.DeviceBuffer(4 * g.V)
            try:
                g.pagerank_noalloc(buf, 0.85, 20)
                np.testing.assert_array_equal(buf.to_host(np.uint32, g.V), ref.view(np.uint32), err_msg=str(cfg))
            finally:
                buf.free()


def hub_graph(rng, V, hubs, leaves, extra):
    """Symmetric graph whose first `hubs` vertices each join `leaves` random
    others, plus `extra` random edges: the hub partition is cut into many
    chunks."""
    a = [np.repeat(np.arange(hubs), leaves), rng.integers(0, V, extra)]
    b = [rng.integers(0, V, hubs * leaves), rng.integers(0, V, extra)]
    s, d = np.concatenate(a), np.concatenate(b)
    keep = s != d
    s, d = np.concatenate([s[keep], d[keep]]), np.concatenate([d[keep], s[keep]])
    pairs = np.unique(s.astype(np.int64) * V + d)
    s, d = pairs // V, pairs % V
    return np.searchsorted(s, np.arange(V)).astype(np.uint32), d.astype(np.uint32)


def pagerank_layouts(ptr, col, cfg, iters, want=None):
    """Ranks of the in-list pull and of the binned layout (forced and
    automatic) on the same graph: the same bits, within 1e-6 L1 of FP64."""
    if want is None:
        want = oracle.pagerank_f64(ptr, col, 0.85, iters)
    out = {}
    for layout in (pg.PR_PULL, pg.PR_BINNED, pg.PR_AUTO):
        with pg.Graph(ptr, col, config=pg.default_config(pr_layout=layout, **cfg)) as g:
            before = g.device_bytes
            r, _ = g.pagerank(0.85, iters)
            grown = g.device_bytes - before
            if layout == pg.PR_BINNED and iters and col.size:
                assert grown >= 2 * col.size, grown  # its 16-bit edge entries are on the device
            assert float(np.abs(r.astype(np.float64) - want).sum()) <= PR_L1, (layout, cfg)
            out[layout] = r
    for layout in (pg.PR_BINNED, pg.PR_AUTO):
        np.testing.assert_array_equal(out[layout].view(np.uint32), out[pg.PR_PULL].view(np.uint32),
                                      err_msg="layout %d %s" % (layout, cfg))
    return out[pg.PR_PULL]


@pytest.mark.parametrize("case", ["rmat16", "rmat15_no_relabel", "hubs", "directed", "sparse", "odd_sizes"])
def test_pagerank_binned_equals_pull(case):
    """Binned PageRank (destination partitions of 16384 vertices, LDS sums,
    16-bit edge entries) against the in-list pull: the same bits for 0, 1, 7
    and 20 iterations, on R-MAT with and without the degree order, a graph
    whose hub partition is cut into many chunks, a directed graph (the binned
    layout reads out-lists, the pull the transpose), a graph too sparse for
    the automatic layout and sizes around the partition width."""
    rng = np.random.default_rng(11)
    cfg = {}
    if case == "rmat16":
        csr = pg.gen_rmat(16, 16, seed=3)
        graphs = [(csr.node_pointers, csr.cols)]
    elif case == "rmat15_no_relabel":
        csr = pg.gen_rmat(15, 16, seed=4)
        graphs = [(csr.node_pointers, csr.cols)]
        cfg = {"flags": pg.FLAG_NO_RELABEL}
    elif case == "hubs":
        graphs = [hub_graph(rng, 70000, 40, 30000, 200000)]
    elif case == "directed":
        ptr, col, _ = random_directed(rng, 50000, 600000, 1)
        graphs = [(ptr, col)]
    elif case == "sparse":
        graphs = [hub_graph(rng, 120000, 0, 0, 30000)]
    else:
        graphs = [hub_graph(rng, V, 3, 500, 4 * V) for V in (16383, 16384, 16385, 32769)]
    for ptr, col in graphs:
        say("case pagerank layouts %s V=%d E=%d" % (case, ptr.size, col.size))
        for iters in (0, 1, 7, 20):
            pagerank_layouts(ptr, col, cfg, iters)


def test_pagerank_layout_tiny_and_dangling():
    """One vertex, vertices without edges, and a graph where every vertex but
    one is dangling: both layouts give the oracle's ranks, the same bits."""
    cases = [
        (np.zeros(1, np.uint32), np.zeros(0, np.uint32)),
        (np.zeros(5, np.uint32), np.zeros(0, np.uint32)),
        (np.array([0, 4, 4, 4, 4], np.uint32), np.array([1, 2, 3, 4], np.uint32)),  # star out of 0
    ]
    for ptr, col in cases:
        say("case pagerank tiny V=%d E=%d" % (ptr.size, col.size))
        for iters in (0, 1, 5):
            pagerank_layouts(ptr, col, {}, iters)


def test_pagerank_autotune_workgroup_size():
    """The tuner's PageRank mode picks the layout (in-list pull or binned)
    and, for the pull, the gather's workgroup size (256, 512 or 1024 threads,
    the LDS hub stage growing with it); the ranks keep their bits, and
    set_config refuses other sizes and layouts."""
    csr = pg.gen_rmat(15, 16, seed=4)
    with pg.Graph.from_csr(csr, weighted=False) as g:
        want, _ = g.pagerank(0.85, 20)
        say("case pagerank autotune")
        c = g.autotune(np.array([0], np.uint32), what=pg.TUNE_PAGERANK)
        assert c.pr_layout in (pg.PR_PULL, pg.PR_BINNED) and g.get_config().pr_layout == c.pr_layout
        assert c.pr_block in (256, 512, 1024) and g.get_config().pr_block == c.pr_block
        np.testing.assert_array_equal(g.pagerank(0.85, 20)[0].view(np.uint32), want.view(np.uint32))
        for layout in (pg.PR_PULL, pg.PR_BINNED, pg.PR_AUTO):
            g.tune(pr_layout=layout)
            np.testing.assert_array_equal(g.pagerank(0.85, 20)[0].view(np.uint32), want.view(np.uint32))
        for bad in (128, 300, 2048):
            with pytest.raises(pg.PtgError) as e:
                g.tune(pr_block=bad)
            assert e.value.code == pg.ERR_INVALID_ARG
        with pytest.raises(pg.PtgError) as e:
            g.tune(pr_layout=3)
        assert e.value.code == pg.ERR_INVALID_ARG


def random_directed(rng, V, E, wmax):
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ptr = np.searchsorted(src, np.arange(V)).astype(np.uint32)
    w = rng.integers(0, wmax + 1, E).astype(np.uint32)
    return ptr, dst.astype(np.uint32), w


@pytest.mark.parametrize("flags", [0, pg.FLAG_NO_RELABEL], ids=["relabel", "no_relabel"])
@pytest.mark.parametrize("seed", [1, 2])
def test_directed_graph(seed, flags):
    rng = np.random.default_rng(seed)
    ptr, col, w = random_directed(rng, 3000, 9000, 9)
    say("case directed seed %d flags %d" % (seed, flags))
    with pg.Graph(ptr, col, w, config=pg.default_config(flags=flags)) as g:
        assert not g.is_symmetric()
        for s in (0, 17, 2999):
            lv, _ = g.bfs(s)
            np.testing.assert_array_equal(lv, oracle.bfs(ptr, col, s)[0])
            d, _ = g.sssp(s)
            np.testing.assert_array_equal(d, oracle.dijkstra(ptr, col, w, s))
        lab, _ = g.cc()
        np.testing.assert_array_equal(lab, oracle.cc_reach_min(ptr, col))
        r, _ = g.pagerank(0.85, 20)
        assert pr_l1(ptr, col, r) <= PR_L1


def test_tiny_graphs():
    say("case tiny graphs")
    with pg.Graph(np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32)) as g:
        assert g.bfs(0)[0].tolist() == [1]
        assert g.sssp(0)[0].tolist() == [1]
        assert g.cc()[0].tolist() == [0]
        r, _ = g.pagerank(0.85, 20)
        assert abs(float(r[0]) - 1.0) < 1e-6
    with pg.Graph(np.array([0, 1], np.uint32), np.array([1, 0], np.uint32), np.array([4, 4], np.uint32)) as g:
        assert g.bfs(1)[0].tolist() == [2, 1]
        assert g.sssp(0)[0].tolist() == [1, 5]
        with pytest.raises(pg.PtgError) as e:
            g.bfs(2)
        assert e.value.code == pg.ERR_INVALID_ARG


def test_unweighted_graph_rejects_sssp():
    say("case unweighted sssp")
    with pg.Graph(np.array([0, 1], np.uint32), np.array([1, 0], np.uint32)) as g:
        with pytest.raises(pg.PtgError) as e:
            g.sssp(0)
        assert e.value.code == pg.ERR_UNSUPPORTED


def test_reference_named_drivers(rmat16):
    g, _ = rmat16
    say("case reference-named drivers")
    L = pg.lib()
    ptr = np.ascontiguousarray(g.node_pointers)
    col = np.ascontiguousarray(g.cols)
    edges = np.ascontiguousarray(np.stack([g.cols, g.weights], axis=1).astype(np.uint32))
    ep = edges.ctypes.data_as(__import__("ctypes").POINTER(pg.EdgeWithWeight))
    for fn in ("bfsCaculateInShare", "bfsCaculateInOpt", "bfsCaculateInAsync", "bfsCaculateInAsyncSwap"):
        assert getattr(L, fn)(g.V, g.E, pg._u32(ptr), pg._u32(col), 0) >= 0
    assert L.ssspCaculateInShare(g.V, g.E, pg._u32(ptr), ep, 0) >= 0
    assert L.ssspCaculateInOpt(g.V, g.E, pg._u32(ptr), ep, 0) >= 0
    for fn in ("ccCaculateInShare", "ccCaculateInOpt", "prCaculateInShare", "prCaculateInOpt"):
        assert getattr(L, fn)(g.V, g.E, pg._u32(ptr), pg._u32(col)) >= 0
    bad = ptr.copy()
    bad[3] = bad[4] + 1
    assert L.bfsCaculateInShare(g.V, g.E, pg._u32(bad), pg._u32(col), 0) == pg.ERR_INVALID_GRAPH


def test_memcpy_d2d():
    say("case memcpy d2d")
    rng = np.random.default_rng(5)
    n = (1 << 20) + 37
    src = rng.integers(0, 256, n, dtype=np.uint8)
    a, b = pg.DeviceBuffer(n), pg.DeviceBuffer(n)
    try:
        a.from_host(src)
        for nb in (0, 1, 15, 16, 4097, n):
            b.from_host(np.zeros(n, np.uint8))
            b.copy_from(a, nb)
            got = b.to_host(np.uint8, n)
            np.testing.assert_array_equal(got[:nb], src[:nb])
            assert not got[nb:].any()
        # unaligned pointers take the byte path
        b.from_host(np.zeros(n, np.uint8))
        assert pg.lib().ptg_memcpy_d2d(b.ptr + 1, a.ptr + 3, n - 3) == pg.OK
        got = b.to_host(np.uint8, n)
        np.testing.assert_array_equal(got[1:n - 2], src[3:])
        assert got[0] == 0 and not got[n - 2:].any()
    finally:
        a.free()
        b.free()
    host = np.zeros(16, np.uint8)
    assert pg.lib().ptg_memcpy_d2d(host.ctypes.data, host.ctypes.data, 16) == pg.ERR_INVALID_ARG


def test_set_flags_timing(rmat16):
    g, gg = rmat16
    say("case set_flags")
    gg.set_flags(pg.FLAG_TIMING)
    _, st = gg.bfs(0)
    assert st.main_launches > 0 and st.main_kernel_ns > 0
    gg.set_flags(0)
    lv, st = gg.bfs(0)
    assert st.main_launches == 0
    want, _, _ = oracle.ref_loop(oracle.BFS, g.node_pointers, g.cols, src=0, threads=8)
    np.testing.assert_array_equal(lv, want)
    with pytest.raises(pg.PtgError) as e:
        gg.set_flags(pg.FLAG_NO_RELABEL)
    assert e.value.code == pg.ERR_INVALID_ARG


def test_sssp_steps_rmat16(rmat16):
    """Light/heavy rounds for several steps (the split CSR is rebuilt per step)."""
    g, gg = rmat16
    ptr, col, w = g.node_pointers, g.cols, g.weights
    deg = np.diff(g.offsets.astype(np.int64))
    srcs = [0] + [int(x) for x in np.random.default_rng(9).choice(np.nonzero(deg > 0)[0], 3, replace=False)]
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    try:
        for delta in (1, 4, 16, 63, 64, 65, 4096):
            gg.tune(sssp_delta=delta)
            say("case rmat16 sssp delta %d" % delta)
            for s in srcs:
                d, st = gg.sssp(s)
                np.testing.assert_array_equal(d, want[s], err_msg="delta %d src %d" % (delta, s))
    finally:
        gg.tune(sssp_delta=0)


def test_sssp_deferral_rmat16(rmat16):
    """Deferral of high-degree vertices, with and without buckets."""
    g, gg = rmat16
    ptr, col, w = g.node_pointers, g.cols, g.weights
    deg = np.diff(g.offsets.astype(np.int64))
    srcs = [0] + [int(x) for x in np.random.default_rng(11).choice(np.nonzero(deg > 0)[0], 3, replace=False)]
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    rounds = {}
    try:
        for delta in (0, 24):
            for ddeg, win in ((32, 1), (64, 2), (64, 100), (500, 1), (0xFFFFFFFF, 0)):
                gg.tune(sssp_delta=delta, sssp_defer_degree=ddeg, sssp_defer_window=win)
                say("case rmat16 sssp delta %d defer degree %d window %d" % (delta, ddeg, win))
                total = 0
                for s in srcs:
                    d, st = gg.sssp(s)
                    np.testing.assert_array_equal(d, want[s], err_msg="delta %d defer %d/%d src %d" % (
                        delta, ddeg, win, s))
                    total += st.iterations
                rounds[(delta, ddeg, win)] = total
        say("rounds %s" % rounds)
        # one unbounded bucket takes fewer rounds than buckets of 24
        assert rounds[(0, 64, 2)] < rounds[(24, 64, 2)]
    finally:
        gg.tune(sssp_delta=0, sssp_defer_degree=0, sssp_defer_window=0)


@pytest.mark.parametrize("lo,hi", [(0, 4), (1, 1), (1, 100000), (1000, 1000000)])
def test_sssp_weight_ranges(lo, hi):
    """The automatic deferral window follows the mean weight; values stay exact
    for zero weights, unit weights and weights far above 16 bits."""
    g = pg.gen_rmat(14, 16, seed=3).add_weights(lo, hi, seed=4)
    ptr, col, w = g.node_pointers, g.cols, g.weights
    deg = np.diff(g.offsets.astype(np.int64))
    srcs = [int(x) for x in np.random.default_rng(5).choice(np.nonzero(deg > 0)[0], 4, replace=False)]
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph.from_csr(g) as gg:
        for cfg in ({}, {"sssp_defer_degree": 16}):
            gg.tune(**cfg)
            for s in srcs:
                say("case rmat14 weights [%d, %d] %s src %d" % (lo, hi, cfg, s))
                d, _ = gg.sssp(s)
                np.testing.assert_array_equal(d, want[s], err_msg="weights [%d, %d] %s src %d" % (lo, hi, cfg, s))
        rows, _ = gg.sssp_batch(np.array(srcs, dtype=np.uint32))
        for k, s in enumerate(srcs):
            np.testing.assert_array_equal(rows[k], want[s], err_msg="batch weights [%d, %d] src %d" % (lo, hi, s))


def test_frontier_sizes_match_levels(rmat16):
    g, gg = rmat16
    say("case frontier sizes")
    for s in (0, 100):
        lv, st = gg.bfs(s)
        fr = gg.frontier_sizes()
        hist = np.bincount(lv[lv != 0xFFFFFFFF])[1:]  # vertices per level 1, 2, ...
        assert fr.size == st.iterations == hist.size
        np.testing.assert_array_equal(fr, hist)
        assert int(fr.sum()) == st.frontier_sum == st.reached


@pytest.mark.parametrize("flags", [0, pg.FLAG_NO_RELABEL], ids=["relabel", "no_relabel"])
def test_depth_changes_between_queries(flags):
    """A query's first launch batch and its closing work (end event, reach
    statistics, output delivery) are sized from the previous query's depth;
    shallow and deep queries in turn, to host and device outputs, must match
    the oracle with exact statistics."""
    rng = np.random.default_rng(11)
    src, dst = list(range(299)), list(range(1, 300))  # path 0 -> ... -> 299
    for leaf in range(301, 400):  # star around 300, both directions
        src += [300, leaf]
        dst += [leaf, 300]
    order = np.argsort(np.array(src), kind="stable")
    s_arr, col = np.array(src)[order], np.array(dst, dtype=np.uint32)[order]
    V, E = 400, col.size
    ptr = np.searchsorted(s_arr, np.arange(V)).astype(np.uint32)
    w = rng.integers(1, 50, E, dtype=np.uint32)
    deg = ptr_degrees(ptr, E)
    buf = pg.DeviceBuffer(4 * V)
    with pg.Graph(ptr, col, w, config=pg.default_config(flags=flags)) as g:
        for k, s in enumerate([300, 0, 300, 150, 399, 0, 0, 300]):
            say("case depth change flags=%d query %d src %d" % (flags, k, s))
            want_b, n_lev, n_reached, n_edges = oracle.bfs(ptr, col, s)
            want_d = oracle.dijkstra(ptr, col, w, s)
            if k % 2:
                _, st_b = g.bfs(s, out=buf)
                lv = buf.to_host(np.uint32, V)
                _, st_d = g.sssp(s, out=buf)
                d = buf.to_host(np.uint32, V)
            else:
                lv, st_b = g.bfs(s)
                d, st_d = g.sssp(s)
            np.testing.assert_array_equal(lv, want_b, err_msg="bfs src %d" % s)
            np.testing.assert_array_equal(d, want_d, err_msg="sssp src %d" % s)
            assert (st_b.reached, st_b.iterations) == (n_reached, n_lev)
            assert st_b.edges_traversed == n_edges
            r_ss = want_d != oracle.SSSP_SENTINEL_SHARE
            assert st_d.reached == int(r_ss.sum())
            assert st_d.edges_traversed == int(deg[r_ss].sum())


@pytest.mark.parametrize("flags", [0, pg.FLAG_NO_RELABEL], ids=["relabel", "no_relabel"])
def test_device_outputs_at_unaligned_offsets(flags):
    """Results written to device memory at offsets that are not 16-byte
    aligned (the value array of a graph without relabeling is the caller's
    buffer, filled in place) match the host results, and the words around
    the window keep their contents."""
    rng = np.random.default_rng(21)
    V = 1003
    ptr, col, w = random_directed(rng, V, 6000, 40)
    pad = 8
    guard = np.full(V + 2 * pad, 0xA5A5A5A5, dtype=np.uint32)
    buf = pg.DeviceBuffer(4 * guard.size)
    with pg.Graph(ptr, col, w, config=pg.default_config(flags=flags)) as g:
        want = {"bfs": g.bfs(5)[0], "sssp": g.sssp(5)[0], "cc": g.cc()[0]}
        for off in (1, 2, 3, 5):
            for kind in ("bfs", "sssp", "cc"):
                say("case unaligned output flags=%d offset %d %s" % (flags, off, kind))
                buf.from_host(guard)
                view = buf.view(4 * (pad + off), 4 * V)
                if kind == "cc":
                    g.cc(out=view)
                else:
                    getattr(g, kind)(5, out=view)
                got = buf.to_host(np.uint32, guard.size)
                np.testing.assert_array_equal(got[pad + off:pad + off + V], want[kind], err_msg="%s at +%d" % (kind, off))
                np.testing.assert_array_equal(got[:pad + off], guard[:pad + off])
                np.testing.assert_array_equal(got[pad + off + V:], guard[pad + off + V:])


def symmetric_csr(V, pairs):
    """Reference-layout CSR (V offsets) of the symmetric closure of `pairs`,
    without self-loops or duplicates."""
    a = np.array([p[0] for p in pairs] + [p[1] for p in pairs], dtype=np.int64)
    b = np.array([p[1] for p in pairs] + [p[0] for p in pairs], dtype=np.int64)
    keep = a != b
    e = np.unique(np.stack([a[keep], b[keep]], axis=1), axis=0)
    ptr = np.searchsorted(e[:, 0], np.arange(V)).astype(np.uint32)
    return ptr, e[:, 1].astype(np.uint32)


def cc_shapes(rng):
    V = 3000
    shapes = {}
    # two dense halves of equal size: no single giant component
    pairs = [tuple(rng.integers(0, 1500, 2)) for _ in range(9000)]
    pairs += [tuple(rng.integers(1500, 3000, 2)) for _ in range(9000)]
    shapes["two_giants"] = pairs
    # 300 random trees of 10 vertices
    perm = rng.permutation(V)
    shapes["many_small"] = [(perm[10 * t + i], perm[10 * t + rng.integers(0, i)]) for t in range(300) for i in range(1, 10)]
    # a dense cluster on random ids plus chains of 2-5 vertices elsewhere
    ids = rng.permutation(V)
    cl, rest = ids[:2000], ids[2000:]
    pairs = [(cl[rng.integers(0, 2000)], cl[rng.integers(0, 2000)]) for _ in range(12000)]
    i = 0
    while i + 1 < rest.size:
        n = int(rng.integers(2, 6))
        pairs += [(rest[j], rest[j + 1]) for j in range(i, min(i + n, rest.size) - 1)]
        i += n
    shapes["giant_plus_chains"] = pairs
    # hubs in a ring whose two lowest-id neighbours are leaves of their own
    # (ids 0-19): without relabeling the hubs meet only through edges past
    # each one's first two neighbours
    pairs = []
    for h in range(10):
        hub = 100 + 290 * h
        pairs += [(hub, 2 * h), (hub, 2 * h + 1)]
        pairs += [(hub, leaf) for leaf in range(hub + 1, hub + 200)]
        pairs.append((hub, 100 + 290 * ((h + 1) % 10)))
    shapes["hubs_leaves_first"] = pairs
    return V, shapes


@pytest.mark.parametrize("flags", [0, pg.FLAG_NO_RELABEL], ids=["relabel", "no_relabel"])
def test_cc_component_shapes(flags):
    """Union-find CC (first neighbours linked, the most sampled component
    skipped, the rest linked) against the oracle on symmetric graphs with no
    giant component, many small ones, a giant one with chains beside it, and
    hubs joined only through their later edges."""
    V, shapes = cc_shapes(np.random.default_rng(31))
    for name, pairs in shapes.items():
        ptr, col = symmetric_csr(V, pairs)
        want = oracle.cc_union_find(ptr, col)
        say("case cc shape %s flags=%d components %d" % (name, flags, len(np.unique(want))))
        with pg.Graph(ptr, col, config=pg.default_config(flags=flags)) as g:
            assert g.is_symmetric()
            for _ in range(2):
                lab, st = g.cc()
                np.testing.assert_array_equal(lab, want, err_msg=name)
                assert st.reached == len(np.unique(want))


# Repeat counts matter to PageRank: 0->1 twice against 1->0 once has every
# reverse edge present, but it is not symmetric as a multigraph, so the pull
# must use the transposed lists (vertex 1 gets 0's contribution twice).
MULTI_UNEQUAL = ([0, 2, 4, 5], [1, 1, 0, 2, 1])       # adj 0:[1,1] 1:[0,2] 2:[1]
MULTI_EQUAL = ([0, 3, 5, 6], [1, 2, 1, 0, 0, 0])      # adj 0:[1,2,1] 1:[0,0] 2:[0]


@pytest.mark.parametrize("case", ["unequal", "equal"])
@pytest.mark.parametrize("mode", ["relabel", "no_relabel", "stream"])
def test_repeat_counts_and_symmetry(case, mode):
    ptr, col = MULTI_UNEQUAL if case == "unequal" else MULTI_EQUAL
    ptr = np.array(ptr[:-1], np.uint32)
    col = np.array(col, np.uint32)
    w = np.arange(1, col.size + 1, dtype=np.uint32)
    flags = {"relabel": 0, "no_relabel": pg.FLAG_NO_RELABEL, "stream": pg.FLAG_STREAM}[mode]
    say("case repeat counts %s %s" % (case, mode))
    cfg = pg.default_config(flags=flags)
    with pg.Graph(ptr, col, None if mode == "stream" else w, config=cfg) as g:
        assert g.is_symmetric() == (case == "equal")
        if mode == "stream" and case == "unequal":
            # streamed PageRank pulls over the out-lists: refused, not wrong
            with pytest.raises(pg.PtgError) as e:
                g.pagerank(0.85, 20)
            assert e.value.code == pg.ERR_UNSUPPORTED
        else:
            r, _ = g.pagerank(0.85, 20)
            assert pr_l1(ptr, col, r) <= PR_L1
        for s in range(ptr.size):
            np.testing.assert_array_equal(g.bfs(s)[0], oracle.bfs(ptr, col, s)[0])
        if mode != "stream":
            lab, _ = g.cc()
            np.testing.assert_array_equal(lab, oracle.cc_reach_min(ptr, col))


def test_repeat_counts_rmat_pagerank():
    """R-MAT-12 with every edge of vertex 3 doubled in one direction only:
    the resident PageRank follows the FP64 oracle (with and without
    relabeling), and the streamed graph refuses PageRank."""
    csr = pg.gen_rmat(12, 16, seed=1)
    ptr, col = csr.node_pointers.astype(np.int64), csr.cols
    E = col.size
    nxt = np.append(ptr[1:], E)
    lists = [col[ptr[u]:nxt[u]] for u in range(ptr.size)]
    lists[3] = np.repeat(lists[3], 2)
    col2 = np.concatenate(lists).astype(np.uint32)
    ptr2 = np.concatenate([[0], np.cumsum([l.size for l in lists])[:-1]]).astype(np.uint32)
    say("case repeat counts rmat12")
    for flags in (0, pg.FLAG_NO_RELABEL, pg.FLAG_STREAM):
        with pg.Graph(ptr2, col2, config=pg.default_config(flags=flags)) as g:
            assert not g.is_symmetric()
            if flags == pg.FLAG_STREAM:
                with pytest.raises(pg.PtgError) as e:
                    g.pagerank(0.85, 20)
                assert e.value.code == pg.ERR_UNSUPPORTED
                continue
            r, _ = g.pagerank(0.85, 20)
            assert pr_l1(ptr2, col2, r) <= PR_L1
