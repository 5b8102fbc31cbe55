"""GPU parity of the batched multi-source queries (ptg_bfs_batch,
ptg_sssp_batch): row k of a batch must equal the oracle (and the
single-source path) from srcs[k], bit for bit."""
import glob
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

CONFIGS = {
    "default": {},
    "tiny_bins": {"small_degree": 1, "large_degree": 64},
    "bottom_up": {"td_to_bu": 1e9},
    "top_down": {"td_to_bu": 1e-9},
    "one_block": {"grid_blocks": 1},
    "no_relabel": {"flags": pg.FLAG_NO_RELABEL},
    "delta1": {"sssp_delta": 1},
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    pg.init(0)
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def batch_sources(V, rng, n=64):
    if V <= n:
        s = np.arange(V)
        return np.concatenate([s, s[: min(V, 3)]])[:n]  # duplicates in one batch
    return rng.choice(V, n, replace=False)


def check_rows(ptr, col, w, srcs, lv, dist, sentinel=oracle.SSSP_SENTINEL_SHARE):
    for k, s in enumerate(srcs):
        want_b = oracle.bfs(ptr, col, int(s))[0]
        np.testing.assert_array_equal(lv[k], want_b, err_msg="bfs row %d src %d" % (k, s))
        if dist is not None:
            want_d = oracle.dijkstra(ptr, col, w, int(s), sentinel)
            np.testing.assert_array_equal(dist[k], want_d, err_msg="sssp row %d src %d" % (k, s))


@pytest.mark.parametrize("cfg", list(CONFIGS), ids=list(CONFIGS))
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_batch_fixture_parity(path, cfg):
    f = load(path)
    ptr, col, w = f["ptr"], f["col"], f["w"]
    rng = np.random.default_rng(7)
    srcs = batch_sources(ptr.size, rng)
    say("case batch %s cfg=%s V=%d n=%d" % (os.path.basename(path), cfg, ptr.size, srcs.size))
    with pg.Graph(ptr, col, w, config=pg.default_config(**CONFIGS[cfg])) as g:
        lv, st = g.bfs_batch(srcs)
        dist, st2 = g.sssp_batch(srcs)
        check_rows(ptr, col, w, srcs, lv, dist)
        reached = int((lv != 0xFFFFFFFF).sum())
        assert st.reached == reached
        assert st2.reached == int((dist != oracle.SSSP_SENTINEL_SHARE).sum())
        deg = np.diff(np.append(ptr.astype(np.int64), col.size))
        assert st.edges_traversed == int(((lv != 0xFFFFFFFF) * deg[None, :]).sum())


def test_batch_rmat16_vs_single_and_replays():
    say("case batch rmat16")
    csr = pg.gen_rmat(16, 16, seed=1).add_weights(1, 64, seed=2)
    ptr, col, w = csr.node_pointers, csr.cols, csr.weights
    deg = np.diff(csr.offsets.astype(np.int64))
    rng = np.random.default_rng(3)
    srcs = rng.choice(np.nonzero(deg > 0)[0], 64, replace=False)
    srcs[5] = int(np.nonzero(deg == 0)[0][0])  # an isolated source
    srcs[9] = srcs[8]                          # a duplicate
    with pg.Graph.from_csr(csr) as g:
        # three batches in a row (graph replays reused), shrinking n
        for n in (64, 17, 64, 1):
            say("case batch rmat16 n=%d" % n)
            lv, st = g.bfs_batch(srcs[:n])
            dist, _ = g.sssp_batch(srcs[:n])
            for k in range(n):
                s = int(srcs[k])
                np.testing.assert_array_equal(lv[k], g.bfs(s)[0], err_msg="bfs row %d" % k)
                np.testing.assert_array_equal(dist[k], g.sssp(s)[0], err_msg="sssp row %d" % k)
            for k in (0, 5, n - 1):
                if k < n:
                    check_rows(ptr, col, w, srcs[k:k + 1], lv[k:k + 1], dist[k:k + 1])
        # device output buffer
        buf = pg.DeviceBuffer(4 * 64 * csr.V)
        try:
            g.bfs_batch(srcs, out=buf)
            got = buf.to_host(np.uint32, 64 * csr.V).reshape(64, csr.V)
            np.testing.assert_array_equal(got, g.bfs_batch(srcs)[0])
            g.sssp_batch(srcs, out=buf)
            got = buf.to_host(np.uint32, 64 * csr.V).reshape(64, csr.V)
            np.testing.assert_array_equal(got, g.sssp_batch(srcs)[0])
        finally:
            buf.free()


def symmetric_unequal_weights(csr, seed):
    """The graph's structure with an independent random weight on each
    direction of every edge (in-pairs then differ from the out-pairs)."""
    ptr, col = csr.node_pointers, csr.cols
    w = np.random.default_rng(seed).integers(1, 65, col.size).astype(np.uint32)
    return ptr, col, w


@pytest.mark.parametrize("pull", [-1.0, 0.0, 1e-9], ids=["never", "auto", "every_round"])
@pytest.mark.parametrize("weights", ["same_both_ways", "differ_by_direction"])
def test_batch_sssp_pull_rounds(pull, weights):
    """Batched SSSP pull rounds: the first batch on a graph pushes and
    records which rounds were dense; later batches pull in those rounds
    (every round but the first with batch_pull = 1e-9). Rows of three batches
    in a row (different sources, then a smaller batch) equal the single
    queries and, for three rows each, the oracle."""
    csr = pg.gen_rmat(15, 16, seed=21).add_weights(1, 64, seed=22)
    if weights == "same_both_ways":
        ptr, col, w = csr.node_pointers, csr.cols, csr.weights
    else:
        ptr, col, w = symmetric_unequal_weights(csr, 23)
    deg = np.diff(np.append(ptr.astype(np.int64), col.size))
    rng = np.random.default_rng(24)
    with pg.Graph(ptr, col, w, config=pg.default_config(batch_pull=pull)) as g:
        assert g.is_symmetric()
        for b, n in enumerate((64, 64, 23)):
            srcs = rng.choice(np.nonzero(deg > 0)[0], n, replace=False).astype(np.uint32)
            say("case batch sssp pull=%g weights=%s batch %d n=%d" % (pull, weights, b, n))
            dist, st = g.sssp_batch(srcs)
            for k in range(n):
                np.testing.assert_array_equal(dist[k], g.sssp(int(srcs[k]))[0], err_msg="row %d" % k)
            for k in (0, n // 2, n - 1):
                np.testing.assert_array_equal(dist[k], oracle.dijkstra(ptr, col, w, int(srcs[k])),
                                              err_msg="oracle row %d" % k)
            assert st.reached == int((dist != oracle.SSSP_SENTINEL_SHARE).sum())


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_batch_fixture_pull_every_round(path):
    """Every fixture (directed ones push only) with pulls in every round after
    the first: two batches, the second equal to the oracle."""
    f = load(path)
    ptr, col, w = f["ptr"], f["col"], f["w"]
    srcs = batch_sources(ptr.size, np.random.default_rng(7))
    say("case batch pull %s V=%d" % (os.path.basename(path), ptr.size))
    with pg.Graph(ptr, col, w, config=pg.default_config(batch_pull=1e-9)) as g:
        first, _ = g.sssp_batch(srcs)
        dist, _ = g.sssp_batch(srcs)
        np.testing.assert_array_equal(dist, first)
        check_rows(ptr, col, w, srcs, g.bfs_batch(srcs)[0], dist)


def random_directed(rng, V, E, wmax):
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ptr = np.searchsorted(src, np.arange(V)).astype(np.uint32)
    w = rng.integers(0, wmax + 1, E).astype(np.uint32)
    return ptr, dst.astype(np.uint32), w


@pytest.mark.parametrize("flags", [0, pg.FLAG_NO_RELABEL], ids=["relabel", "no_relabel"])
def test_batch_directed(flags):
    rng = np.random.default_rng(11)
    ptr, col, w = random_directed(rng, 2000, 5000, 9)
    srcs = rng.choice(2000, 40, replace=False)
    say("case batch directed flags %d" % flags)
    with pg.Graph(ptr, col, w, config=pg.default_config(flags=flags)) as g:
        assert not g.is_symmetric()
        lv, _ = g.bfs_batch(srcs)
        dist, _ = g.sssp_batch(srcs)
        check_rows(ptr, col, w, srcs, lv, dist)


def test_batch_deep_graph_falls_back():
    # a path of 100 vertices is deeper than the levels a batch keeps as masks
    say("case batch deep path")
    V = 100
    src = np.repeat(np.arange(V), 2)[1:-1]
    dst = np.empty_like(src)
    dst[0::2] = src[0::2] + 1
    dst[1::2] = src[1::2] - 1
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ptr = np.searchsorted(src, np.arange(V)).astype(np.uint32)
    col = dst.astype(np.uint32)
    w = np.full(col.size, 3, np.uint32)
    srcs = np.array([0, 99, 50, 0], np.uint32)
    with pg.Graph(ptr, col, w) as g:
        lv, _ = g.bfs_batch(srcs)
        dist, _ = g.sssp_batch(srcs)
        check_rows(ptr, col, w, srcs, lv, dist)
        assert lv[0][99] == 100


def test_batch_sentinel_opt():
    say("case batch opt sentinel")
    f = load(os.path.join(GOLDEN, "hand_components.npz")) if os.path.exists(
        os.path.join(GOLDEN, "hand_components.npz")) else load(FIXTURES[0])
    ptr, col, w = f["ptr"], f["col"], f["w"]
    cfg = pg.default_config(sssp_sentinel=pg.SSSP_SENTINEL_OPT)
    srcs = np.arange(ptr.size, dtype=np.uint32)[:64]
    with pg.Graph(ptr, col, w, config=cfg) as g:
        dist, _ = g.sssp_batch(srcs)
        for k, s in enumerate(srcs):
            np.testing.assert_array_equal(dist[k], oracle.dijkstra(ptr, col, w, int(s), oracle.SSSP_SENTINEL_OPT))


def test_batch_argument_errors():
    say("case batch errors")
    with pg.Graph(np.array([0, 1], np.uint32), np.array([1, 0], np.uint32), np.array([2, 2], np.uint32)) as g:
        for bad in ([], list(range(65)), [2]):
            srcs = np.array([s % 2 for s in bad] if len(bad) == 65 else bad, np.uint32)
            with pytest.raises(pg.PtgError) as e:
                g.bfs_batch(srcs)
            assert e.value.code == pg.ERR_INVALID_ARG
    with pg.Graph(np.array([0, 1], np.uint32), np.array([1, 0], np.uint32)) as g:
        with pytest.raises(pg.PtgError) as e:
            g.sssp_batch(np.array([0], np.uint32))
        assert e.value.code == pg.ERR_UNSUPPORTED
        assert g.bfs_batch(np.array([0, 1], np.uint32))[0].tolist() == [[1, 2], [2, 1]]


def test_autotune_and_set_config():
    say("case autotune rmat14")
    csr = pg.gen_rmat(14, 16, seed=1).add_weights(1, 64, seed=2)
    deg = np.diff(csr.offsets.astype(np.int64))
    srcs = np.random.default_rng(3).choice(np.nonzero(deg > 0)[0], 64, replace=False).astype(np.uint32)
    with pg.Graph.from_csr(csr) as g:
        want_b, _ = g.bfs_batch(srcs)
        want_d, _ = g.sssp_batch(srcs)
        c = g.autotune(srcs)
        got = g.get_config()
        for f in ("td_to_bu", "bu_to_td", "batch_td_to_bu", "batch_grid", "batch_sssp_grid", "batch_delta",
                  "batch_chunk"):
            assert getattr(c, f) == getattr(got, f) and getattr(c, f) > 0, f
        assert c.batch_pull == got.batch_pull and c.batch_pull != 0  # -1 = push only
        for f in ("sssp_delta", "sssp_defer_degree", "sssp_defer_window"):  # 0 = automatic
            assert getattr(c, f) == getattr(got, f), f
        say("case autotune result %s" % {f: getattr(c, f) for f, _ in c._fields_})
        # results do not depend on tuning
        np.testing.assert_array_equal(g.bfs_batch(srcs)[0], want_b)
        np.testing.assert_array_equal(g.sssp_batch(srcs)[0], want_d)
        for k in (0, 63):
            np.testing.assert_array_equal(g.bfs(int(srcs[k]))[0], want_b[k])
            np.testing.assert_array_equal(g.sssp(int(srcs[k]))[0], want_d[k])
        # every tuning knob accepted; structural fields refused
        g.tune(batch_chunk=32, batch_delta=7, batch_grid=3, batch_sssp_grid=5, sssp_delta=5, td_to_bu=1.5)
        np.testing.assert_array_equal(g.sssp_batch(srcs)[0], want_d)
        np.testing.assert_array_equal(g.bfs_batch(srcs)[0], want_b)
        np.testing.assert_array_equal(g.sssp(int(srcs[3]))[0], want_d[3])
        for bad in ({"grid_blocks": 7}, {"small_degree": 3}, {"sssp_sentinel": 5}, {"td_to_bu": -1.0}):
            with pytest.raises(pg.PtgError) as e:
                g.tune(**bad)
            assert e.value.code == pg.ERR_INVALID_ARG


def test_batch_er16_and_cc():
    """Batched BFS/SSSP on a uniform random (Erdos-Renyi) graph of scale 16:
    no hubs, so nearly all work falls in the small-degree bins. Rows equal
    the single queries; CC equals the oracle's labels."""
    say("case batch er16")
    csr = pg.gen_er(16, 16, seed=5).add_weights(1, 64, seed=6)
    ptr, col, w = csr.node_pointers, csr.cols, csr.weights
    deg = np.diff(csr.offsets.astype(np.int64))
    srcs = np.random.default_rng(7).choice(np.nonzero(deg > 0)[0], 64, replace=False).astype(np.uint32)
    with pg.Graph.from_csr(csr) as g:
        assert g.is_symmetric()
        lv, _ = g.bfs_batch(srcs)
        dist, _ = g.sssp_batch(srcs)
        for k in range(0, 64, 9):
            s = int(srcs[k])
            np.testing.assert_array_equal(lv[k], g.bfs(s)[0], err_msg="bfs row %d" % k)
            np.testing.assert_array_equal(dist[k], g.sssp(s)[0], err_msg="sssp row %d" % k)
        check_rows(ptr, col, w, srcs[:2], lv[:2], dist[:2])
        lab, st = g.cc()
        np.testing.assert_array_equal(lab, oracle.cc_union_find(ptr, col))
        assert st.reached == len(np.unique(lab))
