"""ptg_pagerank_noalloc (per-iteration buffers in a fixed per-GPU device
array, no allocation) against ptg_pagerank on the same graphs: bit-identical
ranks on two calls in a row, the second on a smaller graph, and against the
FP64 oracle within 1e-6 L1; sizes beyond the array and non-device outputs
are refused."""
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


def ranks_noalloc(g, iters=20):
    buf = pg.DeviceBuffer(4 * g.V)
    try:
        g.pagerank_noalloc(buf, 0.85, iters)
        return buf.to_host(np.float32, g.V)
    finally:
        buf.free()


@pytest.mark.parametrize("flags", [0, pg.FLAG_NO_RELABEL], ids=["relabel", "no_relabel"])
def test_noalloc_equals_pagerank_large_then_small(flags):
    big = pg.gen_rmat(16, 16, seed=1)
    small = pg.gen_rmat(10, 8, seed=7)
    cfg = pg.default_config(flags=flags)
    for csr in (big, small):  # the second call reuses the array with a smaller graph
        say("case noalloc V=%d flags %d" % (csr.V, flags))
        with pg.Graph.from_csr(csr, weighted=False, config=cfg) as g:
            got = ranks_noalloc(g)
            want, _ = g.pagerank(0.85, 20)
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
            l1 = float(np.abs(got.astype(np.float64) - oracle.pagerank_f64(csr.node_pointers, csr.cols)).sum())
            assert l1 <= PR_L1
            np.testing.assert_array_equal(ranks_noalloc(g, 3).view(np.uint32), g.pagerank(0.85, 3)[0].view(np.uint32))


def test_noalloc_directed_needs_in_edges():
    rng = np.random.default_rng(3)
    V, E = 2000, 9000
    src = np.sort(rng.integers(0, V, E))
    dst = rng.integers(0, V, E).astype(np.uint32)
    ptr = np.searchsorted(src, np.arange(V)).astype(np.uint32)
    say("case noalloc directed")
    with pg.Graph(ptr, dst) as g:
        assert not g.is_symmetric()
        buf = pg.DeviceBuffer(4 * V)
        try:
            with pytest.raises(pg.PtgError) as e:
                g.pagerank_noalloc(buf)
            assert e.value.code == pg.ERR_UNSUPPORTED
            g.tune(pr_layout=pg.PR_BINNED)
            want, _ = g.pagerank(0.85, 20)  # the binned layout builds no in-edges
            with pytest.raises(pg.PtgError) as e:
                g.pagerank_noalloc(buf)
            assert e.value.code == pg.ERR_UNSUPPORTED
            g.tune(pr_layout=pg.PR_PULL)
            pulled, _ = g.pagerank(0.85, 20)  # the in-list pull builds the in-edges
            np.testing.assert_array_equal(pulled.view(np.uint32), want.view(np.uint32))
            g.pagerank_noalloc(buf)
            np.testing.assert_array_equal(buf.to_host(np.uint32, V), want.view(np.uint32))
        finally:
            buf.free()


def test_noalloc_refuses_host_output_and_oversize():
    csr = pg.gen_rmat(8, 4, seed=1)
    say("case noalloc refusals")
    with pg.Graph.from_csr(csr, weighted=False) as g:
        host = np.zeros(g.V, np.float32)
        with pytest.raises(pg.PtgError) as e:
            g.pagerank_noalloc(host)
        assert e.value.code == pg.ERR_INVALID_ARG
    V = 9 << 20  # 2 x 4V bytes of contributions alone exceed the 64 MiB array
    with pg.Graph(np.zeros(V, np.uint32), np.zeros(0, np.uint32)) as g:
        buf = pg.DeviceBuffer(4 * V)
        try:
            with pytest.raises(pg.PtgError) as e:
                g.pagerank_noalloc(buf)
            assert e.value.code == pg.ERR_BUDGET
        finally:
            buf.free()
