"""Full-size parity on the bench graph, R-MAT scale 22 (V = 4,194,304, about
128.3 M directed edges, weights in [1, 64]), against the oracle:
  BFS   from the 8 bench sources, bit-exact against the oracle's restatement
        of the reference loop (orc_ref_loop on the host threads), and the
        shortest-path certificate (orc_certify_paths) on the same levels;
  SSSP  from the same sources, bit-exact against Dijkstra (one source per
        host thread), and the certificate;
  CC    bit-exact against union-find: every label is the smallest id of its
        component, which catches two components merged under one label;
  PageRank (d = 0.85, 20 iterations) within 1e-6 L1 of the FP64 oracle.
Batched rows (64 sources) are compared with the oracle as well."""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import ptgraph as pg

pytestmark = pytest.mark.gpu
INF = 0xFFFFFFFF
SENT = oracle.SSSP_SENTINEL_SHARE
PR_L1 = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


@pytest.fixture(scope="module")
def big():
    pg.init(0)
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()
    T = host_threads()
    csr = pg.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1, threads=T).add_weights(1, 64, seed=2, threads=T)
    g = pg.Graph.from_csr(csr)
    yield csr, g
    g.close()


def bench_sources(csr, n=8):
    """The bench's rule: seed 3, drawn among vertices of degree > 0."""
    deg = np.diff(csr.offsets.astype(np.int64))
    rng = np.random.default_rng(3)
    srcs = [int(s) for s in rng.choice(np.nonzero(deg > 0)[0], n, replace=False)]
    srcs64 = np.array(rng.choice(np.nonzero(deg > 0)[0], 64, replace=False), dtype=np.uint32)
    return srcs, srcs64


def test_fullsize_bfs_against_oracle(big):
    csr, g = big
    ptr, col = csr.node_pointers, csr.cols
    T = host_threads()
    srcs, _ = bench_sources(csr)
    for s in srcs:
        say("case fullsize bfs src %d" % s)
        lv, st = g.bfs(s)
        want, iters, node_sum = oracle.ref_loop(oracle.BFS, ptr, col, src=s, threads=T)
        np.testing.assert_array_equal(lv, want)
        assert st.reached == node_sum
        assert st.iterations == iters  # eccentricity + 1, a pinned invariant of the reference loop
        assert oracle.certify_paths(ptr, col, None, s, lv, INF, threads=T)


def test_fullsize_sssp_against_oracle(big):
    csr, g = big
    ptr, col, w = csr.node_pointers, csr.cols, csr.weights
    T = host_threads()
    srcs, _ = bench_sources(csr)
    got = {}
    for s in srcs:
        say("case fullsize sssp src %d" % s)
        got[s] = g.sssp(s)[0]
    # ctypes releases the GIL, so the Dijkstra runs overlap
    with ThreadPoolExecutor(max_workers=min(T, len(srcs))) as ex:
        want = dict(zip(srcs, ex.map(lambda s: oracle.dijkstra(ptr, col, w, s, SENT), srcs)))
    for s in srcs:
        np.testing.assert_array_equal(got[s], want[s])
        assert oracle.certify_paths(ptr, col, w, s, got[s], SENT, threads=T)


def test_fullsize_cc_against_oracle(big):
    csr, g = big
    say("case fullsize cc")
    lab, st = g.cc()
    want = oracle.cc_union_find(csr.node_pointers, csr.cols)
    np.testing.assert_array_equal(lab, want)
    assert st.reached == int(np.unique(want).size)


def test_fullsize_pagerank_against_oracle(big):
    csr, g = big
    say("case fullsize pagerank")
    r, _ = g.pagerank(0.85, 20)
    want = oracle.pagerank_f64_sym(csr.node_pointers, csr.cols, 0.85, 20, threads=host_threads())
    l1 = float(np.abs(r.astype(np.float64) - want).sum())
    say("fullsize pagerank L1 %.3e" % l1)
    assert l1 <= PR_L1


def test_fullsize_batched_rows_against_oracle(big):
    csr, g = big
    ptr, col, w = csr.node_pointers, csr.cols, csr.weights
    T = host_threads()
    _, srcs = bench_sources(csr)
    say("case fullsize batched")
    buf = pg.DeviceBuffer(4 * 64 * csr.V)
    try:
        g.bfs_batch(srcs, out=buf)
        rows = buf.to_host(np.uint32, 64 * csr.V).reshape(64, csr.V)
        for k in (0, 21, 63):
            want = oracle.ref_loop(oracle.BFS, ptr, col, src=int(srcs[k]), threads=T)[0]
            np.testing.assert_array_equal(rows[k], want)
        for k in range(64):
            assert oracle.certify_paths(ptr, col, None, int(srcs[k]), rows[k], INF, threads=T), k
        g.sssp_batch(srcs, out=buf)
        rows = buf.to_host(np.uint32, 64 * csr.V).reshape(64, csr.V)
        with ThreadPoolExecutor(max_workers=3) as ex:
            want = list(ex.map(lambda k: oracle.dijkstra(ptr, col, w, int(srcs[k]), SENT), (0, 42, 63)))
        for k, wk in zip((0, 42, 63), want):
            np.testing.assert_array_equal(rows[k], wk)
        for k in range(64):
            assert oracle.certify_paths(ptr, col, w, int(srcs[k]), rows[k], SENT, threads=T), k
    finally:
        buf.free()
