"""Full-size checks (R-MAT scale 22, the bench graph) through certificates
that do not need the oracle to finish at this size:
  BFS   every edge joins levels that differ by at most 1 (both reached, as
        the graph is symmetric) and every reached vertex but the source has
        a neighbour one level lower; unreached vertices have no reached
        neighbour;
  SSSP  no edge can improve a distance (dist[v] <= dist[u] + w) and every
        reached vertex but the source has a tight predecessor
        (dist[u] + w == dist[v]) - together, exact shortest distances;
  CC    labels are equal across every edge and each label is the smallest
        id with that label;
  PageRank  ranks sum to 1 and one more iteration in FP64 moves them by
        less than the contraction bound.
Batched rows are compared with the single queries."""
import os
import sys

import numpy as np
import pytest

import ptgraph as pg

pytestmark = pytest.mark.gpu
INF = 0xFFFFFFFF


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def big():
    pg.init(0)
    csr = pg.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1, threads=16).add_weights(1, 64, seed=2, threads=16)
    off = csr.offsets.astype(np.int64)
    src_of_edge = np.repeat(np.arange(csr.V, dtype=np.uint32), np.diff(off))
    g = pg.Graph.from_csr(csr)
    yield csr, src_of_edge, g
    g.close()


def sources(csr, n):
    deg = np.diff(csr.offsets.astype(np.int64))
    return [int(s) for s in np.random.default_rng(3).choice(np.nonzero(deg > 0)[0], n, replace=False)]


def check_bfs(lv, s, src_of_edge, col):
    a = lv[src_of_edge].astype(np.int64)
    b = lv[col].astype(np.int64)
    ra, rb = a != INF, b != INF
    assert np.array_equal(ra, rb)  # symmetric graph: an edge never leaves the reached set
    assert np.abs(a[ra] - b[ra]).max(initial=0) <= 1
    assert lv[s] == 1
    # every reached vertex but the source has a neighbour one level lower
    has_parent = np.zeros(lv.size, bool)
    m = ra & (b == a - 1)
    has_parent[src_of_edge[m]] = True
    reached = lv != INF
    reached[s] = False
    assert has_parent[reached].all()


def check_sssp(d, s, src_of_edge, col, w):
    du = d[src_of_edge].astype(np.int64)
    dv = d[col].astype(np.int64)
    wt = w.astype(np.int64)
    ru = d[src_of_edge] != 0xFFFFFFFE
    assert (dv[ru] <= du[ru] + wt[ru]).all()  # no edge improves
    assert d[s] == 1
    tight = ru & (du + wt == dv)
    has_pred = np.zeros(d.size, bool)
    has_pred[col[tight]] = True
    reached = d != 0xFFFFFFFE
    reached[s] = False
    assert has_pred[reached].all()


def test_fullsize_bfs_sssp_certificates(big):
    csr, soe, g = big
    col, w = csr.cols, csr.weights
    for s in sources(csr, 3):
        say("case fullsize src %d" % s)
        lv, st = g.bfs(s)
        check_bfs(lv, s, soe, col)
        assert st.reached == int((lv != INF).sum())
        d, _ = g.sssp(s)
        check_sssp(d, s, soe, col, w)


def test_fullsize_cc_and_pagerank(big):
    csr, soe, g = big
    say("case fullsize cc")
    lab, st = g.cc()
    assert np.array_equal(lab[soe], lab[csr.cols])
    assert (lab <= np.arange(csr.V)).all()
    assert np.array_equal(lab[lab], lab)  # the label is itself labelled with itself (the minimum)
    assert st.reached == int(np.unique(lab).size)
    say("case fullsize pagerank")
    r, _ = g.pagerank(0.85, 20)
    r64 = r.astype(np.float64)
    assert abs(r64.sum() - 1.0) < 1e-5
    # one FP64 iteration from r: with damping 0.85 the 20th iterate is within
    # 0.85^20 * 2 of the fixed point in L1, so one more step moves it by less
    off = csr.offsets.astype(np.int64)
    deg = np.diff(off)
    contrib = np.where(deg > 0, r64 / np.maximum(deg, 1), 0.0)
    dang = r64[deg == 0].sum()
    nxt = np.bincount(csr.cols, weights=contrib[soe], minlength=csr.V)
    nxt = 0.15 / csr.V + 0.85 * (nxt + dang / csr.V)
    assert np.abs(nxt - r64).sum() < 2 * 0.85 ** 20 + 1e-5


def test_fullsize_batched_rows(big):
    csr, soe, g = big
    srcs = np.array(sources(csr, 64), np.uint32)
    say("case fullsize batched")
    buf = pg.DeviceBuffer(4 * 64 * csr.V)
    try:
        g.bfs_batch(srcs, out=buf)
        rows = buf.to_host(np.uint32, 64 * csr.V).reshape(64, csr.V)
        for k in (0, 21, 63):
            np.testing.assert_array_equal(rows[k], g.bfs(int(srcs[k]))[0])
        check_bfs(rows[5], int(srcs[5]), soe, csr.cols)
        g.sssp_batch(srcs, out=buf)
        rows = buf.to_host(np.uint32, 64 * csr.V).reshape(64, csr.V)
        for k in (0, 42, 63):
            np.testing.assert_array_equal(rows[k], g.sssp(int(srcs[k]))[0])
        check_sssp(rows[7], int(srcs[7]), soe, csr.cols, csr.weights)
    finally:
        buf.free()
