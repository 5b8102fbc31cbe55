"""SSSP pull rounds (ptg_config.sssp_pull) through the HIP path.

On a symmetric graph with one unbounded bucket, a large round may pull
instead of push: every vertex above the frontier minimum + the smallest
weight takes min(val[u] + w) over its in-edges. The distances must equal the
oracle's Dijkstra bit for bit whether pulls never happen (-1), follow the
default rule (0), or happen in every large round (1e-9), on:
  * R-MAT-16 with one weight per undirected pair (the out-pairs serve as
    in-pairs), with and without relabeling and deferral;
  * R-MAT-14 with each direction of an edge weighted on its own and some
    edges repeated in both directions (the in-pairs are built by
    transposition and must carry w(u -> v), not w(v -> u));
  * zero weights with the 0xFFFFFFFF sentinel;
  * weights so large that a path plus an edge could wrap (pulls refused);
  * a directed graph (pulls refused).
A frontier is pulled only once it is held as byte flags (more than 16384
frontier edges per queue shard), so the graphs are sized to reach that."""
import os
import sys

import numpy as np
import pytest

import oracle
import ptgraph as pg

pytestmark = pytest.mark.gpu
SETTINGS = (-1.0, 0.0, 1e-9, 0.5, 4.0)


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


def check_settings(g, srcs, want, label, settings=SETTINGS, **tune):
    """Runs every source under every pull setting; returns pulled rounds per setting."""
    pulls = {}
    for x in settings:
        g.tune(sssp_pull=x, **tune)
        total = 0
        for s in srcs:
            say("case %s sssp_pull %g src %d" % (label, x, s))
            d, st = g.sssp(s)
            np.testing.assert_array_equal(d, want[s], err_msg="%s sssp_pull %g src %d" % (label, x, s))
            total += st.pull_rounds
        pulls[x] = total
    say("%s pulled rounds %s" % (label, pulls))
    assert pulls[-1.0] == 0
    return pulls


@pytest.mark.parametrize("flags", [0, pg.FLAG_NO_RELABEL], ids=["relabel", "no_relabel"])
def test_pull_rmat16_pair_weights(flags):
    csr = pg.gen_rmat(16, 16, seed=1).add_weights(1, 64, seed=2)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    srcs = sources(csr.offsets, 4, 21)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w, config=pg.default_config(flags=flags)) as g:
        assert g.is_symmetric()
        g.tune(sssp_pull=-1.0)
        g.sssp(srcs[0])
        dev = g.device_bytes
        for ddeg in (0, 0xFFFFFFFF):  # deferral by default / off
            # sweep rounds off: they would build the binned layout, whose
            # bytes this test does not count
            pulls = check_settings(g, srcs, want, "rmat16 pair flags %d defer %d" % (flags, ddeg),
                                   sssp_defer_degree=ddeg, sssp_sweep=-1.0)
            assert pulls[1e-9] > 0
        # each reverse edge has the same weight: the out-pairs serve as
        # in-pairs (only the list of long lists, 8 bytes a vertex, is added)
        assert g.device_bytes - dev <= 8 * (ptr.size + 1)


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


def test_pull_directed_weights_and_repeats():
    ptr, col, w = directed_weights_with_repeats(14, 5)
    srcs = sources(np.append(ptr, col.size), 4, 22)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w) as g:
        assert g.is_symmetric()
        g.tune(sssp_pull=-1.0)
        g.sssp(srcs[0])
        dev = g.device_bytes
        pulls = check_settings(g, srcs, want, "rmat14 directed weights")
        assert pulls[1e-9] > 0
        # the in-pairs were built: 8 bytes per edge
        assert g.device_bytes >= dev + 8 * col.size


def test_pull_zero_weights_opt_sentinel():
    csr = pg.gen_rmat(14, 16, seed=3).add_weights(0, 4, seed=4)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    srcs = sources(csr.offsets, 3, 23)
    want = {s: oracle.dijkstra(ptr, col, w, s, oracle.SSSP_SENTINEL_OPT) for s in srcs}
    cfg = pg.default_config(sssp_sentinel=pg.SSSP_SENTINEL_OPT)
    with pg.Graph(ptr, col, w, config=cfg) as g:
        pulls = check_settings(g, srcs, want, "rmat14 weights [0, 4] opt sentinel", settings=(-1.0, 0.0, 1e-9))
        assert pulls[1e-9] > 0


def test_pull_refused_when_values_could_wrap():
    # V = 16384 and weights up to 10^6: a path plus an edge may pass 2^32
    csr = pg.gen_rmat(14, 16, seed=3).add_weights(1000, 1000000, seed=4)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    srcs = sources(csr.offsets, 2, 24)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w) as g:
        pulls = check_settings(g, srcs, want, "rmat14 wide weights", settings=(-1.0, 1e-9))
        assert pulls[1e-9] == 0


def test_pull_refused_on_directed_graph():
    rng = np.random.default_rng(6)
    V, E = 20000, 600000
    a, b = rng.integers(0, V, E), rng.integers(0, V, E)
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    ptr = np.searchsorted(a, np.arange(V)).astype(np.uint32)
    col = b.astype(np.uint32)
    w = rng.integers(1, 65, E).astype(np.uint32)
    srcs = sources(np.append(ptr, E), 2, 25)
    want = {s: oracle.dijkstra(ptr, col, w, s) for s in srcs}
    with pg.Graph(ptr, col, w) as g:
        assert not g.is_symmetric()
        pulls = check_settings(g, srcs, want, "directed", settings=(-1.0, 1e-9))
        assert pulls[1e-9] == 0
