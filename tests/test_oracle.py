"""Pins the CPU oracle (oracle/oracle.c) before results are checked against it.

Each family of the oracle is checked against sources it does not share code
with: the committed golden fixtures (pure-Python brute force, scipy.csgraph,
dense numpy PageRank; tests/golden/make_golden.py), the other oracle families,
the C brute force, and closed forms. The reference itself ships no tests or
golden vectors for this path, and is CUDA (no nvcc here), so it cannot be run.
"""
import glob
import hashlib
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HAND = sorted(glob.glob(os.path.join(GOLDEN, "hand_*.npz")))
RMAT = [os.path.join(GOLDEN, "rmat8.npz"), os.path.join(GOLDEN, "rmat10.npz")]
SENT = oracle.SSSP_SENTINEL_SHARE


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def is_symmetric(ptr, col):
    V, E = ptr.size, col.size
    nxt = np.append(ptr[1:], E)
    pairs = set()
    for u in range(V):
        for k in range(ptr[u], nxt[u]):
            pairs.add((u, int(col[k])))
    return all((v, u) in pairs for (u, v) in pairs)


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()


@pytest.mark.parametrize("path", HAND + RMAT, ids=os.path.basename)
def test_bfs_families_match_golden(path):
    f = load(path)
    for i, s in enumerate(f["srcs"]):
        want = f["bfs"][i]
        for t in (1, 3):
            got, iters, node_sum = oracle.ref_loop(oracle.BFS, f["ptr"], f["col"], src=int(s), threads=t)
            np.testing.assert_array_equal(got, want)
            reached = want != 0xFFFFFFFF
            # pinned invariants of the reference loop: iter = ecc + 1, nodeSum = |reached|
            assert iters == int(want[reached].max())
            assert node_sum == int(reached.sum())
        lv, nlev, nreach, _ = oracle.bfs(f["ptr"], f["col"], int(s))
        np.testing.assert_array_equal(lv, want)
        assert nlev == int(want[want != 0xFFFFFFFF].max())


@pytest.mark.parametrize("path", HAND + RMAT, ids=os.path.basename)
def test_sssp_families_match_golden(path):
    f = load(path)
    for i, s in enumerate(f["srcs"]):
        want = f["sssp"][i]
        got, _, _ = oracle.ref_loop(oracle.SSSP, f["ptr"], f["col"], f["w"], src=int(s), threads=2)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(oracle.dijkstra(f["ptr"], f["col"], f["w"], int(s)), want)
        if f["ptr"].size <= 64:
            np.testing.assert_array_equal(oracle.floyd_warshall(f["ptr"], f["col"], f["w"], int(s), SENT), want)


@pytest.mark.parametrize("path", HAND + RMAT, ids=os.path.basename)
def test_cc_families_match_golden(path):
    f = load(path)
    want = f["cc"]
    got, _, _ = oracle.ref_loop(oracle.CC, f["ptr"], f["col"], threads=2)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(oracle.cc_reach_min(f["ptr"], f["col"]), want)
    if f["ptr"].size <= 64:
        np.testing.assert_array_equal(oracle.closure_min(f["ptr"], f["col"]), want)
    if is_symmetric(f["ptr"], f["col"]):
        np.testing.assert_array_equal(oracle.cc_union_find(f["ptr"], f["col"]), want)


@pytest.mark.parametrize("path", HAND + RMAT, ids=os.path.basename)
def test_pagerank_f64_matches_golden(path):
    f = load(path)
    got = oracle.pagerank_f64(f["ptr"], f["col"], 0.85, 20)
    np.testing.assert_allclose(got, f["pr"], rtol=1e-12, atol=1e-15)
    assert abs(got.sum() - 1.0) < 1e-12  # mass is conserved with dangling redistribution


def test_directed_fixture_distinguishes_cc_semantics():
    # On a non-symmetric graph the reference's cc_kernel fixed point (min id
    # that reaches v) differs from undirected components; the fixture holds
    # the former, so a union-find oracle would fail here.
    f = load(os.path.join(GOLDEN, "hand_directed.npz"))
    assert not is_symmetric(f["ptr"], f["col"])
    assert not np.array_equal(oracle.cc_union_find(f["ptr"], f["col"]), f["cc"])


def test_pagerank_dangling_term_matters():
    # Dropping D/V (a plausible mistake) changes results on graphs with
    # isolated vertices, so the golden PageRank pins the dangling rule.
    f = load(os.path.join(GOLDEN, "hand_components.npz"))
    ptr, col = f["ptr"], f["col"]
    V, E = ptr.size, col.size
    deg = np.append(ptr[1:], E).astype(np.int64) - ptr
    r = np.full(V, 1.0 / V)
    for _ in range(20):
        acc = np.zeros(V)
        for u in range(V):
            for k in range(ptr[u], ptr[u] + deg[u]):
                acc[col[k]] += r[u] / deg[u]
        r = 0.15 / V + 0.85 * acc
    assert np.abs(r - f["pr"]).sum() > 1e-3


def path_graph(n):
    ptr, col = [], []
    for i in range(n):
        ptr.append(len(col))
        if i > 0:
            col.append(i - 1)
        if i + 1 < n:
            col.append(i + 1)
    return np.array(ptr, np.uint32), np.array(col, np.uint32)


def cycle_graph(n):
    ptr = np.arange(0, 2 * n, 2, dtype=np.uint32)
    col = np.array([x for i in range(n) for x in ((i - 1) % n, (i + 1) % n)], np.uint32)
    return ptr, col


def test_closed_forms():
    ptr, col = path_graph(50)
    for s in (0, 17, 49):
        lv, _, _ = oracle.ref_loop(oracle.BFS, ptr, col, src=s)
        np.testing.assert_array_equal(lv, np.abs(np.arange(50) - s) + 1)
        w = np.full(col.size, 3, np.uint32)
        d, _, _ = oracle.ref_loop(oracle.SSSP, ptr, col, w, src=s)
        np.testing.assert_array_equal(d, 3 * np.abs(np.arange(50) - s) + 1)
    ptr, col = cycle_graph(31)
    for s in (0, 5):
        lv, _, _ = oracle.ref_loop(oracle.BFS, ptr, col, src=s)
        k = np.abs(np.arange(31) - s)
        np.testing.assert_array_equal(lv, np.minimum(k, 31 - k) + 1)
    # regular graphs keep the uniform rank; an edgeless graph stays at 1/V
    np.testing.assert_allclose(oracle.pagerank_f64(ptr, col), np.full(31, 1 / 31), rtol=1e-13)
    n = 7
    kptr = np.arange(0, n * (n - 1), n - 1, dtype=np.uint32)
    kcol = np.array([j for i in range(n) for j in range(n) if j != i], np.uint32)
    np.testing.assert_allclose(oracle.pagerank_f64(kptr, kcol), np.full(n, 1 / n), rtol=1e-13)
    eptr = np.zeros(9, np.uint32)
    np.testing.assert_allclose(oracle.pagerank_f64(eptr, np.zeros(0, np.uint32)), np.full(9, 1 / 9), rtol=1e-13)
    # CC on a path: a single component labelled 0
    lab, _, _ = oracle.ref_loop(oracle.CC, *path_graph(20))
    assert (lab == 0).all()


def random_graph(rng, V, p, symmetric, wmax):
    A = rng.random((V, V)) < p
    if symmetric:
        A = A | A.T
    ptr, col, w = [], [], []
    W = rng.integers(0, wmax + 1, (V, V))
    if symmetric:
        W = np.minimum(W, W.T)
    for u in range(V):
        ptr.append(len(col))
        for v in np.nonzero(A[u])[0]:
            col.append(v)
            w.append(W[u, v])
    return np.array(ptr, np.uint32), np.array(col, np.uint32), np.array(w, np.uint32)


@pytest.mark.parametrize("seed", range(12))
def test_random_graphs_all_families_agree(seed):
    rng = np.random.default_rng(seed)
    V = int(rng.integers(1, 48))
    ptr, col, w = random_graph(rng, V, float(rng.uniform(0.02, 0.2)), bool(seed % 2), 6)
    for s in range(0, V, max(1, V // 5)):
        lv, _, _ = oracle.ref_loop(oracle.BFS, ptr, col, src=s, threads=1 + seed % 3)
        np.testing.assert_array_equal(lv, oracle.bfs(ptr, col, s)[0])
        np.testing.assert_array_equal(lv, oracle.floyd_warshall(ptr, col, None, s, 0xFFFFFFFF))
        for sent in (oracle.SSSP_SENTINEL_SHARE, oracle.SSSP_SENTINEL_OPT):
            d, _, _ = oracle.ref_loop(oracle.SSSP, ptr, col, w, src=s, sentinel=sent, threads=2)
            np.testing.assert_array_equal(d, oracle.dijkstra(ptr, col, w, s, sent))
            np.testing.assert_array_equal(d, oracle.floyd_warshall(ptr, col, w, s, sent))
    lab, _, _ = oracle.ref_loop(oracle.CC, ptr, col, threads=2)
    np.testing.assert_array_equal(lab, oracle.cc_reach_min(ptr, col))
    np.testing.assert_array_equal(lab, oracle.closure_min(ptr, col))


def test_rmat16_golden_hashes():
    import ptgraph as pg

    meta = {}
    with open(os.path.join(GOLDEN, "rmat16_bfs0.txt")) as fh:
        for line in fh:
            if not line.startswith("#"):
                k, v = line.split()
                meta[k] = v
    g = pg.gen_rmat(16, 16, seed=1)
    h = hashlib.sha256()
    h.update(g.offsets.tobytes())
    h.update(g.cols.tobytes())
    assert h.hexdigest() == meta["csr_sha256"]
    lv, iters, node_sum = oracle.ref_loop(oracle.BFS, g.node_pointers, g.cols, src=0, threads=4)
    assert hashlib.sha256(lv.tobytes()).hexdigest() == meta["levels_sha256"]
    assert node_sum == int(meta["reached"])
    assert iters == int(meta["max_level"])
    _, _, reached, edges = oracle.bfs(g.node_pointers, g.cols, 0)
    assert reached == int(meta["reached"]) and edges == int(meta["edges_traversed"])


# --- family 4: full-size helpers -------------------------------------------

def planted_mistakes(rng, val, src, unreached, step):
    """Copies of a correct `val` with one plausible mistake each."""
    reached = np.nonzero((val != unreached) & (np.arange(val.size) != src))[0]
    unr = np.nonzero(val == unreached)[0]
    out = []
    if reached.size:
        for delta in (step, -1):
            v = int(rng.choice(reached))
            b = val.copy()
            b[v] = np.uint32(int(b[v]) + delta)
            out.append(("shift %+d at %d" % (delta, v), b))
        b = val.copy()
        b[int(rng.choice(reached))] = unreached
        out.append(("dropped", b))
    if unr.size:
        b = val.copy()
        b[int(rng.choice(unr))] = np.uint32(int(val[val != unreached].max()) + 1)
        out.append(("unreachable reached", b))
    b = val.copy()
    b[src] = 2
    out.append(("source", b))
    return out


@pytest.mark.parametrize("seed", range(10))
def test_certificate_accepts_exact_and_rejects_planted(seed):
    rng = np.random.default_rng(100 + seed)
    V = int(rng.integers(2, 60))
    ptr, col, w = random_graph(rng, V, float(rng.uniform(0.03, 0.2)), bool(seed % 2), 9)
    w = w + 1  # the certificate needs weights >= 1
    for s in range(0, V, max(1, V // 4)):
        lv = oracle.bfs(ptr, col, s)[0]
        assert oracle.certify_paths(ptr, col, None, s, lv, threads=1 + seed % 4)
        for name, bad in planted_mistakes(rng, lv, s, 0xFFFFFFFF, 1):
            assert not oracle.certify_paths(ptr, col, None, s, bad, threads=2), name
        d = oracle.dijkstra(ptr, col, w, s, SENT)
        assert oracle.certify_paths(ptr, col, w, s, d, SENT, threads=3)
        for name, bad in planted_mistakes(rng, d, s, SENT, 1):
            assert not oracle.certify_paths(ptr, col, w, s, bad, SENT), name


def test_certificate_refuses_zero_weights():
    ptr, col = path_graph(4)
    with pytest.raises(ValueError):
        oracle.certify_paths(ptr, col, np.zeros(col.size, np.uint32), 0, np.arange(1, 5, dtype=np.uint32))


def test_certificate_on_golden_fixtures():
    for path in HAND + RMAT:
        f = load(path)
        for i, s in enumerate(f["srcs"]):
            assert oracle.certify_paths(f["ptr"], f["col"], None, int(s), f["bfs"][i], threads=2)
            if (f["w"] > 0).all():
                assert oracle.certify_paths(f["ptr"], f["col"], f["w"], int(s), f["sssp"][i], SENT, threads=2)


@pytest.mark.parametrize("path", HAND + RMAT, ids=os.path.basename)
def test_pagerank_sym_equals_push_form(path):
    f = load(path)
    ptr, col = f["ptr"], f["col"]
    if not is_symmetric(ptr, col):
        pytest.skip("the pull over out-lists needs a symmetric graph")
    want = oracle.pagerank_f64(ptr, col, 0.85, 20)
    for t in (1, 3, 8):
        np.testing.assert_allclose(oracle.pagerank_f64_sym(ptr, col, 0.85, 20, t), want, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(oracle.pagerank_f64_sym(ptr, col, 0.85, 20, 2), f["pr"], rtol=1e-12, atol=1e-16)


def test_pagerank_sym_multigraph():
    # equal repeat counts both ways: 0-1 twice, 0-2 once
    ptr = np.array([0, 3, 5], np.uint32)
    col = np.array([1, 2, 1, 0, 0, 0], np.uint32)
    np.testing.assert_allclose(oracle.pagerank_f64_sym(ptr, col, 0.85, 20, 2), oracle.pagerank_f64(ptr, col, 0.85, 20),
                               rtol=1e-13)


def test_degree_pass_and_overload_gather():
    import ptgraph as pg

    g = pg.gen_rmat(12, 16, seed=1)
    ptr, col = g.node_pointers, g.cols
    E = col.size
    deg, ns = oracle.degree_pass(ptr, E)
    np.testing.assert_array_equal(deg, np.diff(np.append(ptr.astype(np.int64), E)))
    assert ns >= 0
    rng = np.random.default_rng(5)
    off = np.append(ptr.astype(np.int64), E)
    for n, T in ((0, 20), (7, 20), (49, 20), (50, 20), (1000, 20), (3000, 3)):
        active = np.sort(rng.choice(ptr.size, n, replace=False)).astype(np.uint32)
        got, _ = oracle.overload_gather(ptr, col, deg, active, threads=T)
        want = np.concatenate([col[off[v]:off[v + 1]] for v in active]) if n else np.zeros(0, np.uint32)
        np.testing.assert_array_equal(got, want)
