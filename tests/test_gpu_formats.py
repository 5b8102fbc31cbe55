"""The reference's on-disk formats through the HIP path: R-MAT-16 written as
.bcsr and .bwcsr (ptGraph/main.cu:98-116, :452-471, :2575-2614) and R-MAT-12
as the big-endian Java format (ptGraph/main.cu:2479-2573), read back with the
library's readers, loaded on the GPU and queried; BFS levels, SSSP distances
and CC labels equal the oracle's on the generated arrays, PageRank is within
1e-6 L1 of the FP64 oracle."""
import os
import struct
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


def check_queries(g, ptr, col, w):
    deg = np.diff(np.append(ptr.astype(np.int64), col.size))
    for s in (0, int(np.argmax(deg)), ptr.size - 1):
        np.testing.assert_array_equal(g.bfs(s)[0], oracle.bfs(ptr, col, s)[0])
        if w is not None:
            np.testing.assert_array_equal(g.sssp(s)[0], oracle.dijkstra(ptr, col, w, s))
    np.testing.assert_array_equal(g.cc()[0], oracle.cc_union_find(ptr, col))
    r, _ = g.pagerank(0.85, 20)
    assert float(np.abs(r.astype(np.float64) - oracle.pagerank_f64(ptr, col, 0.85, 20)).sum()) <= PR_L1


def test_bcsr_and_bwcsr_files_on_gpu(tmp_path):
    csr = pg.gen_rmat(16, 16, seed=1).add_weights(1, 64, seed=2)
    ptr, col, w = csr.node_pointers.copy(), csr.cols.copy(), csr.weights.copy()
    say("case formats bcsr")
    p = tmp_path / "rmat16.bcsr"
    pg.write_graph(p, csr, "bcsr")
    h = pg.read_graph(p, "bcsr")
    with pg.Graph(h.node_pointers, h.cols) as g:
        check_queries(g, ptr, col, None)
    say("case formats bwcsr")
    q = tmp_path / "rmat16.bwcsr"
    pg.write_graph(q, csr, "bwcsr")
    h = pg.read_graph(q, "bwcsr")
    np.testing.assert_array_equal(h.weights, w)
    aos = np.stack([h.cols, h.weights], axis=1)  # EdgeWithWeight {toNode, weight} (globals.cuh:2-5)
    with pg.Graph.from_aos(h.node_pointers, aos) as g:
        check_queries(g, ptr, col, w)


def test_java_big_endian_file_on_gpu(tmp_path):
    csr = pg.gen_rmat(12, 16, seed=3)
    ptr, col = csr.node_pointers, csr.cols
    say("case formats java")
    blob = struct.pack(">IQ", csr.V, csr.E) + ptr.astype(">u8").tobytes() + col.astype(">u4").tobytes()
    p = tmp_path / "rmat12.java"
    p.write_bytes(blob)
    h = pg.read_graph(p, "java")
    with pg.Graph(h.node_pointers, h.cols) as g:
        check_queries(g, ptr, col, None)
