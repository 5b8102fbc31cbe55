"""Writes the golden fixtures under tests/golden/ (run from the repo root).

The reference ships no tests, fixtures or golden vectors for this path, so the
expected outputs here come from sources independent of oracle/oracle.c:
  * hand graphs: pure-Python Floyd-Warshall / closure (multi-edges and
    self-loops kept, minimum over parallel edges) and a dense numpy PageRank;
  * R-MAT scale 8 and 10: scipy.sparse.csgraph (shortest_path unweighted for
    BFS, dijkstra for SSSP, connected_components for CC) and dense numpy
    PageRank;
  * R-MAT scale 16 (the CPU-runnable BFS workload): SHA-256 of the CSR and of
    the BFS levels from vertex 0 computed by scipy.
Output conventions follow the reference: BFS source = 1, unreached
0xFFFFFFFF (ptGraph/main.cu:1068-1076); SSSP source = 1, unreached
0xFFFFFFFE (ptGraph/main.cu:584-588); CC label = min vertex id that reaches v
(fixed point of ptGraph/gpu_kernels.cu:136-150 from value[v] = v).
PageRank: r0 = 1/V, r' = (1-d)/V + d*(sum r(u)/outdeg(u) + D/V), d = 0.85.
"""
import hashlib
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "test-world_amd"))

UNREACHED = 0xFFFFFFFF
SENTINEL = 0xFFFFFFFE
DAMPING, PR_ITERS = 0.85, 20


def edges_to_csr(V, edges):
    """edges: list of (u, v[, w]) kept in the given order per source."""
    by = [[] for _ in range(V)]
    for e in edges:
        by[e[0]].append(e)
    ptr, col, w = [], [], []
    for u in range(V):
        ptr.append(len(col))
        for e in by[u]:
            col.append(e[1])
            w.append(e[2] if len(e) > 2 else 1)
    return np.array(ptr, np.uint32), np.array(col, np.uint32), np.array(w, np.uint32)


def deg(ptr, E):
    V = ptr.size
    nxt = np.append(ptr[1:].astype(np.int64), E)
    return nxt - ptr.astype(np.int64)


def brute_paths(ptr, col, w, src):
    V, E = ptr.size, col.size
    INF = float("inf")
    D = [[0 if i == j else INF for j in range(V)] for i in range(V)]
    d = deg(ptr, E)
    for u in range(V):
        for k in range(int(ptr[u]), int(ptr[u] + d[u])):
            v = int(col[k])
            c = int(w[k]) if w is not None else 1
            if c < D[u][v]:
                D[u][v] = c
    for k in range(V):
        for i in range(V):
            for j in range(V):
                if D[i][k] + D[k][j] < D[i][j]:
                    D[i][j] = D[i][k] + D[k][j]
    return D[src]


def brute_reach_min(ptr, col):
    V, E = ptr.size, col.size
    d = deg(ptr, E)
    R = [{i} for i in range(V)]
    for u in range(V):
        for k in range(int(ptr[u]), int(ptr[u] + d[u])):
            R[u].add(int(col[k]))
    changed = True
    while changed:
        changed = False
        for u in range(V):
            new = set(R[u])
            for x in R[u]:
                new |= R[x]
            if new != R[u]:
                R[u] = new
                changed = True
    return np.array([min(u for u in range(V) if v in R[u]) for v in range(V)], np.uint32)


def dense_pagerank(ptr, col):
    V, E = ptr.size, col.size
    d = deg(ptr, E)
    M = np.zeros((V, V))
    for u in range(V):
        for k in range(int(ptr[u]), int(ptr[u] + d[u])):
            M[int(col[k]), u] += 1.0 / d[u]
    dangling = d == 0
    r = np.full(V, 1.0 / V)
    for _ in range(PR_ITERS):
        D = r[dangling].sum()
        r = (1 - DAMPING) / V + DAMPING * (M @ r + D / V)
    return r


def levels_from(dist_row, sentinel):
    out = np.empty(len(dist_row), np.uint32)
    for i, x in enumerate(dist_row):
        out[i] = sentinel if x == float("inf") or x != x else int(x) + 1
    return out


def hand_graphs():
    g = {}
    g["path6"] = (6, [(i, i + 1, 3) for i in range(5)] + [(i + 1, i, 3) for i in range(5)])
    g["star8"] = (8, [(0, i, i) for i in range(1, 8)] + [(i, 0, i) for i in range(1, 8)])
    cyc = []
    for i in range(9):
        j = (i + 1) % 9
        cyc += [(i, j, 1 + i % 4), (j, i, 1 + i % 4)]
    g["cycle9"] = (9, cyc)
    # vertex 0 isolated, triangle 1-2-3, edge 4-5, last vertex 6 isolated
    tri = [(1, 2, 5), (2, 1, 5), (2, 3, 1), (3, 2, 1), (1, 3, 9), (3, 1, 9), (4, 5, 2), (5, 4, 2)]
    g["components"] = (7, tri)
    # directed, not symmetric: min-ancestor labels differ from components
    g["directed"] = (6, [(0, 1, 1), (2, 1, 1), (3, 2, 2), (5, 3, 1), (1, 4, 7)])
    # zero weights (the reference's converter draws rand() % 5, main.cu:2596)
    g["zero_w"] = (6, [(0, 1, 0), (1, 0, 0), (1, 2, 0), (2, 1, 0), (2, 3, 4), (3, 2, 4), (0, 4, 1), (4, 0, 1),
                       (4, 3, 0), (3, 4, 0)])
    k5 = [(i, j, 1 + (i * j) % 5) for i in range(5) for j in range(5) if i != j]
    g["complete5"] = (5, k5)
    # self-loops and parallel edges left in place
    g["multi"] = (5, [(0, 0, 1), (0, 1, 7), (0, 1, 2), (1, 0, 2), (1, 2, 1), (2, 1, 1), (2, 2, 3), (1, 0, 9),
                      (3, 4, 1), (4, 3, 1)])
    return g


def write_hand():
    for name, (V, edges) in hand_graphs().items():
        ptr, col, w = edges_to_csr(V, edges)
        srcs = np.arange(V, dtype=np.uint32)
        bfs = np.stack([levels_from(brute_paths(ptr, col, None, s), UNREACHED) for s in srcs])
        sssp = np.stack([levels_from(brute_paths(ptr, col, w, s), SENTINEL) for s in srcs])
        cc = brute_reach_min(ptr, col)
        pr = dense_pagerank(ptr, col)
        np.savez(os.path.join(HERE, "hand_%s.npz" % name), ptr=ptr, col=col, w=w, srcs=srcs, bfs=bfs,
                 sssp=sssp, cc=cc, pr=pr)


def scipy_outputs(ptr, col, w, srcs):
    V, E = ptr.size, col.size
    offs = np.append(ptr, E).astype(np.int64)
    A1 = sp.csr_matrix((np.ones(E), col.astype(np.int64), offs), shape=(V, V))
    Aw = sp.csr_matrix((w.astype(np.float64), col.astype(np.int64), offs), shape=(V, V))
    hop = csg.shortest_path(A1, method="D", unweighted=True, indices=srcs)
    dist = csg.dijkstra(Aw, indices=srcs)
    bfs = np.stack([levels_from(r, UNREACHED) for r in hop])
    sssp = np.stack([levels_from(r, SENTINEL) for r in dist])
    ncomp, lab = csg.connected_components(A1, directed=False)
    mins = np.full(ncomp, V, np.int64)
    np.minimum.at(mins, lab, np.arange(V))
    cc = mins[lab].astype(np.uint32)
    return bfs, sssp, cc


def write_rmat():
    import ptgraph as pg

    for scale in (8, 10):
        g = pg.gen_rmat(scale, 16, seed=1)
        g.add_weights(1, 64, seed=2)
        ptr = g.node_pointers.copy()
        col = g.cols.copy()
        w = g.weights.copy()
        d = deg(ptr, col.size)
        rng = np.random.default_rng(3)
        srcs = np.concatenate([[0], rng.choice(np.nonzero(d > 0)[0], 7, replace=False)]).astype(np.uint32)
        bfs, sssp, cc = scipy_outputs(ptr, col, w, srcs)
        pr = dense_pagerank(ptr, col)
        np.savez_compressed(os.path.join(HERE, "rmat%d.npz" % scale), ptr=ptr, col=col, w=w, srcs=srcs, bfs=bfs,
                            sssp=sssp, cc=cc, pr=pr)


def write_rmat16():
    import ptgraph as pg

    g = pg.gen_rmat(16, 16, seed=1)
    ptr = g.node_pointers.copy()
    col = g.cols.copy()
    bfs, _, _ = scipy_outputs(ptr, col, np.ones(col.size, np.uint32), np.array([0], np.uint32))
    lv = bfs[0]
    h = hashlib.sha256()
    h.update(g.offsets.tobytes())
    h.update(col.tobytes())
    reached = lv != UNREACHED
    with open(os.path.join(HERE, "rmat16_bfs0.txt"), "w") as f:
        f.write("# R-MAT scale 16, edge factor 16, a=0.57 b=0.19 c=0.19, seed 1; BFS from vertex 0 (scipy)\n")
        f.write("V %d\nE %d\n" % (g.V, g.E))
        f.write("csr_sha256 %s\n" % h.hexdigest())
        f.write("levels_sha256 %s\n" % hashlib.sha256(lv.tobytes()).hexdigest())
        f.write("reached %d\n" % int(reached.sum()))
        f.write("edges_traversed %d\n" % int(deg(ptr, col.size)[reached].sum()))
        f.write("max_level %d\n" % int(lv[reached].max()))


if __name__ == "__main__":
    write_hand()
    write_rmat()
    write_rmat16()
    print("fixtures written to", HERE)
