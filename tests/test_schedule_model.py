"""CPU model of the single-query SSSP schedule: one unbounded bucket, and
vertices of degree > defer_deg waiting in the frontier while their value is
at or above the frontier's smallest value + window.

The model runs the rounds synchronously on small graphs (numpy, no GPU) and
checks the two properties the device kernels rely on: the values equal
Dijkstra's for every degree/window setting (only the order of relaxations
changes), and every round makes progress (the vertex holding the frontier
minimum is always expanded), so the number of rounds stays bounded."""
import numpy as np
import pytest

import oracle
import ptgraph as pg

SENTINEL = 0xFFFFFFFE


def schedule_rounds(off, col, w, src, defer_deg, window):
    """Returns (values, rounds, scanned edges) of the deferred schedule."""
    V = off.size - 1
    deg = np.diff(off)
    val = np.full(V, SENTINEL, dtype=np.int64)
    val[src] = 1
    frontier = np.array([src], dtype=np.int64)
    rounds = scanned = 0
    while frontier.size:
        rounds += 1
        assert rounds <= 4 * V + 8, "no progress"
        fmin = val[frontier].min()
        wait = (deg[frontier] > defer_deg) & (val[frontier] >= fmin + window)
        active = frontier[~wait]
        assert active.size > 0, "the frontier minimum must always be expanded"
        nxt = set(frontier[wait].tolist())
        base = val[active].copy()  # values as read at the start of the round
        for u, bu in zip(active.tolist(), base.tolist()):
            for e in range(off[u], off[u + 1]):
                scanned += 1
                t, nv = col[e], bu + w[e]
                if nv < val[t]:
                    val[t] = nv
                    nxt.add(t)
        frontier = np.array(sorted(nxt), dtype=np.int64)
    return val.astype(np.uint32), rounds, scanned


@pytest.mark.parametrize("defer_deg,window", [(0xFFFFFFFF, 0), (16, 1), (16, 2), (4, 1), (64, 4), (0, 1)])
def test_deferred_schedule_matches_dijkstra(defer_deg, window):
    g = pg.gen_rmat(10, 16, seed=1).add_weights(1, 64, seed=2)
    off = g.offsets.astype(np.int64)
    col, w = g.cols.astype(np.int64), g.weights.astype(np.int64)
    deg = np.diff(off)
    srcs = [int(s) for s in np.random.default_rng(7).choice(np.nonzero(deg > 0)[0], 3, replace=False)]
    for s in srcs:
        got, rounds, _ = schedule_rounds(off, col, w, s, defer_deg, window)
        np.testing.assert_array_equal(got, oracle.dijkstra(g.node_pointers, g.cols, g.weights, s))
        assert rounds >= 1


def test_deferral_cuts_rescans():
    """On R-MAT, deferring high-degree vertices scans fewer edges than
    relaxing every improvement at once (the reason for the schedule)."""
    g = pg.gen_rmat(12, 16, seed=1).add_weights(1, 64, seed=2)
    off = g.offsets.astype(np.int64)
    col, w = g.cols.astype(np.int64), g.weights.astype(np.int64)
    s = int(np.argmax(np.diff(off) > 0))
    _, r_off, scanned_off = schedule_rounds(off, col, w, s, 0xFFFFFFFF, 0)
    val, r_on, scanned_on = schedule_rounds(off, col, w, s, 64, 2)
    np.testing.assert_array_equal(val, oracle.dijkstra(g.node_pointers, g.cols, g.weights, s))
    assert scanned_on < scanned_off
