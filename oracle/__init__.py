"""ctypes view of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always to check results against or as the timed
CPU baseline, never as the computation under test.

Every function takes a graph in the reference layout (ptGraph/main.cu:257-264):
`ptr` holds V entries and the last vertex's degree is E - ptr[V-1]. The C file
documents which reference lines each routine restates and how each is pinned.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

BFS, SSSP, CC = 0, 1, 2
SSSP_SENTINEL_SHARE = 0xFFFFFFFE  # ptGraph/main.cu:584
SSSP_SENTINEL_OPT = 0xFFFFFFFF  # ptGraph/main.cu:742

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_lib = None


def build():
    """Compile oracle.c with its Makefile (gcc)."""
    import subprocess

    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is missing: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.orc_ref_loop.argtypes = [i32, u32, u64, _u32p, _u32p, _u32p, u32, u32, i32, _u32p, _u32p, _u64p]
        L.orc_ref_loop_timed.argtypes = L.orc_ref_loop.argtypes
        L.orc_ref_loop_timed.restype = ctypes.c_int64
        L.orc_bfs_queue.argtypes = [u32, u64, _u32p, _u32p, u32, _u32p, _u32p, _u32p, _u64p]
        L.orc_dijkstra.argtypes = [u32, u64, _u32p, _u32p, _u32p, u32, u32, _u32p]
        L.orc_cc_reach_min.argtypes = [u32, u64, _u32p, _u32p, _u32p]
        L.orc_cc_union_find.argtypes = [u32, u64, _u32p, _u32p, _u32p]
        L.orc_pagerank_f64.argtypes = [u32, u64, _u32p, _u32p, ctypes.c_double, i32, _f64p]
        L.orc_floyd_warshall.argtypes = [u32, u64, _u32p, _u32p, _u32p, u32, u32, _u32p]
        L.orc_closure_min.argtypes = [u32, u64, _u32p, _u32p, _u32p]
        L.orc_certify_paths.argtypes = [u32, u64, _u32p, _u32p, _u32p, u32, u32, _u32p, i32]
        L.orc_pagerank_f64_sym.argtypes = [u32, u64, _u32p, _u32p, ctypes.c_double, i32, i32, _f64p]
        L.orc_degree_pass.argtypes = [u32, u64, _u32p, _u32p]
        L.orc_degree_pass.restype = ctypes.c_int64
        L.orc_overload_gather.argtypes = [i32, u32, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p]
        L.orc_overload_gather.restype = ctypes.c_int64
        _lib = L
    return _lib


def _p(a, t=_u32p):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _graph(ptr, col, w=None):
    ptr = np.ascontiguousarray(ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    if col.size == 0:
        col = np.zeros(1, dtype=np.uint32)
        E = 0
    else:
        E = col.size
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.uint32)
        if w.size == 0:
            w = np.zeros(1, dtype=np.uint32)
    return ptr, col, w, E


def ref_loop(algo, ptr, col, w=None, src=0, sentinel=SSSP_SENTINEL_SHARE, threads=1):
    """Reference frontier loop; returns (values, iterations, nodeSum)."""
    ptr, col, w, E = _graph(ptr, col, w)
    V = ptr.size
    out = np.empty(V, dtype=np.uint32)
    it = ctypes.c_uint32(0)
    ns = ctypes.c_uint64(0)
    rc = lib().orc_ref_loop(algo, V, E, _p(ptr), _p(col), _p(w), src, sentinel, threads,
                            _p(out), ctypes.byref(it), ctypes.byref(ns))
    if rc:
        raise ValueError("orc_ref_loop failed")
    return out, it.value, ns.value


def ref_loop_timed(algo, ptr, col, w=None, src=0, sentinel=SSSP_SENTINEL_SHARE, threads=1):
    """Reference loop timed in C (ns, setup excluded); returns (values, ns, iters, nodeSum)."""
    ptr, col, w, E = _graph(ptr, col, w)
    V = ptr.size
    out = np.empty(V, dtype=np.uint32)
    it = ctypes.c_uint32(0)
    ns = ctypes.c_uint64(0)
    t = lib().orc_ref_loop_timed(algo, V, E, _p(ptr), _p(col), _p(w), src, sentinel, threads,
                                 _p(out), ctypes.byref(it), ctypes.byref(ns))
    if t < 0:
        raise ValueError("orc_ref_loop_timed failed")
    return out, t, it.value, ns.value


def bfs(ptr, col, src):
    """Queue BFS; returns (levels, n_levels, reached, edges_traversed)."""
    ptr, col, _, E = _graph(ptr, col)
    V = ptr.size
    out = np.empty(V, dtype=np.uint32)
    lv, rc_, ed = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint64(0)
    rc = lib().orc_bfs_queue(V, E, _p(ptr), _p(col), src, _p(out), ctypes.byref(lv),
                             ctypes.byref(rc_), ctypes.byref(ed))
    if rc:
        raise ValueError("orc_bfs_queue failed")
    return out, lv.value, rc_.value, ed.value


def dijkstra(ptr, col, w, src, sentinel=SSSP_SENTINEL_SHARE):
    ptr, col, w, E = _graph(ptr, col, w)
    out = np.empty(ptr.size, dtype=np.uint32)
    rc = lib().orc_dijkstra(ptr.size, E, _p(ptr), _p(col), _p(w), src, sentinel, _p(out))
    if rc < 0:
        raise ValueError("orc_dijkstra failed")
    return out


def cc_reach_min(ptr, col):
    ptr, col, _, E = _graph(ptr, col)
    out = np.empty(ptr.size, dtype=np.uint32)
    if lib().orc_cc_reach_min(ptr.size, E, _p(ptr), _p(col), _p(out)):
        raise ValueError("orc_cc_reach_min failed")
    return out


def cc_union_find(ptr, col):
    ptr, col, _, E = _graph(ptr, col)
    out = np.empty(ptr.size, dtype=np.uint32)
    if lib().orc_cc_union_find(ptr.size, E, _p(ptr), _p(col), _p(out)):
        raise ValueError("orc_cc_union_find failed")
    return out


def pagerank_f64(ptr, col, damping=0.85, iters=20):
    ptr, col, _, E = _graph(ptr, col)
    out = np.empty(ptr.size, dtype=np.float64)
    if lib().orc_pagerank_f64(ptr.size, E, _p(ptr), _p(col), damping, iters, _p(out, _f64p)):
        raise ValueError("orc_pagerank_f64 failed")
    return out


def floyd_warshall(ptr, col, w, src, sentinel):
    ptr, col, w, E = _graph(ptr, col, w)
    out = np.empty(ptr.size, dtype=np.uint32)
    if lib().orc_floyd_warshall(ptr.size, E, _p(ptr), _p(col), _p(w), src, sentinel, _p(out)):
        raise ValueError("orc_floyd_warshall failed")
    return out


def closure_min(ptr, col):
    ptr, col, _, E = _graph(ptr, col)
    out = np.empty(ptr.size, dtype=np.uint32)
    if lib().orc_closure_min(ptr.size, E, _p(ptr), _p(col), _p(out)):
        raise ValueError("orc_closure_min failed")
    return out


def certify_paths(ptr, col, w, src, val, unreached=0xFFFFFFFF, threads=1):
    """True when `val` is exactly the BFS levels (w None) or SSSP distances from
    `src` (reference convention, weights >= 1), checked by the shortest-path
    certificate in oracle.c; raises on bad arguments or a zero weight."""
    ptr, col, w, E = _graph(ptr, col, w)
    val = np.ascontiguousarray(val, dtype=np.uint32)
    if val.size != ptr.size:
        raise ValueError("certify_paths: val has the wrong length")
    rc = lib().orc_certify_paths(ptr.size, E, _p(ptr), _p(col), _p(w), src, unreached, _p(val), threads)
    if rc < 0:
        raise ValueError("orc_certify_paths: bad arguments")
    return rc == 0


def pagerank_f64_sym(ptr, col, damping=0.85, iters=20, threads=1):
    """pagerank_f64 for a graph that is symmetric as a multigraph, pulled over
    the out-lists on `threads` threads (the caller vouches for symmetry)."""
    ptr, col, _, E = _graph(ptr, col)
    out = np.empty(ptr.size, dtype=np.float64)
    if lib().orc_pagerank_f64_sym(ptr.size, E, _p(ptr), _p(col), damping, iters, threads, _p(out, _f64p)):
        raise ValueError("orc_pagerank_f64_sym failed")
    return out


def degree_pass(ptr, E):
    """The reference's degree loop; returns (degrees, ns)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.uint32)
    deg = np.zeros(ptr.size, dtype=np.uint32)
    ns = lib().orc_degree_pass(ptr.size, E, _p(ptr), _p(deg))
    return deg, int(ns)


def overload_gather(ptr, col, degree, active, threads=20):
    """The reference's overload gather of the adjacency of `active` with its
    thread fan-out; returns (gathered edges, ns)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    degree = np.ascontiguousarray(degree, dtype=np.uint32)
    active = np.ascontiguousarray(active, dtype=np.uint32)
    d = degree[active].astype(np.uint64)
    aptr = np.concatenate([[0], np.cumsum(d)[:-1]]).astype(np.uint32) if active.size else np.zeros(1, np.uint32)
    total = int(d.sum())
    out = np.empty(max(total, 1), dtype=np.uint32)
    ns = lib().orc_overload_gather(threads, active.size, _p(degree), _p(aptr), _p(ptr),
                                   _p(active if active.size else np.zeros(1, np.uint32)), _p(col), _p(out))
    return out[:total], int(ns)
