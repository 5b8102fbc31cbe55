"""Python view of libptgraph.so, the C-ABI declared in include/ptg.h.

The graph path itself is C++ and HIP; this module only marshals numpy arrays
through ctypes so that tests and the benchmark drive the same entry points a
C/C++ caller of the reference's drivers would (ptGraph/main.cu:24-56).

If libptgraph.so is missing this module raises at load time: there is no
Python or CPU fallback for any query.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libptgraph.so")

OK = 0
ERR_NO_DEVICE = -1
ERR_INVALID_ARG = -2
ERR_INVALID_GRAPH = -3
ERR_OUT_OF_MEMORY = -4
ERR_BUDGET = -5
ERR_HIP = -6
ERR_UNSUPPORTED = -7
ERR_REJECTED = -8
ERR_IO = -9

FLAG_NO_VALIDATE = 0x1
FLAG_TIMING = 0x2
FLAG_NO_RELABEL = 0x4
FLAG_STREAM = 0x8
TUNE_BFS, TUNE_SSSP, TUNE_BFS_BATCH, TUNE_SSSP_BATCH, TUNE_STREAM, TUNE_PAGERANK = 0x1, 0x2, 0x4, 0x8, 0x10, 0x20
PR_AUTO, PR_PULL, PR_BINNED = 0, 1, 2  # Config.pr_layout
Q_BFS, Q_SSSP, Q_CC, Q_PAGERANK = 0, 1, 2, 3

BFS_UNREACHED = 0xFFFFFFFF  # ptGraph/main.cu:1068
SSSP_SENTINEL_SHARE = 0xFFFFFFFE  # ptGraph/main.cu:584
SSSP_SENTINEL_OPT = 0xFFFFFFFF  # ptGraph/main.cu:742


class PtgError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        L = lib()
        msg = "%s failed: %s (%d): %s" % (
            where,
            L.ptg_strerror(code).decode(),
            code,
            L.ptg_last_error().decode(errors="replace"),
        )
        super().__init__(msg)


class EdgeWithWeight(ctypes.Structure):
    """ptGraph/globals.cuh:2-5."""

    _fields_ = [("toNode", ctypes.c_uint32), ("weight", ctypes.c_uint32)]


class Config(ctypes.Structure):
    _fields_ = [
        ("device_budget_bytes", ctypes.c_uint64),
        ("chunk_bytes", ctypes.c_uint64),
        ("sssp_sentinel", ctypes.c_uint32),
        ("grid_blocks", ctypes.c_uint32),
        ("small_degree", ctypes.c_uint32),
        ("large_degree", ctypes.c_uint32),
        ("td_to_bu", ctypes.c_float),
        ("bu_to_td", ctypes.c_float),
        ("sssp_delta", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("batch_delta", ctypes.c_uint32),
        ("batch_grid", ctypes.c_uint32),
        ("batch_td_to_bu", ctypes.c_float),
        ("batch_chunk", ctypes.c_uint32),
        ("order_hub_degree", ctypes.c_uint32),
        ("sssp_defer_degree", ctypes.c_uint32),
        ("sssp_defer_window", ctypes.c_uint32),
        ("stream_pinned", ctypes.c_uint32),
        ("batch_sssp_grid", ctypes.c_uint32),
        ("sssp_pull", ctypes.c_float),
        ("batch_pull", ctypes.c_float),
        ("pr_block", ctypes.c_uint32),
        ("pr_layout", ctypes.c_uint32),
        ("sssp_sweep", ctypes.c_float),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("reached", ctypes.c_uint32),
        ("edges_traversed", ctypes.c_uint64),
        ("frontier_sum", ctypes.c_uint64),
        ("kernel_ns", ctypes.c_uint64),
        ("host_link_bytes", ctypes.c_uint64),
        ("host_link_ns", ctypes.c_uint64),
        ("main_kernel_ns", ctypes.c_uint64),
        ("main_launches", ctypes.c_uint32),
        ("chunk_loads", ctypes.c_uint32),
        ("chunk_hits", ctypes.c_uint32),
        ("pull_rounds", ctypes.c_uint32),
        ("sweep_rounds", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("state_budget_bytes", ctypes.c_uint64),
        ("max_batch", ctypes.c_uint32),
        ("max_inflight", ctypes.c_uint32),
        ("coalesce", ctypes.c_uint32),
        ("queue_cap", ctypes.c_uint32),
        ("min_batch", ctypes.c_uint32),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("graph", ctypes.c_uint32),
        ("kind", ctypes.c_uint32),
        ("source", ctypes.c_uint32),
        ("iterations", ctypes.c_uint32),
        ("damping", ctypes.c_float),
        ("pad", ctypes.c_uint32),
        ("out", ctypes.c_void_p),
        ("tag", ctypes.c_uint64),
    ]


class Completion(ctypes.Structure):
    _fields_ = [
        ("tag", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("batch_size", ctypes.c_uint32),
        ("checksum", ctypes.c_uint64),
        ("submit_ns", ctypes.c_uint64),
        ("start_ns", ctypes.c_uint64),
        ("done_ns", ctypes.c_uint64),
    ]


class ServerSummary(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("launches", ctypes.c_uint64),
        ("p50_ns", ctypes.c_uint64),
        ("p99_ns", ctypes.c_uint64),
        ("max_ns", ctypes.c_uint64),
        ("mean_batch", ctypes.c_double),
        ("first_submit_ns", ctypes.c_uint64),
        ("last_done_ns", ctypes.c_uint64),
        ("state_bytes_peak", ctypes.c_uint64),
        ("evictions", ctypes.c_uint64),
    ]


class _HostCsr(ctypes.Structure):
    _fields_ = [
        ("V", ctypes.c_uint32),
        ("E", ctypes.c_uint64),
        ("offsets", ctypes.POINTER(ctypes.c_uint32)),
        ("cols", ctypes.POINTER(ctypes.c_uint32)),
        ("weights", ctypes.POINTER(ctypes.c_uint32)),
    ]


_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_lib = None

# Every exported function of include/ptg.h: name -> (restype, argtypes).
SIGNATURES = {
    "ptg_init": (ctypes.c_int, [ctypes.c_int]),
    "ptg_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "ptg_last_error": (ctypes.c_char_p, []),
    "ptg_config_default": (None, [ctypes.POINTER(Config)]),
    "ptg_device_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.POINTER(ctypes.c_uint64)]),
    "ptg_device_alloc": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "ptg_device_free": (ctypes.c_int, [_vp]),
    "ptg_memcpy_d2d": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64]),
    "ptg_graph_stream_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                             ctypes.POINTER(ctypes.c_uint64)]),
    "ptg_host_link_probe": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "ptg_memcpy_d2h": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64]),
    "ptg_memcpy_h2d": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64]),
    "ptg_validate_csr": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, _u32p, _u32p,
                                        ctypes.POINTER(EdgeWithWeight)]),
    "ptg_graph_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, _u32p, _u32p,
                                        ctypes.POINTER(EdgeWithWeight), ctypes.POINTER(Config),
                                        ctypes.POINTER(_vp)]),
    "ptg_graph_create_soa": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, _u32p, _u32p, _u32p,
                                            ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
    "ptg_graph_destroy": (None, [_vp]),
    "ptg_graph_estimate_bytes": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                                   ctypes.POINTER(Config)]),
    "ptg_graph_device_bytes": (ctypes.c_uint64, [_vp]),
    "ptg_graph_is_symmetric": (ctypes.c_int, [_vp]),
    "ptg_graph_set_flags": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "ptg_graph_frontier_sizes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32,
                                                ctypes.POINTER(ctypes.c_uint32)]),
    "ptg_graph_set_config": (ctypes.c_int, [_vp, ctypes.POINTER(Config)]),
    "ptg_graph_get_config": (ctypes.c_int, [_vp, ctypes.POINTER(Config)]),
    "ptg_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]),
    "ptg_server_add_graph": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_uint32)]),
    "ptg_server_submit": (ctypes.c_int, [_vp, ctypes.POINTER(Request)]),
    "ptg_server_replay": (ctypes.c_int, [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.c_uint64]),
    "ptg_server_poll": (ctypes.c_int, [_vp, ctypes.POINTER(Completion), ctypes.c_int, ctypes.c_int]),
    "ptg_server_drain": (ctypes.c_int, [_vp]),
    "ptg_server_pause": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ptg_server_error": (ctypes.c_char_p, [_vp]),
    "ptg_server_summary_get": (ctypes.c_int, [_vp, ctypes.POINTER(ServerSummary)]),
    "ptg_server_destroy": (None, [_vp]),
    "ptg_autotune": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int, ctypes.c_uint32,
                                    ctypes.POINTER(Config)]),
    "ptg_bfs": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp, ctypes.POINTER(Stats)]),
    "ptg_sssp": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp, ctypes.POINTER(Stats)]),
    "ptg_cc": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Stats)]),
    "ptg_pagerank": (ctypes.c_int, [_vp, ctypes.c_float, ctypes.c_int, _vp, ctypes.POINTER(Stats)]),
    "ptg_pagerank_noalloc": (ctypes.c_int, [_vp, ctypes.c_float, ctypes.c_int, _vp, ctypes.POINTER(Stats)]),
    "ptg_bfs_batch": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int, _vp, ctypes.POINTER(Stats)]),
    "ptg_sssp_batch": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int, _vp, ctypes.POINTER(Stats)]),
    "bfsCaculateInShare": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint32]),
    "bfsCaculateInOpt": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint32]),
    "bfsCaculateInAsync": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint32]),
    "bfsCaculateInAsyncSwap": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint32]),
    "ssspCaculateInShare": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                            ctypes.POINTER(EdgeWithWeight), ctypes.c_uint32]),
    "ssspCaculateInOpt": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                          ctypes.POINTER(EdgeWithWeight), ctypes.c_uint32]),
    "ccCaculateInShare": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]),
    "ccCaculateInOpt": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]),
    "prCaculateInShare": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]),
    "prCaculateInOpt": (ctypes.c_long, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]),
    "ptg_gen_rmat": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_HostCsr)]),
    "ptg_gen_er": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                  ctypes.POINTER(_HostCsr)]),
    "ptg_gen_weights": (ctypes.c_int, [ctypes.POINTER(_HostCsr), ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint64, ctypes.c_int]),
    "ptg_host_csr_free": (None, [ctypes.POINTER(_HostCsr)]),
    "ptg_read_bcsr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_HostCsr)]),
    "ptg_write_bcsr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_HostCsr)]),
    "ptg_read_bwcsr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_HostCsr)]),
    "ptg_write_bwcsr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_HostCsr)]),
    "ptg_read_java_be": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_HostCsr)]),
}


def lib():
    """Loads libptgraph.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libptgraph.so is not built: run __graft_entry__.build() "
                               "or `make -C test-world_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, where):
    if rc != OK:
        raise PtgError(rc, where)


def _u32(a):
    return a.ctypes.data_as(_u32p)


def init(device=0):
    _check(lib().ptg_init(device), "ptg_init")


def default_config(**overrides):
    c = Config()
    lib().ptg_config_default(ctypes.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def host_link_probe(nbytes=256 << 20, reps=8):
    """Pinned host -> device copy rate in GB/s."""
    x = ctypes.c_double(0)
    _check(lib().ptg_host_link_probe(nbytes, reps, ctypes.byref(x)), "ptg_host_link_probe")
    return x.value


def device_info():
    n = ctypes.c_int(0)
    tot, free = ctypes.c_uint64(0), ctypes.c_uint64(0)
    _check(lib().ptg_device_info(ctypes.byref(n), ctypes.byref(tot), ctypes.byref(free)), "ptg_device_info")
    return {"num_cus": n.value, "total_mem": tot.value, "free_mem": free.value}


class HostCsr:
    """A CSR owned by the C library (V+1 offsets), exposed as numpy views."""

    def __init__(self):
        self._c = _HostCsr()
        self._owned = False

    def _adopt(self):
        self._owned = True
        V, E = self._c.V, self._c.E
        self.offsets = np.ctypeslib.as_array(self._c.offsets, shape=(V + 1,))
        self.cols = np.ctypeslib.as_array(self._c.cols, shape=(max(E, 1),))[:E]
        self.weights = (np.ctypeslib.as_array(self._c.weights, shape=(max(E, 1),))[:E]
                        if self._c.weights else None)
        return self

    @property
    def V(self):
        return self._c.V

    @property
    def E(self):
        return self._c.E

    @property
    def node_pointers(self):
        """The reference's V-entry pointer array (ptGraph/main.cu:257-264)."""
        return self.offsets[: self.V]

    def add_weights(self, lo=1, hi=64, seed=2, threads=0):
        _check(lib().ptg_gen_weights(ctypes.byref(self._c), lo, hi, seed, threads), "ptg_gen_weights")
        return self._adopt()

    def free(self):
        if self._owned:
            self.offsets = self.cols = self.weights = None
            lib().ptg_host_csr_free(ctypes.byref(self._c))
            self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, threads=0):
    h = HostCsr()
    _check(lib().ptg_gen_rmat(scale, edge_factor, a, b, c, seed, threads, ctypes.byref(h._c)), "ptg_gen_rmat")
    return h._adopt()


def gen_er(scale, edge_factor=16, seed=1, threads=0):
    h = HostCsr()
    _check(lib().ptg_gen_er(scale, edge_factor, seed, threads, ctypes.byref(h._c)), "ptg_gen_er")
    return h._adopt()


def read_graph(path, fmt):
    """fmt: 'bcsr', 'bwcsr' or 'java' (ptGraph/main.cu:98-116, :452-471, :2479-2573)."""
    fn = {"bcsr": "ptg_read_bcsr", "bwcsr": "ptg_read_bwcsr", "java": "ptg_read_java_be"}[fmt]
    h = HostCsr()
    _check(getattr(lib(), fn)(os.fsencode(path), ctypes.byref(h._c)), fn)
    return h._adopt()


def write_graph(path, csr, fmt):
    fn = {"bcsr": "ptg_write_bcsr", "bwcsr": "ptg_write_bwcsr"}[fmt]
    c = _HostCsr()
    offsets = np.ascontiguousarray(csr.offsets, dtype=np.uint32)
    cols = np.ascontiguousarray(csr.cols, dtype=np.uint32)
    w = None if csr.weights is None else np.ascontiguousarray(csr.weights, dtype=np.uint32)
    c.V = offsets.size - 1
    c.E = cols.size
    c.offsets = _u32(offsets)
    c.cols = _u32(cols) if cols.size else None
    c.weights = _u32(w) if w is not None else None
    _check(getattr(lib(), fn)(os.fsencode(path), ctypes.byref(c)), fn)


def estimate_bytes(V, E, weighted, config=None):
    cfg = config if config is not None else default_config()
    return lib().ptg_graph_estimate_bytes(V, E, 1 if weighted else 0, ctypes.byref(cfg))


def validate_csr(node_pointers, cols, E=None):
    ptr = np.ascontiguousarray(node_pointers, dtype=np.uint32)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    E = cols.size if E is None else E
    return lib().ptg_validate_csr(ptr.size, E, _u32(ptr), _u32(cols) if cols.size else None, None)


class DeviceBuffer:
    """HBM buffer from ptg_device_alloc, for outputs that stay on the device."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        p = _vp()
        _check(lib().ptg_device_alloc(nbytes, ctypes.byref(p)), "ptg_device_alloc")
        self.ptr = p.value

    def to_host(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        _check(lib().ptg_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "ptg_memcpy_d2h")
        return out

    def from_host(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _check(lib().ptg_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes), "ptg_memcpy_h2d")

    def copy_from(self, other, nbytes=None):
        n = min(self.nbytes, other.nbytes) if nbytes is None else nbytes
        _check(lib().ptg_memcpy_d2d(self.ptr, other.ptr, n), "ptg_memcpy_d2d")

    def view(self, offset, nbytes):
        """A window [offset, offset + nbytes) of this buffer that does not own
        its memory (an output at any 4-byte offset, for example)."""
        if offset < 0 or nbytes < 0 or offset + nbytes > self.nbytes:
            raise ValueError("view outside the buffer")
        v = DeviceBuffer.__new__(DeviceBuffer)
        v.nbytes, v.ptr, v._owner = nbytes, self.ptr + offset, self
        return v

    def free(self):
        if self.ptr and getattr(self, "_owner", None) is None:
            lib().ptg_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Graph:
    """A graph resident in HBM (ptg_graph_create).

    node_pointers: V entries in the reference layout, or V+1 offsets when
    `offsets_v1=True`; cols: uint32[E]; weights: uint32[E] or None.
    """

    def __init__(self, node_pointers, cols, weights=None, config=None, offsets_v1=False):
        ptr = np.ascontiguousarray(node_pointers, dtype=np.uint32)
        if offsets_v1:
            ptr = np.ascontiguousarray(ptr[:-1])
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        self.V = ptr.size
        self.E = cols.size
        self.weighted = weights is not None
        cfg = config if config is not None else default_config()
        self.sentinel = cfg.sssp_sentinel
        h = _vp()
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint32)
        if w is not None and w.size == 0:
            w = np.zeros(1, np.uint32)  # weighted graph without edges
        rc = lib().ptg_graph_create_soa(self.V, self.E, _u32(ptr), _u32(cols) if cols.size else None,
                                        _u32(w) if w is not None else None,
                                        ctypes.byref(cfg), ctypes.byref(h))
        _check(rc, "ptg_graph_create_soa")
        self._h = h.value

    @classmethod
    def from_csr(cls, csr, weighted=True, config=None):
        return cls(csr.node_pointers, csr.cols, csr.weights if weighted else None, config)

    @classmethod
    def from_aos(cls, node_pointers, edges, config=None):
        """Weighted graph from an EdgeWithWeight array (ptg_graph_create)."""
        ptr = np.ascontiguousarray(node_pointers, dtype=np.uint32)
        edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
        self = cls.__new__(cls)
        self.V, self.E, self.weighted = ptr.size, edges.shape[0], True
        cfg = config if config is not None else default_config()
        self.sentinel = cfg.sssp_sentinel
        h = _vp()
        rc = lib().ptg_graph_create(self.V, self.E, _u32(ptr), None,
                                    edges.ctypes.data_as(ctypes.POINTER(EdgeWithWeight)),
                                    ctypes.byref(cfg), ctypes.byref(h))
        _check(rc, "ptg_graph_create")
        self._h = h.value
        return self

    def _out(self, out, dtype):
        if out is None:
            return np.empty(self.V, dtype=dtype), None
        if isinstance(out, DeviceBuffer):
            return None, out.ptr
        return out, None

    def _run(self, fn, args, out, dtype, name):
        st = Stats()
        host, dev = self._out(out, dtype)
        ptr = dev if dev is not None else host.ctypes.data
        _check(fn(self._h, *args, ptr, ctypes.byref(st)), name)
        return (host if dev is None else out), st

    def bfs(self, src, out=None):
        return self._run(lib().ptg_bfs, (src,), out, np.uint32, "ptg_bfs")

    def sssp(self, src, out=None):
        return self._run(lib().ptg_sssp, (src,), out, np.uint32, "ptg_sssp")

    def cc(self, out=None):
        return self._run(lib().ptg_cc, (), out, np.uint32, "ptg_cc")

    def pagerank(self, damping=0.85, iters=20, out=None):
        return self._run(lib().ptg_pagerank, (ctypes.c_float(damping), iters), out, np.float32, "ptg_pagerank")

    def pagerank_noalloc(self, out, damping=0.85, iters=20):
        """ptg_pagerank_noalloc: no device allocation; `out` is a DeviceBuffer."""
        return self._run(lib().ptg_pagerank_noalloc, (ctypes.c_float(damping), iters), out, np.float32,
                         "ptg_pagerank_noalloc")

    def _run_batch(self, fn, srcs, out, name):
        srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
        n = int(srcs.size)
        st = Stats()
        if out is None:
            host = np.empty((n, self.V), dtype=np.uint32)
            _check(fn(self._h, _u32(srcs), n, host.ctypes.data, ctypes.byref(st)), name)
            return host, st
        if out.nbytes < 4 * n * self.V:
            raise ValueError("output buffer smaller than n x V words")
        _check(fn(self._h, _u32(srcs), n, out.ptr, ctypes.byref(st)), name)
        return out, st

    def bfs_batch(self, srcs, out=None):
        """Levels from up to 64 sources at once: row k == bfs(srcs[k])."""
        return self._run_batch(lib().ptg_bfs_batch, srcs, out, "ptg_bfs_batch")

    def sssp_batch(self, srcs, out=None):
        """Distances from up to 64 sources at once: row k == sssp(srcs[k])."""
        return self._run_batch(lib().ptg_sssp_batch, srcs, out, "ptg_sssp_batch")

    def is_symmetric(self):
        rc = lib().ptg_graph_is_symmetric(self._h)
        if rc < 0:
            raise PtgError(rc, "ptg_graph_is_symmetric")
        return bool(rc)

    def frontier_sizes(self):
        """Frontier size of every round of the last BFS / SSSP / CC query."""
        n = ctypes.c_uint32(0)
        _check(lib().ptg_graph_frontier_sizes(self._h, None, 0, ctypes.byref(n)), "ptg_graph_frontier_sizes")
        buf = (ctypes.c_uint64 * max(1, n.value))()
        _check(lib().ptg_graph_frontier_sizes(self._h, buf, n.value, ctypes.byref(n)), "ptg_graph_frontier_sizes")
        return np.array(buf[: n.value], dtype=np.uint64)

    def set_flags(self, flags):
        """Changes runtime flags (only FLAG_TIMING may change after creation)."""
        _check(lib().ptg_graph_set_flags(self._h, flags), "ptg_graph_set_flags")

    def stream_info(self):
        """(chunks, device slots, bytes per slot) of a PTG_FLAG_STREAM graph."""
        c, n, b = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint64(0)
        _check(lib().ptg_graph_stream_info(self._h, ctypes.byref(c), ctypes.byref(n), ctypes.byref(b)),
               "ptg_graph_stream_info")
        return c.value, n.value, b.value

    def get_config(self):
        c = Config()
        _check(lib().ptg_graph_get_config(self._h, ctypes.byref(c)), "ptg_graph_get_config")
        return c

    def tune(self, **changes):
        """Changes tuning fields (td_to_bu, bu_to_td, sssp_*, batch_*)."""
        c = self.get_config()
        for k, v in changes.items():
            setattr(c, k, v)
        _check(lib().ptg_graph_set_config(self._h, ctypes.byref(c)), "ptg_graph_set_config")

    def autotune(self, srcs, what=0xF):
        """Tunes the graph on sample sources; returns the Config it applied."""
        srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
        c = Config()
        _check(lib().ptg_autotune(self._h, _u32(srcs), int(srcs.size), what, ctypes.byref(c)), "ptg_autotune")
        return c

    @property
    def device_bytes(self):
        return lib().ptg_graph_device_bytes(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().ptg_graph_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop over resident graphs (ptg_server_*): continuous batching,
    memory admission, latency summary."""

    def __init__(self, state_budget_bytes=0, max_batch=64, max_inflight=0, coalesce=True, queue_cap=0, min_batch=0):
        cfg = ServerConfig(state_budget_bytes, max_batch, max_inflight, 1 if coalesce else 0, queue_cap, min_batch)
        h = _vp()
        _check(lib().ptg_server_create(ctypes.byref(cfg), ctypes.byref(h)), "ptg_server_create")
        self._h = h.value
        self.graphs = []

    def add_graph(self, graph):
        gid = ctypes.c_uint32(0)
        _check(lib().ptg_server_add_graph(self._h, graph._h, ctypes.byref(gid)), "ptg_server_add_graph")
        self.graphs.append(graph)
        return gid.value

    @staticmethod
    def request(graph, kind, source=0, iterations=10, damping=0.85, out=None, tag=0):
        r = Request()
        r.graph, r.kind, r.source, r.iterations, r.damping, r.tag = graph, kind, source, iterations, damping, tag
        r.out = out
        return r

    def submit(self, req):
        _check(lib().ptg_server_submit(self._h, ctypes.byref(req)), "ptg_server_submit")

    def replay(self, reqs, arrival_ns):
        arr = (Request * len(reqs))(*reqs)
        t = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        _check(lib().ptg_server_replay(self._h, arr, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(reqs)),
               "ptg_server_replay")

    def drain(self):
        _check(lib().ptg_server_drain(self._h), "ptg_server_drain")

    def error(self):
        return lib().ptg_server_error(self._h).decode()

    def pause(self, paused=True):
        _check(lib().ptg_server_pause(self._h, 1 if paused else 0), "ptg_server_pause")

    def poll(self, max_n=4096, timeout_ms=0):
        buf = (Completion * max_n)()
        k = lib().ptg_server_poll(self._h, buf, max_n, timeout_ms)
        if k < 0:
            raise PtgError(k, "ptg_server_poll")
        return [buf[i] for i in range(k)]

    def completions(self):
        out = []
        while True:
            c = self.poll()
            if not c:
                return out
            out.extend(c)

    def summary(self):
        s = ServerSummary()
        _check(lib().ptg_server_summary_get(self._h, ctypes.byref(s)), "ptg_server_summary_get")
        return s

    def close(self):
        if getattr(self, "_h", None):
            lib().ptg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def checksum(values):
    """sum over v of values[v] * (2v + 1) mod 2^64 (the server's result checksum)."""
    v = np.ascontiguousarray(values).view(np.uint32).astype(np.uint64)
    k = 2 * np.arange(v.size, dtype=np.uint64) + np.uint64(1)
    return int(np.sum(v * k, dtype=np.uint64))
