"""Host-side tests that need no GPU: the C-ABI library loads and exports every
symbol include/ptg.h declares, CSR validation, host-forced failure paths, the
synthetic graph generators, and the reference's on-disk formats."""
import ctypes
import hashlib
import os
import re
import struct

import numpy as np
import pytest

import ptgraph as pg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ptg.h")).read()
    return re.findall(r"^PTG_API\s+[\w\s\*]*?\b(\w+)\s*\(", text, re.M)


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 30
    L = ctypes.CDLL(pg.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    # the Python mirror binds exactly the declared set
    assert set(names) == set(pg.SIGNATURES)


def _gpu_present():
    return pg.lib().ptg_init(0) == pg.OK


def test_no_device_is_an_error_not_a_crash():
    if _gpu_present():
        pytest.skip("a GPU is present")
    assert pg.lib().ptg_init(0) == pg.ERR_NO_DEVICE
    g = pg.gen_rmat(6)
    with pytest.raises(pg.PtgError) as e:
        pg.Graph.from_csr(g, weighted=False)
    assert e.value.code == pg.ERR_NO_DEVICE
    # the reference-named drivers return the code instead of exiting
    ptr = np.ascontiguousarray(g.node_pointers)
    col = np.ascontiguousarray(g.cols)
    rc = pg.lib().bfsCaculateInShare(g.V, g.E, pg._u32(ptr), pg._u32(col), 0)
    assert rc == pg.ERR_NO_DEVICE


def test_validation_rejects_malformed_csr():
    ok_ptr = np.array([0, 2, 3, 3], np.uint32)
    ok_col = np.array([1, 2, 0, 1], np.uint32)
    assert pg.validate_csr(ok_ptr, ok_col) == pg.OK
    assert pg.validate_csr(np.array([0, 3, 2, 3], np.uint32), ok_col) == pg.ERR_INVALID_GRAPH
    assert pg.validate_csr(ok_ptr, np.array([1, 2, 0, 4], np.uint32)) == pg.ERR_INVALID_GRAPH
    assert pg.validate_csr(np.array([0, 2, 3, 5], np.uint32), ok_col) == pg.ERR_INVALID_GRAPH
    assert pg.validate_csr(np.array([], np.uint32), ok_col) == pg.ERR_INVALID_ARG
    # E beyond uint32 (the reference's `uint E`)
    L = pg.lib()
    assert L.ptg_validate_csr(4, 1 << 32, pg._u32(ok_ptr), pg._u32(ok_col), None) == pg.ERR_INVALID_GRAPH
    # graph creation validates before touching any device
    with pytest.raises(pg.PtgError) as e:
        pg.Graph(np.array([0, 3, 2, 3], np.uint32), ok_col)
    assert e.value.code == pg.ERR_INVALID_GRAPH


def test_budget_rejected_before_allocation():
    g = pg.gen_rmat(10)
    need = pg.estimate_bytes(g.V, g.E, True)
    assert need > 4 * (g.V + g.E)
    cfg = pg.default_config(device_budget_bytes=need // 2)
    with pytest.raises(pg.PtgError) as e:
        pg.Graph.from_csr(g.add_weights(), config=cfg)
    assert e.value.code == pg.ERR_BUDGET


@pytest.mark.parametrize("field", ["sssp_pull", "batch_pull", "sssp_sweep"])
def test_nan_tuning_rejected_before_allocation(field):
    """A NaN decision factor would make every comparison false and silently
    change the schedule; graph creation refuses it before any device work."""
    g = pg.gen_rmat(8).add_weights()
    cfg = pg.default_config(**{field: float("nan")})
    with pytest.raises(pg.PtgError) as e:
        pg.Graph.from_csr(g, config=cfg)
    assert e.value.code == pg.ERR_INVALID_ARG


def test_config_defaults():
    c = pg.default_config()
    # PageRank layout automatic (binned where it pays), SSSP sweeps automatic
    assert c.pr_layout == pg.PR_AUTO and c.sssp_sweep == 0.0
    assert c.sssp_sentinel == 0xFFFFFFFE
    assert c.small_degree < c.large_degree
    assert c.chunk_bytes == 256 << 20
    assert c.order_hub_degree == 16
    # SSSP: one unbounded bucket with automatic deferral degree and window
    assert c.sssp_delta == 0 and c.sssp_defer_degree == 0 and c.sssp_defer_window == 0


def test_config_layout_matches_header():
    """The ctypes mirror has the field order and size of ptg_config in ptg.h."""
    import ctypes
    import re

    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "ptg.h")).read()
    body = hdr[hdr.index("typedef struct ptg_config {"):hdr.index("} ptg_config;")]
    names = re.findall(r"^\s*(?:uint64_t|uint32_t|float)\s+(\w+);", body, re.M)
    assert names == [f for f, _ in pg.Config._fields_]
    assert ctypes.sizeof(pg.Config) % 8 == 0


def check_clean_symmetric(g):
    off = g.offsets.astype(np.int64)
    assert off[0] == 0 and off[-1] == g.E and np.all(np.diff(off) >= 0)
    src = np.repeat(np.arange(g.V, dtype=np.int64), np.diff(off))
    dst = g.cols.astype(np.int64)
    assert np.all(dst < g.V)
    assert not np.any(src == dst)  # no self-loops
    # strictly ascending within each list (sorted, no duplicates)
    same = src[1:] == src[:-1]
    assert np.all(dst[1:][same] > dst[:-1][same])
    fwd = np.sort(src * g.V + dst)
    rev = np.sort(dst * g.V + src)
    np.testing.assert_array_equal(fwd, rev)


def test_rmat_generator_invariants_and_determinism():
    a = pg.gen_rmat(12, 16, seed=7, threads=3)
    b = pg.gen_rmat(12, 16, seed=7, threads=1)
    assert a.V == 4096
    np.testing.assert_array_equal(a.offsets, b.offsets)
    np.testing.assert_array_equal(a.cols, b.cols)
    check_clean_symmetric(a)
    deg = np.diff(a.offsets.astype(np.int64))
    assert deg.argmax() == 0  # no permutation: vertex 0 is the hub
    c = pg.gen_rmat(12, 16, seed=8)
    assert c.E != a.E or not np.array_equal(c.cols, a.cols)
    # R-MAT scale 16 size matches the expected ~0.87 x 32 x V directed edges
    g16 = pg.gen_rmat(16)
    assert 0.84 < g16.E / (32 * 65536) < 0.90


def test_er_generator_invariants():
    g = pg.gen_er(12, 16, seed=1, threads=2)
    check_clean_symmetric(g)
    deg = np.diff(g.offsets.astype(np.int64))
    assert deg.max() < 80  # no hubs in a uniform graph
    assert 0.95 < g.E / (32 * 4096) <= 1.0


def test_weights_symmetric_in_range_and_deterministic():
    g = pg.gen_rmat(10).add_weights(1, 64, seed=2)
    w1 = g.weights.copy()
    off = g.offsets.astype(np.int64)
    src = np.repeat(np.arange(g.V), np.diff(off))
    assert w1.min() >= 1 and w1.max() <= 64
    lookup = dict(zip(zip(src.tolist(), g.cols.tolist()), w1.tolist()))
    for (u, v), w in list(lookup.items())[:5000]:
        assert lookup[(v, u)] == w
    g2 = pg.gen_rmat(10).add_weights(1, 64, seed=2)
    np.testing.assert_array_equal(g2.weights, w1)


def test_bcsr_bwcsr_roundtrip(tmp_path):
    g = pg.gen_rmat(9).add_weights(0, 4, seed=5)
    p = tmp_path / "g.bcsr"
    pg.write_graph(p, g, "bcsr")
    raw = p.read_bytes()
    # layout of ptGraph/main.cu:98-116: [u32 V][u32 E][u32 ptr x V][u32 col x E]
    assert len(raw) == 8 + 4 * g.V + 4 * g.E
    assert struct.unpack("<II", raw[:8]) == (g.V, g.E)
    h = pg.read_graph(p, "bcsr")
    np.testing.assert_array_equal(h.offsets, g.offsets)
    np.testing.assert_array_equal(h.cols, g.cols)
    q = tmp_path / "g.bwcsr"
    pg.write_graph(q, g, "bwcsr")
    raw = q.read_bytes()
    assert len(raw) == 8 + 4 * g.V + 8 * g.E  # EdgeWithWeight AoS (globals.cuh:2-5)
    e0 = struct.unpack_from("<II", raw, 8 + 4 * g.V)
    assert e0 == (int(g.cols[0]), int(g.weights[0]))
    h = pg.read_graph(q, "bwcsr")
    np.testing.assert_array_equal(h.cols, g.cols)
    np.testing.assert_array_equal(h.weights, g.weights)


def test_java_big_endian_reader(tmp_path):
    # [be u32 V][be u64 E][be u64 ptr x V][be u32 col x E] (ptGraph/main.cu:2479-2573)
    ptr = [0, 2, 3, 3]
    col = [1, 2, 0, 1]
    blob = struct.pack(">IQ", 4, 4) + b"".join(struct.pack(">Q", p) for p in ptr) + \
        b"".join(struct.pack(">I", c) for c in col)
    p = tmp_path / "g.java"
    p.write_bytes(blob)
    h = pg.read_graph(p, "java")
    assert h.V == 4 and h.E == 4
    np.testing.assert_array_equal(h.offsets, [0, 2, 3, 3, 4])
    np.testing.assert_array_equal(h.cols, col)
    p.write_bytes(blob[:20])
    with pytest.raises(pg.PtgError) as e:
        pg.read_graph(p, "java")
    assert e.value.code == pg.ERR_IO


def test_missing_file_is_io_error(tmp_path):
    with pytest.raises(pg.PtgError) as e:
        pg.read_graph(tmp_path / "nope.bcsr", "bcsr")
    assert e.value.code == pg.ERR_IO


def test_generator_matches_golden_fixture():
    with np.load(os.path.join(ROOT, "tests", "golden", "rmat10.npz"), allow_pickle=False) as z:
        g = pg.gen_rmat(10, 16, seed=1).add_weights(1, 64, seed=2)
        np.testing.assert_array_equal(g.node_pointers, z["ptr"])
        np.testing.assert_array_equal(g.cols, z["col"])
        np.testing.assert_array_equal(g.weights, z["w"])
    assert hashlib.sha256(g.cols.tobytes()).hexdigest()


def test_device_buffer_views_do_not_own_memory():
    """A view is a window of its buffer: bounds are checked, and freeing the
    view never frees the buffer's memory (no device needed: a view is only
    pointer arithmetic)."""
    buf = pg.DeviceBuffer.__new__(pg.DeviceBuffer)
    buf.nbytes, buf.ptr = 4096, 0x10000
    v = buf.view(8, 100)
    assert (v.ptr, v.nbytes) == (0x10008, 100)
    for off, n in ((-4, 8), (4090, 8), (0, 4097)):
        with pytest.raises(ValueError):
            buf.view(off, n)
    v.free()  # must not reach the library
    assert v.ptr is None and buf.ptr == 0x10000
    buf.ptr = None  # the fake buffer owns nothing either


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it exits with a message and
    status 2 before touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
