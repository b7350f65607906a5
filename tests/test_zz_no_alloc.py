"""sg_bfs_noalloc (the allocation-free BFS) through the C ABI: two calls in a
row on one device -- an RMAT-16 graph, then a smaller one -- each bit-exact
against sg_bfs (the engine-based entry point) on the same graph and source,
and against the oracle's FIFO BFS (samples/hybrid_bfs/hybrid_bfs_host.cpp:33-60).
The output buffers are the caller's, allocated here through the HIP runtime."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("HIP runtime library not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def _noalloc(sg):
    f = sg.lib().sg_bfs_noalloc
    f.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _bfs_noalloc(sg, hip, g, source, stream):
    n = g.host.n
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), n * 4) == 0
    try:
        # poison the caller's buffer: every entry must be written
        assert hip.hipMemset(buf, 0x5A, n * 4) == 0
        rc = _noalloc(sg)(g.h, int(source), buf, stream, None)
        if rc != 0:
            raise sg.SepGraphError(rc, sg.lib().sg_last_error().decode())
        out = np.empty(n, np.int32)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), buf, n * 4, 2) == 0  # device to host
        return out
    finally:
        hip.hipFree(buf)


def _bfs_engine(sg, g, source):
    e = sg.Engine(g)
    try:
        e.bfs(int(source))
        return e.gather(np.int32)
    finally:
        e.close()


def test_noalloc_two_calls_match_sg_bfs(sg, orc, gpu):
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        cases = [
            (sg.rmat_graph(16, 16, symmetrize=False, seed=11), (0, 5, 77)),
            (sg.rmat_graph(12, 8, symmetrize=True, seed=12), (3, 1000)),
        ]
        for hg, sources in cases:
            g = sg.Graph(hg)
            try:
                for s in sources:
                    print(f"[noalloc] n={hg.n} m={hg.m} source={s}", flush=True)
                    got = _bfs_noalloc(sg, hip, g, s, stream)
                    want = _bfs_engine(sg, g, s)
                    ref = orc.bfs(hg.row, hg.col, int(s))
                    np.testing.assert_array_equal(got, want)
                    np.testing.assert_array_equal(got, ref)
            finally:
                g.close()
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_default_stream_and_relabelled_graph(sg, orc, gpu):
    hip = _hip()
    hg = sg.rmat_graph(14, 16, symmetrize=True, seed=13)
    g = sg.Graph(hg, relabel=True)
    try:
        for s in (0, 9):
            got = _bfs_noalloc(sg, hip, g, s, None)
            np.testing.assert_array_equal(got, orc.bfs(hg.row, hg.col, int(s)))
    finally:
        g.close()


def test_noalloc_rejects_bad_arguments(sg, gpu):
    hg = sg.rmat_graph(10, 4, symmetrize=True, seed=14)
    g = sg.Graph(hg)
    f = _noalloc(sg)
    try:
        assert f(g.h, hg.n, C.c_void_p(1), None, None) == -2  # SG_ERR_RANGE
        assert f(g.h, 0, None, None, None) == -1  # SG_ERR_INVALID
        o = sg.default_options(block_size=256)
        assert f(g.h, 0, C.c_void_p(1), None, C.byref(o)) == -8  # SG_ERR_UNSUPPORTED
    finally:
        g.close()
