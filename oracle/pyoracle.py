"""ctypes binding of the CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's cpu_baseline leg may import this
module, and only as the checker or the timed CPU baseline; the product path
(sep-graph_amd/) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_P = C.c_void_p
_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE], check=True, capture_output=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32, i32, dbl = C.c_uint32, C.c_int, C.c_double
        sig = {
            "or_csc_from_csr": [u32, _P, _P, _P, _P, _P, _P],
            "or_bfs": [u32, _P, _P, u32, _P],
            "or_sssp_fifo": [u32, _P, _P, _P, u32, _P],
            "or_sssp_dijkstra": [u32, _P, _P, _P, u32, _P],
            "or_pagerank": [u32, _P, _P, dbl, dbl, i32, _P, C.POINTER(C.c_int), C.POINTER(C.c_double)],
            "or_pagerank_residual": [u32, _P, _P, dbl, dbl, _P],
            "or_bc_single": [u32, _P, _P, u32, _P, _P],
            "or_bc": [u32, _P, _P, _P, i32, _P],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with {rc}")


def csc_from_csr(row, col, w=None):
    row, col = _u32(row), _u32(col)
    n = len(row) - 1
    cs = np.empty(n + 1, np.uint32)
    src = np.empty(max(len(col), 1), np.uint32)
    cw = None if w is None else np.empty(max(len(col), 1), np.uint32)
    _ok(lib().or_csc_from_csr(n, _p(row), _p(col), _p(None if w is None else _u32(w)), _p(cs), _p(src), _p(cw)), "csc")
    return cs, src[: len(col)], (None if cw is None else cw[: len(col)])


def bfs(row, col, source: int) -> np.ndarray:
    row, col = _u32(row), _u32(col)
    out = np.empty(len(row) - 1, np.int32)
    _ok(lib().or_bfs(len(row) - 1, _p(row), _p(col), source, _p(out)), "bfs")
    return out


def sssp_fifo(row, col, w, source: int) -> np.ndarray:
    row, col, w = _u32(row), _u32(col), _u32(w)
    out = np.empty(len(row) - 1, np.uint32)
    _ok(lib().or_sssp_fifo(len(row) - 1, _p(row), _p(col), _p(w), source, _p(out)), "sssp_fifo")
    return out


def sssp_dijkstra(row, col, w, source: int) -> np.ndarray:
    row, col, w = _u32(row), _u32(col), _u32(w)
    out = np.empty(len(row) - 1, np.uint32)
    _ok(lib().or_sssp_dijkstra(len(row) - 1, _p(row), _p(col), _p(w), source, _p(out)), "sssp_dijkstra")
    return out


def pagerank(row, col, damping=0.85, tol=1e-6, max_iter=1000):
    row, col = _u32(row), _u32(col)
    x = np.empty(len(row) - 1, np.float64)
    it, res = C.c_int(), C.c_double()
    _ok(lib().or_pagerank(len(row) - 1, _p(row), _p(col), damping, tol, max_iter, _p(x), C.byref(it), C.byref(res)), "pagerank")
    return x, it.value, res.value


def pagerank_residual(row, col, alpha=0.85, error=0.01) -> np.ndarray:
    row, col = _u32(row), _u32(col)
    r = np.empty(len(row) - 1, np.float64)
    _ok(lib().or_pagerank_residual(len(row) - 1, _p(row), _p(col), alpha, error, _p(r)), "pagerank_residual")
    return r


def bc_single(row, col, source: int):
    row, col = _u32(row), _u32(col)
    n = len(row) - 1
    hd = np.empty(n, np.float64)
    sg = np.empty(n, np.float64)
    _ok(lib().or_bc_single(n, _p(row), _p(col), source, _p(hd), _p(sg)), "bc_single")
    return hd, sg


def bc(row, col, sources) -> np.ndarray:
    row, col, s = _u32(row), _u32(col), _u32(sources)
    out = np.empty(len(row) - 1, np.float64)
    _ok(lib().or_bc(len(row) - 1, _p(row), _p(col), _p(s), len(s), _p(out)), "bc")
    return out
