"""ctypes binding of libsepgraph.so (the C ABI declared in include/sepgraph.h).

This module is plumbing for tests, the smoke check and bench.py: every call
goes straight to the native library, which runs on the GPU. There is no CPU
fallback: if the library is missing or no GPU is visible, the GPU entry points
raise.

The life cycle mirrors the reference's Engine (include/framework/framework.cuh:56-797):
Graph(...) = LoadGraph, Engine.set_options = SetOptions, Engine.bfs(src) =
InitGraph(src) + Start(), Engine.gather() = GatherValue().
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsepgraph.so")

SG_OK = 0
ERRORS = {
    -1: "SG_ERR_INVALID",
    -2: "SG_ERR_RANGE",
    -3: "SG_ERR_NOMEM",
    -4: "SG_ERR_HIP",
    -5: "SG_ERR_CAPACITY",
    -6: "SG_ERR_TIMEOUT",
    -7: "SG_ERR_STATE",
    -8: "SG_ERR_UNSUPPORTED",
    -9: "SG_ERR_IO",
}
CSR_SYMMETRIZE, CSR_DROP_SELF_LOOPS, CSR_DEDUP = 1, 2, 4
PR_AUTO, PR_PULL_TD, PR_PUSH_DD, PR_ITERATIVE = 0, 1, 2, 3
GRAPH_UNDIRECTED, GRAPH_RELABEL, GRAPH_NO_RELABEL = 1, 2, 4
FORMATS = {"gr": 0, "metis": 1, "market": 2}
CLEAN_SELF_LOOPS, CLEAN_DEDUP, CLEAN_TO_UNDIRECTED = 1, 2, 4

VARIANTS = {
    "auto": -1,
    "async_push_td": 0,
    "async_push_dd": 1,
    "async_pull_td": 2,
    "async_pull_dd": 3,
    "sync_push_td": 4,
    "sync_push_dd": 5,
    "sync_pull_td": 6,
    "sync_pull_dd": 7,
}


class SepGraphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class EngineOptions(C.Structure):
    """sg_engine_options (EngineOptions, include/framework/common.h:203-321)."""

    _fields_ = [
        ("force_variant", C.c_int32),
        ("fused", C.c_int32),
        ("priority", C.c_int32),
        ("priority_threshold", C.c_float),
        ("cut_threshold", C.c_float),
        ("lb_push", C.c_int32),
        ("lb_pull", C.c_int32),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("edge_factor", C.c_double),
        ("max_iteration", C.c_int32),
        ("block_size", C.c_int32),
        ("large_degree", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("stage_hubs", C.c_int32),
        ("reserved", C.c_int32 * 6),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("rounds", C.c_uint32),
        ("push_rounds", C.c_uint32),
        ("pull_rounds", C.c_uint32),
        ("reached", C.c_uint32),
        ("reached_edges", C.c_uint64),
        ("policy_trace", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("residual", C.c_double),
        ("iterations", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("status", C.c_int32),
        ("reserved", C.c_uint32 * 7),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class RoundTrace(C.Structure):
    _fields_ = [
        ("round", C.c_uint32),
        ("variant", C.c_int32),
        ("in_nodes", C.c_uint32),
        ("pad", C.c_uint32),
        ("in_edges", C.c_uint64),
        ("time_us", C.c_double),
    ]


_P = C.c_void_p
_U32P = C.POINTER(C.c_uint32)
_lib = None


def lib() -> C.CDLL:
    """Load libsepgraph.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        sig = {
            "sg_last_error": ([], C.c_char_p),
            "sg_version": ([], C.c_int),
            "sg_device_count": ([], C.c_int),
            "sg_gen_rmat_edges": ([C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, _P, _P], C.c_int),
            "sg_gen_grid_edges": ([C.c_uint32, C.c_uint32, C.c_double, C.c_uint64, _P, _P, C.POINTER(C.c_uint64)], C.c_int),
            "sg_build_csr": ([C.c_uint32, C.c_uint64, _P, _P, C.c_int, _P, _P, C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
            "sg_gen_weights": ([C.c_uint32, _P, _P, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, _P], C.c_int),
            "sg_csc_from_csr": ([C.c_uint32, _P, _P, _P, _P, _P, _P], C.c_int),
            "sg_pick_sources": ([C.c_uint32, _P, C.c_int, C.c_uint64, _P], C.c_int),
            "sg_gen_weights_std": ([C.c_uint64, C.c_int32, _P], C.c_int),
            "sg_file_read": ([C.c_char_p, C.c_int, C.POINTER(_P)], C.c_int),
            "sg_file_from_csr": ([C.c_uint32, _P, _P, _P, C.POINTER(_P)], C.c_int),
            "sg_file_free": ([_P], C.c_int),
            "sg_file_info": ([_P, _U32P, C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_uint64)], C.c_int),
            "sg_file_copy": ([_P, _P, _P, _P], C.c_int),
            "sg_file_clean": ([_P, C.c_int], C.c_int),
            "sg_file_write": ([_P, C.c_char_p, C.c_int], C.c_int),
            "sg_file_write_edgelist": ([_P, C.c_char_p, C.c_char_p], C.c_int),
            "sg_file_write_out_degree": ([_P, C.c_char_p, C.c_char_p], C.c_int),
            "sg_graph_create": ([C.c_uint32, C.c_uint64, _P, _P, _P, _P, _P, _P, C.c_int, C.POINTER(_P)], C.c_int),
            "sg_graph_create_ex": ([C.c_uint32, C.c_uint64, _P, _P, _P, _P, _P, _P, C.c_uint32, C.POINTER(_P)], C.c_int),
            "sg_graph_destroy": ([_P], C.c_int),
            "sg_graph_info": ([_P, _U32P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
            "sg_engine_options_default": ([C.POINTER(EngineOptions)], None),
            "sg_engine_create": ([_P, C.POINTER(_P)], C.c_int),
            "sg_engine_destroy": ([_P], C.c_int),
            "sg_engine_set_stream": ([_P, _P], C.c_int),
            "sg_engine_set_options": ([_P, C.POINTER(EngineOptions)], C.c_int),
            "sg_bfs": ([_P, C.c_uint32], C.c_int),
            "sg_sssp": ([_P, C.c_uint32], C.c_int),
            "sg_pagerank": ([_P, C.c_double, C.c_double, C.c_int32, C.c_int32], C.c_int),
            "sg_bc": ([_P, _P, C.c_int32, C.c_int32], C.c_int),
            "sg_bc_dependency": ([_P, _P, C.c_int32, C.c_int32], C.c_int),
            "sg_bfs_multi": ([_P, _P, C.c_int32, C.c_int32], C.c_int),
            "sg_sssp_multi": ([_P, _P, C.c_int32, C.c_int32], C.c_int),
            "sg_sssp_concurrent": ([_P, _P, C.c_int32], C.c_int),
            "sg_engine_sync": ([_P], C.c_int),
            "sg_engine_value_ptr": ([_P, C.POINTER(_P), C.POINTER(C.c_uint64)], C.c_int),
            "sg_engine_gather_value": ([_P, _P, C.c_uint64], C.c_int),
            "sg_engine_get_stats": ([_P, C.POINTER(Stats)], C.c_int),
            "sg_engine_get_trace": ([_P, C.POINTER(RoundTrace), C.c_int, C.POINTER(C.c_int)], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Names of every function declared in include/sepgraph.h."""
    import re

    hdr = os.path.join(_HERE, "..", "include", "sepgraph.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(sg_[a-z0-9_]+)\s*\(", text)))


def _check(rc: int) -> None:
    if rc != SG_OK:
        raise SepGraphError(rc, lib().sg_last_error().decode())


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_P)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


# ------------------------------------------------------------------ generators
def gen_rmat_edges(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1):
    count = edge_factor << scale
    src = np.empty(count, np.uint32)
    dst = np.empty(count, np.uint32)
    _check(lib().sg_gen_rmat_edges(scale, edge_factor, a, b, c, seed, _ptr(src), _ptr(dst)))
    return src, dst


def gen_grid_edges(rows: int, cols: int, p_delete: float = 0.1, seed: int = 1):
    cap = 2 * rows * cols
    src = np.empty(cap, np.uint32)
    dst = np.empty(cap, np.uint32)
    cnt = C.c_uint64()
    _check(lib().sg_gen_grid_edges(rows, cols, p_delete, seed, _ptr(src), _ptr(dst), C.byref(cnt)))
    return src[: cnt.value], dst[: cnt.value]


def build_csr(n: int, src, dst, symmetrize=False, drop_self_loops=True, dedup=True):
    src, dst = _u32(src), _u32(dst)
    flags = (CSR_SYMMETRIZE if symmetrize else 0) | (CSR_DROP_SELF_LOOPS if drop_self_loops else 0)
    flags |= CSR_DEDUP if dedup else 0
    cap = len(src) * (2 if symmetrize else 1)
    row = np.empty(n + 1, np.uint32)
    col = np.empty(max(cap, 1), np.uint32)
    m = C.c_uint64()
    _check(lib().sg_build_csr(n, len(src), _ptr(src), _ptr(dst), flags, _ptr(row), _ptr(col), cap, C.byref(m)))
    return row, col[: m.value].copy()


def gen_weights(row, col, lo: int, hi: int, seed: int = 2, symmetric=False):
    row, col = _u32(row), _u32(col)
    w = np.empty(max(len(col), 1), np.uint32)
    _check(lib().sg_gen_weights(len(row) - 1, _ptr(row), _ptr(col), lo, hi, seed, int(symmetric), _ptr(w)))
    return w[: len(col)]


def csc_from_csr(row, col, w=None):
    row, col = _u32(row), _u32(col)
    n = len(row) - 1
    cs = np.empty(n + 1, np.uint32)
    src = np.empty(max(len(col), 1), np.uint32)
    cw = None if w is None else np.empty(max(len(col), 1), np.uint32)
    _check(lib().sg_csc_from_csr(n, _ptr(row), _ptr(col), _ptr(None if w is None else _u32(w)), _ptr(cs), _ptr(src), _ptr(cw)))
    return cs, src[: len(col)], (None if cw is None else cw[: len(col)])


def pick_sources(row, k: int, seed: int = 3):
    row = _u32(row)
    out = np.empty(k, np.uint32)
    _check(lib().sg_pick_sources(len(row) - 1, _ptr(row), k, seed, _ptr(out)))
    return out


@dataclass
class HostGraph:
    """A host CSR (+ optional weights) with the metadata of its generator."""

    n: int
    row: np.ndarray
    col: np.ndarray
    w: np.ndarray | None = None
    undirected: bool = False
    name: str = ""
    # edge count a reference graph-file reader records (None: not from a file)
    header_edges: int | None = None

    @property
    def m(self) -> int:
        return int(self.row[-1])


def rmat_graph(scale: int, edge_factor: int = 16, symmetrize=False, seed=1, weights=None, weight_seed=2) -> HostGraph:
    """RMAT (a,b,c,d)=(0.57,0.19,0.19,0.05), self-loops and duplicates removed."""
    s, d = gen_rmat_edges(scale, edge_factor, seed=seed)
    n = 1 << scale
    row, col = build_csr(n, s, d, symmetrize=symmetrize)
    del s, d
    w = None
    if weights is not None:
        w = gen_weights(row, col, weights[0], weights[1], seed=weight_seed, symmetric=symmetrize)
    kind = "sym" if symmetrize else "dir"
    return HostGraph(n, row, col, w, undirected=symmetrize, name=f"rmat{scale}-ef{edge_factor}-{kind}")


def grid_graph(rows=2048, cols=2048, p_delete=0.1, seed=1, weights=(1, 1000), weight_seed=2) -> HostGraph:
    """4-neighbour grid, each undirected edge deleted with probability p_delete."""
    s, d = gen_grid_edges(rows, cols, p_delete, seed)
    n = rows * cols
    row, col = build_csr(n, s, d, symmetrize=True)
    w = None if weights is None else gen_weights(row, col, weights[0], weights[1], seed=weight_seed, symmetric=True)
    return HostGraph(n, row, col, w, undirected=True, name=f"grid{rows}x{cols}-p{p_delete}")


# ----------------------------------------------------------------- graph files
def gen_weights_std(m: int, weight_range: int = 100) -> np.ndarray:
    """The reference drivers' --gen_weights sequence (traversal.h:174-186)."""
    w = np.empty(max(m, 1), np.uint32)
    _check(lib().sg_gen_weights_std(m, weight_range, _ptr(w)))
    return w[:m]


def _fmt(fmt: str) -> int:
    if fmt not in FORMATS:
        raise ValueError(f"unknown graph format {fmt!r} (gr, metis, market)")
    return FORMATS[fmt]


def _host_csr_out(h) -> tuple[np.ndarray, np.ndarray, np.ndarray | None, int]:
    L = lib()
    n, m, wt, hdr = C.c_uint32(), C.c_uint64(), C.c_int(), C.c_uint64()
    _check(L.sg_file_info(h, C.byref(n), C.byref(m), C.byref(wt), C.byref(hdr)))
    row = np.empty(n.value + 1, np.uint32)
    col = np.empty(max(m.value, 1), np.uint32)
    w = np.empty(max(m.value, 1), np.uint32) if wt.value else None
    _check(L.sg_file_copy(h, _ptr(row), _ptr(col), _ptr(w)))
    return row, col[: m.value], (None if w is None else w[: m.value]), hdr.value


def _host_csr_in(hg: HostGraph):
    h = _P()
    _check(lib().sg_file_from_csr(hg.n, _ptr(_u32(hg.row)), _ptr(_u32(hg.col)),
                                  _ptr(None if hg.w is None else _u32(hg.w)), C.byref(h)))
    return h


def read_graph(path: str, fmt: str = "gr", clean: int = 0, undirected=False) -> HostGraph:
    """Load a graph file as the reference's --graphfile/--format do
    (src/utils/utils.cpp:39-58), then apply the gtools cleanups in `clean`
    (CLEAN_* flags). `header_edges` records the edge count the reference's
    reader keeps."""
    L = lib()
    h = _P()
    _check(L.sg_file_read(os.fsencode(path), _fmt(fmt), C.byref(h)))
    try:
        if clean:
            _check(L.sg_file_clean(h, clean))
        row, col, w, hdr = _host_csr_out(h)
    finally:
        L.sg_file_free(h)
    return HostGraph(len(row) - 1, row, col, w, undirected=undirected, name=os.path.basename(path), header_edges=hdr)


def clean_graph(hg: HostGraph, clean: int) -> HostGraph:
    """gtools cleanups (samples/tools/gtools.cpp:56-327) on a host CSR."""
    L = lib()
    h = _host_csr_in(hg)
    try:
        _check(L.sg_file_clean(h, clean))
        row, col, w, _ = _host_csr_out(h)
    finally:
        L.sg_file_free(h)
    return HostGraph(len(row) - 1, row, col, w, undirected=hg.undirected, name=hg.name)


def write_graph(hg: HostGraph, path: str, fmt: str = "gr") -> None:
    L = lib()
    h = _host_csr_in(hg)
    try:
        _check(L.sg_file_write(h, os.fsencode(path), _fmt(fmt)))
    finally:
        L.sg_file_free(h)


def write_edgelist(hg: HostGraph, path: str, separator: str = "\t", out_degree: str | None = None) -> None:
    """gtools --output (edge list) and, if given, --out_degree files."""
    L = lib()
    h = _host_csr_in(hg)
    try:
        _check(L.sg_file_write_edgelist(h, os.fsencode(path), separator.encode()))
        if out_degree:
            _check(L.sg_file_write_out_degree(h, os.fsencode(out_degree), separator.encode()))
    finally:
        L.sg_file_free(h)


# --------------------------------------------------------------- device side
class Graph:
    """A device-resident graph (sg_graph)."""

    def __init__(self, hg: HostGraph, with_weights: bool = True, relabel: bool | None = None):
        """relabel: None lets the library decide (degree order on the device for
        skewed graphs), True / False force it either way."""
        self.host = hg
        h = _P()
        w = _u32(hg.w) if (with_weights and hg.w is not None) else None
        row, col = _u32(hg.row), _u32(hg.col)
        flags = (GRAPH_UNDIRECTED if hg.undirected else 0) | {None: 0, True: GRAPH_RELABEL, False: GRAPH_NO_RELABEL}[relabel]
        _check(lib().sg_graph_create_ex(hg.n, hg.m, _ptr(row), _ptr(col), _ptr(w), None, None, None, flags, C.byref(h)))
        self.h = h

    def info(self):
        n, m, b = C.c_uint32(), C.c_uint64(), C.c_uint64()
        _check(lib().sg_graph_info(self.h, C.byref(n), C.byref(m), C.byref(b)))
        return n.value, m.value, b.value

    def close(self):
        if self.h:
            lib().sg_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_options(**kw) -> EngineOptions:
    o = EngineOptions()
    lib().sg_engine_options_default(C.byref(o))
    for k, v in kw.items():
        if k == "force_variant" and isinstance(v, str):
            v = VARIANTS[v]
        setattr(o, k, v)
    return o


class Engine:
    """sg_engine bound to one Graph."""

    def __init__(self, graph: Graph, stream: int | None = None, **opts):
        self.graph = graph
        h = _P()
        _check(lib().sg_engine_create(graph.h, C.byref(h)))
        self.h = h
        if stream is not None:
            _check(lib().sg_engine_set_stream(self.h, _P(stream)))
        if opts:
            self.set_options(**opts)

    def set_options(self, **kw):
        o = default_options(**kw)
        _check(lib().sg_engine_set_options(self.h, C.byref(o)))

    def bfs(self, source: int):
        _check(lib().sg_bfs(self.h, int(source)))

    def sssp(self, source: int):
        _check(lib().sg_sssp(self.h, int(source)))

    def pagerank(self, damping: float = 0.85, tol: float = 1e-6, max_iter: int = 0, mode: int = 0):
        """Normalised PageRank into the value array (FP32); mode PR_PULL_TD,
        PR_PUSH_DD or PR_ITERATIVE (the reference's ITERATIVE_SCHEME)."""
        _check(lib().sg_pagerank(self.h, float(damping), float(tol), int(max_iter), int(mode)))

    def bc(self, sources, batch: int = 64):
        """0.5 * sum of Brandes dependencies over `sources` into the value array (FP32)."""
        src = _u32(sources)
        _check(lib().sg_bc(self.h, _ptr(src), len(src), int(batch)))

    def bc_dependency(self, sources, batch: int = 64):
        """Per-source 0.5 * delta_s rows (k x n FP32) into the value array."""
        src = _u32(sources)
        _check(lib().sg_bc_dependency(self.h, _ptr(src), len(src), int(batch)))

    def bfs_multi(self, sources, batch: int = 64):
        """BFS levels from every source (k x n int32) into the value array."""
        src = _u32(sources)
        _check(lib().sg_bfs_multi(self.h, _ptr(src), len(src), int(batch)))

    def sssp_multi(self, sources, batch: int = 64):
        """SSSP distances from every source (k x n uint32) into the value array."""
        src = _u32(sources)
        _check(lib().sg_sssp_multi(self.h, _ptr(src), len(src), int(batch)))

    def sssp_concurrent(self, sources):
        """SSSP from every source, one concurrent query per XCD (k x n uint32)."""
        src = _u32(sources)
        _check(lib().sg_sssp_concurrent(self.h, _ptr(src), len(src)))

    def sync(self):
        _check(lib().sg_engine_sync(self.h))

    def value_ptr(self):
        p, b = _P(), C.c_uint64()
        _check(lib().sg_engine_value_ptr(self.h, C.byref(p), C.byref(b)))
        return p.value, b.value

    def gather(self, dtype=np.int32) -> np.ndarray:
        _, nbytes = self.value_ptr()
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        _check(lib().sg_engine_gather_value(self.h, _ptr(out), out.nbytes))
        return out

    def stats(self) -> dict:
        s = Stats()
        _check(lib().sg_engine_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def trace(self, max_rows: int = 4096) -> list[dict]:
        """Per-round rows of the last traversal: round, variant, in_nodes, in_edges, time_us."""
        rows = (RoundTrace * max_rows)()
        nrow = C.c_int()
        _check(lib().sg_engine_get_trace(self.h, rows, max_rows, C.byref(nrow)))
        # variant -1 marks the final row (not SG_VARIANT_AUTO)
        inv = {v: k for k, v in VARIANTS.items() if v >= 0}
        return [
            {"round": r.round, "variant": inv.get(r.variant, "end"), "in_nodes": r.in_nodes,
             "in_edges": r.in_edges, "time_us": r.time_us}
            for r in rows[: nrow.value]
        ]

    def close(self):
        if self.h:
            lib().sg_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -------------------------------------------------------------------- serving
QUERY_BFS, QUERY_SSSP, QUERY_BC_DEP = 0, 1, 2
_QUERY_DTYPE = {QUERY_BFS: np.int32, QUERY_SSSP: np.uint32, QUERY_BC_DEP: np.float32}


SERVE_SLOTS, SERVE_BATCHES = 0, 1


class ServerConfig(C.Structure):
    _fields_ = [
        ("memory_budget", C.c_uint64),
        ("max_rows", C.c_int32),
        ("max_batch", C.c_int32 * 3),
        ("bfs_packed_min", C.c_int32),
        ("sssp_packed_min", C.c_int32),
        ("mode", C.c_int32),
        ("sssp_prio_graphs", C.c_uint32),
        ("reserved", C.c_int32 * 4),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("batches", C.c_uint64),
        ("queued", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("mean_batch", C.c_double),
        ("busy_ms", C.c_double),
        ("rows", C.c_uint32),
        ("free_rows", C.c_uint32),
        ("reserved", C.c_uint32 * 8),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_server_ready = False


def _server_lib():
    global _server_ready
    L = lib()
    if not _server_ready:
        D = C.POINTER(C.c_double)
        sig = {
            "sg_server_config_default": ([C.POINTER(ServerConfig)], None),
            "sg_server_create": ([_P, C.POINTER(ServerConfig), C.POINTER(EngineOptions), C.POINTER(_P)], C.c_int),
            "sg_server_destroy": ([_P], C.c_int),
            "sg_server_submit": ([_P, C.c_int32, C.c_uint32, C.POINTER(C.c_uint64)], C.c_int),
            "sg_server_wait": ([_P, C.c_uint64, C.c_int32], C.c_int),
            "sg_server_result": ([_P, C.c_uint64, C.POINTER(_P), C.POINTER(C.c_uint64), D, D], C.c_int),
            "sg_server_copy_result": ([_P, C.c_uint64, _P, C.c_uint64], C.c_int),
            "sg_server_release": ([_P, C.c_uint64], C.c_int),
            "sg_server_get_stats": ([_P, C.POINTER(ServerStats)], C.c_int),
            "sg_server_replay": ([_P, _P, _P, _P, C.c_int64, _P], C.c_int),
            "sg_server_create_multi": ([_P, C.c_int32, C.POINTER(ServerConfig), C.POINTER(EngineOptions),
                                        C.POINTER(_P)], C.c_int),
            "sg_server_submit_to": ([_P, C.c_int32, C.c_int32, C.c_uint32, C.POINTER(C.c_uint64)], C.c_int),
            "sg_server_replay_multi": ([_P, _P, _P, _P, _P, C.c_int64, _P], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _server_ready = True
    return L


class Server:
    """sg_server: batched BFS / SSSP / BC-dependency queries on a resident graph."""

    def __init__(self, graph, memory_budget: int = 4 << 30, max_rows: int = 4096, max_batch=(64, 64, 64),
                 bfs_packed_min: int = 0, sssp_packed_min: int = 0, mode: int = 0, sssp_prio_graphs: int = 0,
                 **opts):
        """graph: a Graph, or a list of up to four Graphs (queries then name
        theirs by index). mode: SERVE_SLOTS (0, the serving kernel with one
        query slot per XCD) or SERVE_BATCHES (1, bit-packed batches formed
        when the GPU frees up; max_batch and the packed_min thresholds apply
        to it; one graph only). sssp_prio_graphs: bit i turns on the two-level
        SSSP priority schedule for graph i (slots only)."""
        L = _server_lib()
        graphs = list(graph) if isinstance(graph, (list, tuple)) else [graph]
        self.graphs = graphs
        self.graph = graphs[0]
        self.n = graphs[0].host.n
        cfg = ServerConfig()
        L.sg_server_config_default(C.byref(cfg))
        cfg.memory_budget = memory_budget
        cfg.max_rows = max_rows
        for i, b in enumerate(max_batch):
            cfg.max_batch[i] = b
        cfg.bfs_packed_min = bfs_packed_min
        cfg.sssp_packed_min = sssp_packed_min
        cfg.mode = mode
        cfg.sssp_prio_graphs = sssp_prio_graphs
        o = default_options(**opts)
        h = _P()
        arr = (_P * len(graphs))(*[gr.h for gr in graphs])
        _check(L.sg_server_create_multi(arr, len(graphs), C.byref(cfg), C.byref(o), C.byref(h)))
        self.h = h
        self._kind_graph = {}

    def submit(self, kind: int, source: int, graph: int = 0) -> int:
        t = C.c_uint64()
        _check(_server_lib().sg_server_submit_to(self.h, int(graph), int(kind), int(source), C.byref(t)))
        self._kind_graph[t.value] = int(graph)
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        _check(_server_lib().sg_server_wait(self.h, ticket, timeout_ms))

    def result(self, ticket: int, kind: int) -> np.ndarray:
        n = self.graphs[self._kind_graph.get(ticket, 0)].host.n
        out = np.empty(n, _QUERY_DTYPE[kind])
        _check(_server_lib().sg_server_copy_result(self.h, ticket, _ptr(out), out.nbytes))
        return out

    def times(self, ticket: int):
        q, t = C.c_double(), C.c_double()
        _check(_server_lib().sg_server_result(self.h, ticket, None, None, C.byref(q), C.byref(t)))
        return q.value, t.value

    def release(self, ticket: int):
        _check(_server_lib().sg_server_release(self.h, ticket))
        self._kind_graph.pop(ticket, None)

    def stats(self) -> dict:
        s = ServerStats()
        _check(_server_lib().sg_server_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def replay(self, kinds, sources, arrival_s, graphs=None) -> np.ndarray:
        """Open-loop arrivals; graphs[i] (optional) is query i's graph index."""
        k = np.ascontiguousarray(kinds, np.int32)
        s = _u32(sources)
        a = np.ascontiguousarray(arrival_s, np.float64)
        gi = None if graphs is None else np.ascontiguousarray(graphs, np.int32)
        lat = np.empty(len(k), np.float64)
        _check(_server_lib().sg_server_replay_multi(self.h, _ptr(gi), _ptr(k), _ptr(s), _ptr(a), len(k), _ptr(lat)))
        return lat

    def close(self):
        if self.h:
            _server_lib().sg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ autotuner
class TuneConfig(C.Structure):
    _fields_ = [
        ("latency_budget_ms", C.c_double),
        ("tune_bfs", C.c_int32),
        ("tune_sssp", C.c_int32),
        ("tune_batches", C.c_int32),
        ("reserved", C.c_int32 * 8),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("bfs_alpha", C.c_double),
        ("bfs_beta", C.c_double),
        ("bfs_block_size", C.c_int32),
        ("sssp_block_size", C.c_int32),
        ("sssp_alpha", C.c_double),
        ("bfs_grid_blocks", C.c_int32),
        ("sssp_grid_blocks", C.c_int32),
        ("sssp_priority", C.c_int32),
        ("pad0", C.c_int32),
        ("bfs_ms", C.c_double),
        ("bfs_default_ms", C.c_double),
        ("sssp_ms", C.c_double),
        ("sssp_default_ms", C.c_double),
        ("max_batch", C.c_int32 * 3),
        ("pad", C.c_int32),
        ("batch_ms", C.c_double * 3),
        ("batch_qps", C.c_double * 3),
        ("reserved", C.c_int32 * 8),
    ]

    def as_dict(self) -> dict:
        d = {}
        for k, _ in self._fields_:
            if k in ("reserved", "pad"):
                continue
            v = getattr(self, k)
            d[k] = list(v) if hasattr(v, "__len__") else v
        return d


def autotune(engine: Engine, sources, latency_budget_ms: float = 25.0, bfs=True, sssp=True, batches=True) -> dict:
    """sg_autotune: per-graph alpha/beta/block size and batch sizes by measurement."""
    L = lib()
    L.sg_tune_config_default.argtypes = [C.POINTER(TuneConfig)]
    L.sg_tune_config_default.restype = None
    L.sg_autotune.argtypes = [_P, _P, C.c_int32, C.POINTER(TuneConfig), C.POINTER(TuneResult)]
    L.sg_autotune.restype = C.c_int
    cfg = TuneConfig()
    L.sg_tune_config_default(C.byref(cfg))
    cfg.latency_budget_ms = latency_budget_ms
    cfg.tune_bfs, cfg.tune_sssp, cfg.tune_batches = int(bfs), int(sssp), int(batches)
    src = _u32(sources)
    out = TuneResult()
    _check(L.sg_autotune(engine.h, _ptr(src), len(src), C.byref(cfg), C.byref(out)))
    return out.as_dict()
