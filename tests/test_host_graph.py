"""Host-side graph construction of the product library (no GPU needed):
generators, CSR build, weights, CSC transpose, source picking, C-ABI surface."""
import os

import numpy as np
import pytest

import graphs as G


def test_library_exports_every_declared_symbol(sg):
    import ctypes

    L = sg.lib()
    names = sg.declared_symbols()
    assert len(names) >= 20
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing
    assert L.sg_version() >= 100


def test_rmat_deterministic_and_thread_independent(sg, monkeypatch):
    a = sg.gen_rmat_edges(12, 16, seed=5)
    monkeypatch.setenv("SG_THREADS", "1")
    b = sg.gen_rmat_edges(12, 16, seed=5)
    monkeypatch.setenv("SG_THREADS", "7")
    c = sg.gen_rmat_edges(12, 16, seed=5)
    for x, y in ((a, b), (a, c)):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
    d = sg.gen_rmat_edges(12, 16, seed=6)
    assert not np.array_equal(a[0], d[0])


def test_rmat_quadrant_statistics(sg):
    scale = 14
    s, d = sg.gen_rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed=1)
    assert len(s) == 16 << scale
    assert s.max() < (1 << scale) and d.max() < (1 << scale)
    top_s = (s >> (scale - 1)) & 1
    top_d = (d >> (scale - 1)) & 1
    N = len(s)
    tol = 4 * np.sqrt(0.25 / N)
    assert abs(top_s.mean() - 0.24) < tol  # c + d
    assert abs(top_d.mean() - 0.24) < tol  # b + d
    assert abs((top_s & top_d).mean() - 0.05) < tol  # d
    # every level draws independently: the lowest bit has the same marginals
    assert abs((s & 1).mean() - 0.24) < tol
    # vertex 0 (all quadrant-a descents) is the hub
    deg = np.bincount(s, minlength=1 << scale)
    assert deg.argmax() == 0


def _np_csr(n, s, d, sym, dedup=True, noloop=True):
    s = s.astype(np.uint64)
    d = d.astype(np.uint64)
    if sym:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
    if noloop:
        keep = s != d
        s, d = s[keep], d[keep]
    key = (s << 32) | d
    key = np.unique(key) if dedup else np.sort(key)
    src = (key >> 32).astype(np.int64)
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row).astype(np.uint32), (key & 0xFFFFFFFF).astype(np.uint32)


@pytest.mark.parametrize("sym", [False, True])
def test_build_csr_matches_numpy(sg, sym):
    s, d = sg.gen_rmat_edges(11, 8, seed=3)
    n = 1 << 11
    row, col = sg.build_csr(n, s, d, symmetrize=sym)
    r2, c2 = _np_csr(n, s, d, sym)
    assert np.array_equal(row, r2) and np.array_equal(col, c2)
    # without dedup the multiplicity survives
    row3, col3 = sg.build_csr(n, s, d, symmetrize=sym, dedup=False)
    r4, c4 = _np_csr(n, s, d, sym, dedup=False)
    assert np.array_equal(row3, r4) and np.array_equal(col3, c4)


def test_build_csr_errors(sg):
    with pytest.raises(sg.SepGraphError) as ei:
        sg.build_csr(4, np.array([0, 5], np.uint32), np.array([1, 2], np.uint32))
    assert ei.value.code == -2
    # capacity too small -> SG_ERR_CAPACITY (call the C entry directly)
    import ctypes as C

    s = np.array([0, 1, 2], np.uint32)
    d = np.array([1, 2, 3], np.uint32)
    row = np.empty(5, np.uint32)
    col = np.empty(2, np.uint32)
    m = C.c_uint64()
    rc = sg.lib().sg_build_csr(4, 3, s.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p), 6,
                               row.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p), 2, C.byref(m))
    assert rc == -5


def test_grid_generator(sg):
    r, c = 64, 48
    s, d = sg.gen_grid_edges(r, c, 0.1, seed=9)
    full = r * (c - 1) + (r - 1) * c
    assert 0.85 * full < len(s) < 0.95 * full
    assert np.all(s < d)
    diff = d.astype(np.int64) - s
    assert set(np.unique(diff)) <= {1, c}
    assert np.all((diff != 1) | (s % c != c - 1))  # no wrap-around
    s0, d0 = sg.gen_grid_edges(r, c, 0.0, seed=9)
    assert len(s0) == full
    hg = sg.grid_graph(r, c, 0.1, seed=9)
    row, col = hg.row, hg.col
    assert np.diff(row.astype(np.int64)).max() <= 4
    # symmetric weights
    w = hg.w
    assert w.min() >= 1 and w.max() < 1000
    cs, src, cw = sg.csc_from_csr(row, col, w)
    assert np.array_equal(cs, row) and np.array_equal(src, col) and np.array_equal(cw, w)


def test_weights_range_and_symmetry(sg):
    hg = sg.rmat_graph(10, 8, symmetrize=True, weights=(1, 64))
    w = hg.w
    assert w.min() >= 1 and w.max() <= 63
    assert len(np.unique(w)) == 63
    cs, src, cw = sg.csc_from_csr(hg.row, hg.col, w)
    assert np.array_equal(cw, w)  # symmetric graph + symmetric weights -> CSC == CSR
    hd = sg.rmat_graph(10, 8, symmetrize=False, weights=(1, 64))
    assert np.array_equal(hd.w, sg.gen_weights(hd.row, hd.col, 1, 64, seed=2))


@pytest.mark.parametrize("seed", range(3))
def test_csc_matches_oracle(sg, orc, seed):
    hg = sg.rmat_graph(12, 8, symmetrize=False, seed=seed, weights=(1, 64))
    a = sg.csc_from_csr(hg.row, hg.col, hg.w)
    b = orc.csc_from_csr(hg.row, hg.col, hg.w)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_pick_sources(sg):
    hg = sg.rmat_graph(12, 4, symmetrize=False)
    deg = np.diff(hg.row.astype(np.int64))
    s = sg.pick_sources(hg.row, 64, seed=3)
    assert len(set(s.tolist())) == 64
    assert np.all(deg[s] > 0)
    assert np.array_equal(s, sg.pick_sources(hg.row, 64, seed=3))
    row, col = G.star(3, undirected=False)  # only vertex 0 has out-edges
    with pytest.raises(sg.SepGraphError):
        sg.pick_sources(row, 2)


def test_graph_create_validates_before_device(sg):
    import ctypes as C

    L = sg.lib()
    h = C.c_void_p()
    row = np.array([0, 1, 1], np.uint32)
    col = np.array([5], np.uint32)  # out of range
    rc = L.sg_graph_create(2, 1, row.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p), None, None, None, None, 0, C.byref(h))
    assert rc == -2
    rc = L.sg_graph_create(0, 0, row.ctypes.data_as(C.c_void_p), None, None, None, None, None, 0, C.byref(h))
    assert rc == -2
    bad = np.array([0, 2, 1], np.uint32)
    rc = L.sg_graph_create(2, 1, bad.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p), None, None, None, None, 0, C.byref(h))
    assert rc == -1


def _create_with_csc(sg, n, row, col, cs, src, undirected=False):
    import ctypes as C

    h = C.c_void_p()
    rc = sg.lib().sg_graph_create_ex(n, len(col), sg._ptr(row), sg._ptr(col), None, sg._ptr(cs), sg._ptr(src),
                                     None, 1 if undirected else 0, C.byref(h))
    if rc == 0:
        sg.lib().sg_graph_destroy(h)
    return rc


def test_malformed_csc_is_rejected_before_use(sg):
    """A caller's CSC is validated like the CSR (no device needed: the checks
    run before the library looks for a GPU)."""
    n = 4
    # edges 0->1, 0->2, 1->2, 2->3
    row = np.array([0, 2, 3, 4, 4], np.uint32)
    col = np.array([1, 2, 2, 3], np.uint32)
    cs = np.array([0, 0, 1, 3, 4], np.uint32)
    src = np.array([0, 0, 1, 2], np.uint32)
    ERR_INVALID, ERR_RANGE = -1, -2
    bad_first = cs.copy()
    bad_first[0] = 1
    assert _create_with_csc(sg, n, row, col, bad_first, src) == ERR_INVALID
    bad_last = cs.copy()
    bad_last[n] = 3
    assert _create_with_csc(sg, n, row, col, bad_last, src) == ERR_INVALID
    not_monotone = np.array([0, 2, 1, 3, 4], np.uint32)
    assert _create_with_csc(sg, n, row, col, not_monotone, src) == ERR_INVALID
    src_range = np.array([0, 0, 1, 9], np.uint32)
    assert _create_with_csc(sg, n, row, col, cs, src_range) == ERR_RANGE
    # right shape, but vertex 1 gets two in-edges and vertex 2 one
    wrong_degrees = np.array([0, 0, 2, 3, 4], np.uint32)
    assert _create_with_csc(sg, n, row, col, wrong_degrees, src) == ERR_INVALID
    # the checks apply to directed graphs only (an undirected CSC aliases the CSR)
    assert sg.lib().sg_last_error() is not None
