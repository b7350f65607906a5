"""Batched (bit-packed multi-source) queries on the GPU through the C ABI:
BFS levels from up to 64 sources per launch must equal the single-source
oracle BFS (BFSHost, samples/hybrid_bfs/hybrid_bfs_host.cpp:33-60) bit for
bit, row by row; per-source BC dependencies must match the single-source FP64
Brandes oracle (hybrid_bc_host.cpp:16-92) within 1e-5 relative."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu


def rows(x, k, n):
    return x.reshape(k, n)


SMALL = {
    "path": lambda: G.path(70),
    "grid": lambda: G.grid(17, 23),
    "star": lambda: G.star(80),
    "dir_random": lambda: G.random_graph(700, 3000, 3),
    "sparse_dir": lambda: G.random_graph(2000, 1500, 5),
}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("variant", ["auto", "async_push_dd", "async_pull_dd"])
def test_bfs_multi_small(sg, orc, gpu, name, variant):
    row, col = SMALL[name]()
    n = len(row) - 1
    hg = sg.HostGraph(n, row, col, None, undirected=name != "dir_random" and name != "sparse_dir")
    g = sg.Graph(hg)
    e = sg.Engine(g, force_variant=variant)
    rng = np.random.default_rng(7)
    srcs = rng.integers(0, n, 70).astype(np.uint32)
    srcs[5] = srcs[9]  # a repeated source
    for batch in (64, 13, 1):
        e.bfs_multi(srcs, batch)
        out = rows(e.gather(np.int32), len(srcs), n)
        for i, s in enumerate(srcs):
            assert np.array_equal(out[i], orc.bfs(row, col, int(s))), (name, variant, batch, i)


def test_bfs_multi_rmat16(sg, orc, gpu, rmat16_sym):
    hg = rmat16_sym
    g = sg.Graph(hg)
    e = sg.Engine(g)
    srcs = sg.pick_sources(hg.row, 64, seed=3)
    e.bfs_multi(srcs, 64)
    out = rows(e.gather(np.int32), 64, hg.n)
    st = e.stats()
    for i, s in enumerate(srcs):
        assert np.array_equal(out[i], orc.bfs(hg.row, hg.col, int(s))), i
    assert st["iterations"] == 1 and st["status"] == 0


def test_bfs_multi_matches_single_rmat20_directed(sg, gpu):
    hg = sg.rmat_graph(20, 16, symmetrize=False, seed=1)
    g = sg.Graph(hg)
    e = sg.Engine(g)
    srcs = sg.pick_sources(hg.row, 100, seed=3)
    e.bfs_multi(srcs, 64)
    out = rows(e.gather(np.int32), 100, hg.n).copy()
    for i in (0, 1, 37, 63, 64, 99):
        e.bfs(int(srcs[i]))
        assert np.array_equal(out[i], e.gather(np.int32)), i


@pytest.mark.parametrize("batch", [64, 6, 1])
def test_bc_dependency(sg, orc, gpu, batch):
    row, col = G.random_graph(800, 4000, 11, undirected=True)
    n = 800
    hg = sg.HostGraph(n, row, col, None, undirected=True)
    g = sg.Graph(hg)
    e = sg.Engine(g)
    srcs = np.array([0, 5, 5, 17, 799, 400, 3, 2, 1], np.uint32)
    e.bc_dependency(srcs, batch)
    out = rows(e.gather(np.float32), len(srcs), n).astype(np.float64)
    for i, s in enumerate(srcs):
        hd, _ = orc.bc_single(row, col, int(s))
        err = np.abs(out[i] - hd)
        assert np.all(err <= 1e-5 * np.abs(hd) + 1e-6), (i, err.max())
    # the sum over sources equals sg_bc
    e.bc(srcs, batch)
    tot = e.gather(np.float32).astype(np.float64)
    assert np.allclose(tot, out.sum(0), rtol=1e-5, atol=1e-5)


def _weighted_und(row, col, lo, hi, seed):
    """Symmetric random weights for an undirected CSR."""
    n = len(row) - 1
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(n), np.diff(row.astype(np.int64)))
    a, b = np.minimum(src, col), np.maximum(src, col)
    key = a.astype(np.int64) * n + b
    uk, inv = np.unique(key, return_inverse=True)
    return rng.integers(lo, hi, len(uk)).astype(np.uint32)[inv]


@pytest.mark.parametrize("name", sorted(SMALL))
def test_sssp_multi_small(sg, orc, gpu, name):
    row, col = SMALL[name]()
    n = len(row) - 1
    und = name not in ("dir_random", "sparse_dir")
    w = _weighted_und(row, col, 1, 40, 5) if und else np.random.default_rng(5).integers(1, 40, len(col)).astype(np.uint32)
    hg = sg.HostGraph(n, row, col, w, undirected=und)
    g = sg.Graph(hg)
    e = sg.Engine(g)
    srcs = np.random.default_rng(9).integers(0, n, 67).astype(np.uint32)
    srcs[3] = srcs[40]
    for batch in (64, 9):
        e.sssp_multi(srcs, batch)
        out = rows(e.gather(np.uint32), len(srcs), n)
        for i, s in enumerate(srcs):
            assert np.array_equal(out[i], orc.sssp_dijkstra(row, col, w, int(s))), (name, batch, i)


def test_sssp_multi_matches_single_rmat18(sg, gpu):
    hg = sg.rmat_graph(18, 16, symmetrize=False, seed=1, weights=(1, 64))
    g = sg.Graph(hg)
    e = sg.Engine(g)
    srcs = sg.pick_sources(hg.row, 64, seed=3)
    e.sssp_multi(srcs, 64)
    out = rows(e.gather(np.uint32), 64, hg.n).copy()
    for i in (0, 5, 31, 63):
        e.sssp(int(srcs[i]))
        assert np.array_equal(out[i], e.gather(np.uint32)), i


def test_sssp_multi_grid(sg, orc, gpu):
    hg = sg.grid_graph(128, 128, 0.1, seed=1, weights=(1, 1000))
    g = sg.Graph(hg)
    e = sg.Engine(g)
    srcs = sg.pick_sources(hg.row, 16, seed=3)
    e.sssp_multi(srcs, 16)
    out = rows(e.gather(np.uint32), 16, hg.n)
    for i, s in enumerate(srcs):
        assert np.array_equal(out[i], orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s))), i


@pytest.mark.parametrize("prio", [0, 1])
def test_sssp_concurrent_matches_single(sg, orc, gpu, prio):
    """Up to eight SSSP queries at once, one per XCD, each equal to a lone
    sg_sssp run; 13 sources exercise a full and a partial launch."""
    hg = sg.rmat_graph(16, 16, symmetrize=False, seed=1, weights=(1, 64))
    g = sg.Graph(hg)
    e = sg.Engine(g, priority=prio)
    srcs = sg.pick_sources(hg.row, 13, seed=3)
    srcs[4] = srcs[9]
    e.sssp_concurrent(srcs)
    out = rows(e.gather(np.uint32), 13, hg.n).copy()
    for i, s in enumerate(srcs):
        assert np.array_equal(out[i], orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s))), i
    gg = sg.grid_graph(64, 64, 0.1, seed=1, weights=(1, 1000))
    e2 = sg.Engine(sg.Graph(gg), priority=prio)
    s2 = sg.pick_sources(gg.row, 8, seed=4)
    e2.sssp_concurrent(s2)
    out2 = rows(e2.gather(np.uint32), 8, gg.n)
    for i, s in enumerate(s2):
        assert np.array_equal(out2[i], orc.sssp_dijkstra(gg.row, gg.col, gg.w, int(s))), i
