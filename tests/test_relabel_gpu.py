"""Degree-ordered device layout: every entry point takes and returns the
caller's vertex ids, so a graph uploaded with the relabelling forced on must
give exactly the results of the same graph kept in the caller's order (BFS
levels and SSSP distances bit-exact, PageRank and BC within their FP
tolerances), and both must match the oracle."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu


def both(sg, hg, **opts):
    return [sg.Engine(sg.Graph(hg, relabel=r), **opts) for r in (True, False)]


def test_relabel_traversals(sg, orc, gpu):
    hg = sg.rmat_graph(13, 16, symmetrize=False, seed=5, weights=(1, 64))
    srcs = sg.pick_sources(hg.row, 6, seed=2)
    er, en = both(sg, hg)
    for s in srcs:
        s = int(s)
        er.bfs(s)
        en.bfs(s)
        lv = er.gather(np.int32)
        assert np.array_equal(lv, en.gather(np.int32))
        assert np.array_equal(lv, orc.bfs(hg.row, hg.col, s))
        assert er.stats()["reached"] == en.stats()["reached"] == int((lv >= 0).sum())
        er.sssp(s)
        en.sssp(s)
        d = er.gather(np.uint32)
        assert np.array_equal(d, en.gather(np.uint32))
        assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, hg.w, s))
    for e in (er, en):
        e.bfs_multi(srcs, 64)
        out = e.gather(np.int32).reshape(len(srcs), hg.n)
        for i, s in enumerate(srcs):
            assert np.array_equal(out[i], orc.bfs(hg.row, hg.col, int(s)))
        e.sssp_multi(srcs, 64)
        out = e.gather(np.uint32).reshape(len(srcs), hg.n)
        for i, s in enumerate(srcs):
            assert np.array_equal(out[i], orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s)))
        e.sssp_concurrent(srcs)
        out = e.gather(np.uint32).reshape(len(srcs), hg.n)
        for i, s in enumerate(srcs):
            assert np.array_equal(out[i], orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s)))


def test_relabel_pagerank_bc(sg, orc, gpu):
    hg = sg.rmat_graph(13, 16, symmetrize=True, seed=6)
    er, en = both(sg, hg)
    ref, _, _ = orc.pagerank(hg.row, hg.col, 0.85, 1e-13, 100000)
    for e in (er, en):
        e.pagerank(0.85, 1e-9, 0, sg.PR_PULL_TD)
        x = e.gather(np.float32).astype(np.float64)
        assert np.abs(x - ref).sum() / ref.sum() <= 1e-6
    srcs = sg.pick_sources(hg.row, 20, seed=4)
    bref = orc.bc(hg.row, hg.col, srcs)
    for e in (er, en):
        e.bc(srcs, 64)
        x = e.gather(np.float32).astype(np.float64)
        assert np.all(np.abs(x - bref) <= 1e-5 * np.abs(bref) + 1e-6)
        e.bc_dependency(srcs[:3], 64)
        rows = e.gather(np.float32).reshape(3, hg.n).astype(np.float64)
        for i, s in enumerate(srcs[:3]):
            hd, _ = orc.bc_single(hg.row, hg.col, int(s))
            assert np.all(np.abs(rows[i] - hd) <= 1e-5 * np.abs(hd) + 1e-6)


def test_relabel_server(sg, orc, gpu):
    hg = sg.rmat_graph(12, 16, symmetrize=True, seed=7, weights=(1, 64))
    srv = sg.Server(sg.Graph(hg, relabel=True), max_rows=64)
    srcs = sg.pick_sources(hg.row, 12, seed=8)
    kinds = [0, 1, 2] * 4
    t = [srv.submit(k, int(s)) for k, s in zip(kinds, srcs)]
    for i, (k, s) in enumerate(zip(kinds, srcs)):
        srv.wait(t[i], 60000)
        got = srv.result(t[i], k)
        if k == 0:
            assert np.array_equal(got, orc.bfs(hg.row, hg.col, int(s)))
        elif k == 1:
            assert np.array_equal(got, orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s)))
        else:
            hd, _ = orc.bc_single(hg.row, hg.col, int(s))
            assert np.all(np.abs(got.astype(np.float64) - hd) <= 1e-5 * np.abs(hd) + 1e-6)
        srv.release(t[i])
    srv.close()


@pytest.mark.parametrize("name", ["star", "path", "two_parts", "dir_random"])
def test_relabel_small(sg, orc, gpu, name):
    """Tiny graphs (below the automatic threshold) with the relabelling forced:
    hubs at the front of the device order, isolated parts, sources whose
    device id differs from the caller's."""
    row, col = {
        "star": lambda: G.star(25),
        "path": lambda: G.path(33),
        "two_parts": lambda: G.csr_from_pairs(60, [(i, i + 1) for i in range(29)] + [(30 + i, 31 + i) for i in range(29)], True),
        "dir_random": lambda: G.random_graph(400, 2400, 3),
    }[name]()
    n = len(row) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(row.astype(np.int64)))
    lo, hi = np.minimum(src, col), np.maximum(src, col)
    w = ((lo * 31 + hi * 17) % 61 + 1).astype(np.uint32)  # symmetric in (u, v)
    hg = sg.HostGraph(n, row, col, w, undirected=name != "dir_random")
    er, en = both(sg, hg)
    for s in sorted({0, n - 1, n // 2}):
        for e in (er, en):
            e.bfs(s)
            assert np.array_equal(e.gather(np.int32), orc.bfs(row, col, s))
            e.sssp(s)
            assert np.array_equal(e.gather(np.uint32), orc.sssp_dijkstra(row, col, w, s))
