"""The benchmark configurations at their full sizes.

Config 2(b): the 2048 x 2048 four-neighbour grid with 10% of its edges
deleted and weights in [1, 1000) -- the high-diameter case (thousands of
rounds). 64 seeded non-isolated sources: 8 bit-exact against the oracle
(FIFO BFS restating BFSHost, Dijkstra for SSSP), the rest through the
shortest-path characterisation; the two-level priority schedule must give
the same distances; a source in the corner runs more rounds than the trace
holds.
Config 3: directed RMAT-23 PageRank against the FP64 oracle at 1e-6 relative
L1, both schedules.
Config 4: undirected RMAT-20 BC from 256 sources, one source per launch
against 64 per launch."""
import numpy as np
import pytest

from test_bc_gpu import close
from test_bfs_gpu import check_bfs_properties
from test_pagerank_gpu import D, check_td, power_step, rel_l1
from test_policy_gpu import level_profile
from test_sssp_gpu import check_sssp_properties

pytestmark = pytest.mark.gpu

TRACE_ROWS = 4096


@pytest.fixture(scope="module")
def grid2048(sg):
    return sg.grid_graph(2048, 2048, 0.1, seed=1, weights=(1, 1000))


def test_grid2048_bfs_64_sources(sg, orc, gpu, grid2048):
    hg = grid2048
    srcs = [int(s) for s in sg.pick_sources(hg.row, 64, seed=3)]
    e = sg.Engine(sg.Graph(hg))
    deep = 0
    for i, s in enumerate(srcs):
        e.bfs(s)
        lv = e.gather(np.int32)
        st = e.stats()
        assert st["status"] == 0 and st["rounds"] == lv.max() + 1, (s, st)
        if i < 8:
            assert np.array_equal(lv, orc.bfs(hg.row, hg.col, s)), s
        else:
            check_bfs_properties(hg, lv, s)
        deep = max(deep, st["rounds"])
    assert deep > 1000


def test_grid2048_bfs_from_the_corner(sg, orc, gpu, grid2048):
    """From the corner the traversal runs 4095 rounds (depth 4094): levels
    exact, and every trace row (the trace holds 4096) matches the level
    sizes and level out-degree sums."""
    hg = grid2048
    e = sg.Engine(sg.Graph(hg))
    e.bfs(0)
    lv = e.gather(np.int32)
    assert np.array_equal(lv, orc.bfs(hg.row, hg.col, 0))
    st = e.stats()
    assert st["rounds"] == lv.max() + 1 > 4000, st
    t = e.trace()
    assert len(t) == st["rounds"] + 1 and t[-1]["variant"] == "end"
    sizes, sums = level_profile(hg, lv)
    assert [r["in_nodes"] for r in t[:-1]] == sizes.tolist()
    assert [r["in_edges"] for r in t[:-1]] == sums.tolist()


def test_bfs_longer_than_the_trace(sg, orc, gpu):
    """A 5000-vertex path runs more rounds than the trace keeps: levels stay
    exact and the rows kept match the first levels."""
    n = 5000
    row = np.arange(n + 1, dtype=np.uint32)
    row[n] = n - 1
    col = np.arange(1, n, dtype=np.uint32)
    hg = sg.HostGraph(n, row, col)
    e = sg.Engine(sg.Graph(hg))
    e.bfs(0)
    lv = e.gather(np.int32)
    assert np.array_equal(lv, np.arange(n, dtype=np.int32))
    st = e.stats()
    assert st["rounds"] == n and st["status"] == 0
    t = e.trace()
    assert len(t) == TRACE_ROWS
    assert all(r["in_nodes"] == 1 for r in t)


def test_grid2048_sssp_64_sources_with_and_without_priority(sg, orc, gpu, grid2048):
    hg = grid2048
    srcs = [int(s) for s in sg.pick_sources(hg.row, 64, seed=3)]
    g = sg.Graph(hg)
    plain, prio = sg.Engine(g), sg.Engine(g, priority=1)
    for i, s in enumerate(srcs):
        plain.sssp(s)
        d = plain.gather(np.uint32)
        assert plain.stats()["status"] == 0
        if i < 8:
            assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, hg.w, s)), s
        else:
            check_sssp_properties(hg, d, s)
        prio.sssp(s)
        st = prio.stats()
        assert st["status"] == 0 and st["pull_rounds"] == 0
        assert np.array_equal(prio.gather(np.uint32), d), s


def test_pr_rmat23_against_the_oracle(sg, orc, gpu):
    """Config 3 at full size against the FP64 oracle: pull TD stops on the
    oracle's rule and matches it at the same iteration count within 1e-6
    relative L1; data-driven propagation to 1e-8 matches the converged
    oracle within 1e-6."""
    hg = sg.rmat_graph(23, 16, symmetrize=False, seed=1)
    g = sg.Graph(hg)
    e = sg.Engine(g)
    e.pagerank(D, 1e-6, 0, sg.PR_PULL_TD)
    x = e.gather(np.float32)
    check_td(orc, hg, x, e.stats())
    e.pagerank(D, 1e-8, 0, sg.PR_PUSH_DD)
    xd = e.gather(np.float32)
    ref, it, res = orc.pagerank(hg.row, hg.col, D, 1e-12, 100000)
    assert res < 1e-12
    assert rel_l1(xd, ref) <= 1e-6, rel_l1(xd, ref)
    assert np.abs(power_step(hg.row, hg.col, xd) - xd.astype(np.float64)).sum() < 1e-6


def test_bc_rmat20_one_source_per_launch_against_64(sg, orc, gpu):
    """Config 4: the same 256 sources one per launch and 64 per launch give
    the same BC within 1e-5 relative; per-source dependencies of the first
    sources agree with the FP64 oracle either way."""
    hg = sg.rmat_graph(20, 16, symmetrize=True, seed=1)
    srcs = sg.pick_sources(hg.row, 256, seed=3)
    e = sg.Engine(sg.Graph(hg))
    e.bc(srcs, 64)
    x64 = e.gather(np.float32)
    e.bc(srcs, 1)
    x1 = e.gather(np.float32)
    ok, err = close(x1, x64.astype(np.float64))
    assert ok, err
    for s in srcs[:3]:
        want, _ = orc.bc_single(hg.row, hg.col, int(s))
        for batch in (1, 64):
            e.bc_dependency([int(s)], batch)
            ok, err = close(e.gather(np.float32), want)
            assert ok, (int(s), batch, err)
