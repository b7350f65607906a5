"""SSSP on the GPU through the C ABI, bit-exact against the oracle (FIFO
label-correcting SSSPHostNaive, samples/hybrid_sssp/hybrid_sssp_host.cpp:34-60,
and Dijkstra) and against the shortest-path characterisation at full size."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu
INF = np.uint32(0xFFFFFFFF)


def run_sssp(sg, hg, sources, **opts):
    g = sg.Graph(hg)
    e = sg.Engine(g, **opts)
    outs = []
    for s in sources:
        e.sssp(int(s))
        outs.append((e.gather(np.uint32), e.stats()))
    return outs


def check_sssp_properties(hg, d, src):
    """d is the shortest-path distance vector iff d[src]=0, every edge u->v
    with d[u] finite has d[v] <= d[u]+w, and every reached v != src has an
    in-edge with d[v] == d[u]+w (weights are positive)."""
    row = hg.row.astype(np.int64)
    col = hg.col.astype(np.int64)
    w = hg.w.astype(np.int64)
    srcs = np.repeat(np.arange(hg.n), np.diff(row))
    du = d[srcs].astype(np.int64)
    dv = d[col].astype(np.int64)
    fin = du != int(INF)
    assert d[src] == 0
    assert np.all(dv[fin] <= du[fin] + w[fin])
    tight = fin & (dv == du + w)
    ok = np.zeros(hg.n, bool)
    ok[col[tight]] = True
    reached = d != INF
    reached[src] = False
    assert np.all(ok[reached])


def weighted(row, col, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi, len(col)).astype(np.uint32)


SMALL = {
    "path": lambda: G.path(40),
    "grid": lambda: G.grid(19, 21),
    "cycle": lambda: G.cycle(77),
    "star": lambda: G.star(90),
    "dir_random": lambda: G.random_graph(600, 4000, 3),
    "und_random": lambda: G.random_graph(500, 1500, 4, undirected=True),
}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("variant", ["auto", "async_push_dd", "async_pull_dd"])
def test_sssp_small_graphs(sg, orc, gpu, name, variant):
    row, col = SMALL[name]()
    w = weighted(row, col, 1, 50, hash(name) % 1000)
    und = name not in ("dir_random",)
    if und:
        # symmetric weights for undirected graphs: weight of (u,v) = weight of (v,u)
        n = len(row) - 1
        key = {}
        for u in range(n):
            for e in range(row[u], row[u + 1]):
                a, b = min(u, col[e]), max(u, col[e])
                key.setdefault((a, b), w[e])
                w[e] = key[(a, b)]
    hg = sg.HostGraph(len(row) - 1, row, col, w, undirected=und)
    srcs = [0, hg.n - 1, hg.n // 3]
    for s, (d, st) in zip(srcs, run_sssp(sg, hg, srcs, force_variant=variant)):
        ref = orc.sssp_dijkstra(row, col, w, s)
        assert np.array_equal(d, ref), (name, variant, s)
        assert np.array_equal(ref, orc.sssp_fifo(row, col, w, s))
        assert st["reached"] == int((ref != INF).sum())


def test_sssp_unweighted_graph_uses_unit_weights(sg, orc, gpu, rmat16_sym):
    hg = rmat16_sym  # no weights: the all-ones fallback makes SSSP equal BFS
    for s, (d, _) in zip([0, 9], run_sssp(sg, hg, [0, 9])):
        lv = orc.bfs(hg.row, hg.col, s)
        assert np.array_equal(d.astype(np.int64), np.where(lv < 0, int(INF), lv.astype(np.int64)))


def test_sssp_rejects_bad_source(sg, gpu):
    row, col = G.path(5)
    e = sg.Engine(sg.Graph(sg.HostGraph(5, row, col, np.ones(len(col), np.uint32), undirected=True)))
    with pytest.raises(sg.SepGraphError) as ei:
        e.sssp(5)
    assert ei.value.code == -2


@pytest.mark.parametrize("opts", [{}, {"force_variant": "async_push_dd"}, {"force_variant": "async_pull_dd"},
                                  {"block_size": 256}, {"block_size": 512}, {"alpha": 2.0},
                                  {"lb_push": 0}, {"lb_pull": 0, "force_variant": "async_pull_dd"},
                                  {"lb_push": 0, "lb_pull": 0}, {"lb_push": 0, "priority": 1}])
def test_sssp_rmat16_weighted(sg, orc, gpu, opts):
    hg = sg.rmat_graph(16, 16, symmetrize=True, seed=1, weights=(1, 64))
    srcs = sg.pick_sources(hg.row, 8, seed=3)
    for s, (d, st) in zip(srcs, run_sssp(sg, hg, srcs, **opts)):
        assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s))), (opts, s)
        assert st["status"] == 0


def test_sssp_directed_rmat18(sg, orc, gpu):
    hg = sg.rmat_graph(18, 16, symmetrize=False, seed=1, weights=(1, 64))
    srcs = sg.pick_sources(hg.row, 4, seed=3)
    for s, (d, st) in zip(srcs, run_sssp(sg, hg, srcs)):
        assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s)))


def test_sssp_grid_high_diameter(sg, orc, gpu):
    hg = sg.grid_graph(256, 256, 0.1, seed=1, weights=(1, 1000))
    for s, (d, st) in zip([0, 40000], run_sssp(sg, hg, [0, 40000])):
        assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, hg.w, s))


def test_sssp_rmat22_full_size(sg, orc, gpu):
    """Config 2(a) at full size: Dijkstra equality for one source, the
    shortest-path characterisation for three more."""
    hg = sg.rmat_graph(22, 16, symmetrize=False, seed=1, weights=(1, 64))
    srcs = sg.pick_sources(hg.row, 4, seed=3)
    for i, (s, (d, st)) in enumerate(zip(srcs, run_sssp(sg, hg, srcs))):
        if i == 0:
            assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s)))
        check_sssp_properties(hg, d, int(s))


def test_bfs_and_sssp_share_an_engine(sg, orc, gpu):
    hg = sg.rmat_graph(14, 16, symmetrize=True, seed=2, weights=(1, 64))
    e = sg.Engine(sg.Graph(hg))
    for s in (3, 11):
        e.sssp(s)
        assert np.array_equal(e.gather(np.uint32), orc.sssp_dijkstra(hg.row, hg.col, hg.w, s))
        e.bfs(s)
        assert np.array_equal(e.gather(np.int32), orc.bfs(hg.row, hg.col, s))


@pytest.mark.parametrize("delta", [0.0, 1.0, 37.0, 5000.0])
def test_sssp_two_level_priority(sg, orc, gpu, delta):
    """The near/far schedule (SetTwoLevelBasedPriority) changes the order of
    work, never the distances."""
    hg = sg.rmat_graph(14, 16, symmetrize=False, seed=1, weights=(1, 64))
    g = sg.Graph(hg)
    e = sg.Engine(g, priority=1, priority_threshold=delta)
    for s in sg.pick_sources(hg.row, 4, seed=3):
        e.sssp(int(s))
        assert np.array_equal(e.gather(np.uint32), orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s)))
        assert e.stats()["pull_rounds"] == 0
    gg = sg.grid_graph(96, 96, 0.1, seed=1, weights=(1, 1000))
    e2 = sg.Engine(sg.Graph(gg), priority=1, priority_threshold=delta)
    for s in (0, 4000):
        e2.sssp(s)
        assert np.array_equal(e2.gather(np.uint32), orc.sssp_dijkstra(gg.row, gg.col, gg.w, s))
