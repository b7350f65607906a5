"""BFS on the GPU through the C ABI, bit-exact against the oracle (FIFO BFS,
samples/hybrid_bfs/hybrid_bfs_host.cpp:33-60) and against the BFS level
characterisation at full size."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu


def run_bfs(sg, hg, sources, **opts):
    g = sg.Graph(hg)
    e = sg.Engine(g, **opts)
    outs = []
    for s in sources:
        e.bfs(int(s))
        outs.append((e.gather(np.int32), e.stats()))
    return outs


def check_bfs_properties(hg, lv, src):
    """Levels are BFS levels iff: lv[src]=0; for every edge u->v with u reached,
    v is reached and lv[v] <= lv[u]+1; every reached v != src has an in-neighbour
    at lv[v]-1; unreached vertices have no reached in-neighbour."""
    row = hg.row.astype(np.int64)
    col = hg.col.astype(np.int64)
    deg = np.diff(row)
    srcs = np.repeat(np.arange(hg.n), deg)
    lu = lv[srcs].astype(np.int64)
    lvv = lv[col].astype(np.int64)
    reached_u = lu >= 0
    assert lv[src] == 0
    assert np.all(lvv[reached_u] >= 0)
    assert np.all(lvv[reached_u] <= lu[reached_u] + 1)
    parent_ok = np.zeros(hg.n, bool)
    tight = reached_u & (lvv == lu + 1)
    parent_ok[col[tight]] = True
    reached = lv >= 0
    reached_nonsrc = reached.copy()
    reached_nonsrc[src] = False
    assert np.all(parent_ok[reached_nonsrc])


SMALL = {
    "path": lambda: G.path(33),
    "star": lambda: G.star(70),
    "cycle": lambda: G.cycle(129),
    "grid": lambda: G.grid(17, 23),
    "dir_random": lambda: G.random_graph(500, 3000, 1),
    "und_random": lambda: G.random_graph(700, 2000, 2, undirected=True),
    "dir_path": lambda: G.path(100, undirected=False),
}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("variant", ["auto", "async_push_dd", "async_pull_dd"])
def test_bfs_small_graphs(sg, orc, gpu, name, variant):
    row, col = SMALL[name]()
    und = name in ("path", "star", "cycle", "grid", "und_random")
    hg = sg.HostGraph(len(row) - 1, row, col, undirected=und)
    srcs = [0, hg.n - 1, hg.n // 2]
    for s, (lv, st) in zip(srcs, run_bfs(sg, hg, srcs, force_variant=variant)):
        ref = orc.bfs(row, col, s)
        assert np.array_equal(lv, ref), (name, variant, s)
        assert st["reached"] == int((ref >= 0).sum())
        deg = np.diff(row.astype(np.int64))
        assert st["reached_edges"] == int(deg[ref >= 0].sum())


@pytest.mark.parametrize("relabel", [False, True])
def test_bfs_sssp_multi_edges_unsorted(sg, orc, gpu, relabel):
    """A raw CSR with repeated edges, self-loops and unsorted neighbour lists:
    a wave's lanes then hit one frontier word in several separate runs and
    with the same head twice, which the bit claims must handle (each vertex
    claimed once, none lost)."""
    rng = np.random.default_rng(21)
    n = 3000
    deg = rng.integers(0, 40, size=n)
    deg[:8] = 600  # hubs, so some rows span several chunks
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(deg)
    # heads drawn from a narrow window around the tail: many land in the
    # same 64-vertex word, repeats and self-loops included
    tails = np.repeat(np.arange(n), deg)
    col = ((tails + rng.integers(-70, 70, size=len(tails))) % n).astype(np.uint32)
    w = rng.integers(1, 50, size=len(col)).astype(np.uint32)
    hg = sg.HostGraph(n, row, col, w)
    e = sg.Engine(sg.Graph(hg, relabel=relabel))
    for s in (0, 5, 1234, n - 1):
        e.bfs(s)
        assert np.array_equal(e.gather(np.int32), orc.bfs(row, col, s)), s
        e.sssp(s)
        assert np.array_equal(e.gather(np.uint32), orc.sssp_dijkstra(row, col, w, s)), s


def test_bfs_single_vertex_and_isolated_source(sg, orc, gpu):
    hg = sg.HostGraph(1, np.array([0, 0], np.uint32), np.zeros(0, np.uint32))
    (lv, st), = run_bfs(sg, hg, [0])
    assert lv.tolist() == [0] and st["reached"] == 1
    row, col = G.csr_from_pairs(200, [(1, 2), (2, 3)])
    hg = sg.HostGraph(200, row, col)
    (lv, st), = run_bfs(sg, hg, [150])
    assert lv[150] == 0 and (lv >= 0).sum() == 1


def test_bfs_rejects_bad_source(sg, gpu):
    row, col = G.path(5)
    e = sg.Engine(sg.Graph(sg.HostGraph(5, row, col, undirected=True)))
    with pytest.raises(sg.SepGraphError) as ei:
        e.bfs(5)
    assert ei.value.code == -2


@pytest.mark.parametrize("opts", [
    {},
    {"force_variant": "async_push_dd"},
    {"force_variant": "async_pull_dd"},
    {"block_size": 256},
    {"block_size": 1024},
    {"block_size": 512},
    {"max_iteration": 0, "edge_factor": 4.0},
    {"alpha": 0.0},
    {"alpha": 2.0, "beta": 2.0},
    {"lb_push": 0},
    {"lb_pull": 0},
    {"lb_push": 0, "lb_pull": 0, "force_variant": "async_pull_dd"},
    {"lb_push": 0, "force_variant": "async_push_dd", "block_size": 256},
    {"lb_push": 2, "lb_pull": 3, "fused": 1, "large_degree": 256},
])
def test_bfs_rmat16_config1(sg, orc, gpu, rmat16_sym, opts):
    """Config 1: symmetrised RMAT-16, source 0 plus seeded sources, bit-exact."""
    hg = rmat16_sym
    srcs = [0] + sg.pick_sources(hg.row, 15, seed=3).tolist()
    for s, (lv, st) in zip(srcs, run_bfs(sg, hg, srcs, **opts)):
        assert np.array_equal(lv, orc.bfs(hg.row, hg.col, s)), (opts, s)
        assert st["status"] == 0


def test_bfs_rmat16_64_sources_trace(sg, orc, gpu, rmat16_sym):
    hg = rmat16_sym
    srcs = sg.pick_sources(hg.row, 64, seed=3)
    outs = run_bfs(sg, hg, srcs)
    pulls = 0
    for s, (lv, st) in zip(srcs, outs):
        assert np.array_equal(lv, orc.bfs(hg.row, hg.col, int(s)))
        pulls += st["pull_rounds"]
        assert st["rounds"] == st["push_rounds"] + st["pull_rounds"]
        assert st["max_depth"] == lv.max()
    assert pulls > 0  # the alpha/beta rule switches to pull on RMAT


def test_bfs_directed_rmat18(sg, orc, gpu):
    hg = sg.rmat_graph(18, 16, symmetrize=False, seed=1)
    srcs = sg.pick_sources(hg.row, 6, seed=3)
    for s, (lv, st) in zip(srcs, run_bfs(sg, hg, srcs)):
        assert np.array_equal(lv, orc.bfs(hg.row, hg.col, int(s)))


def test_bfs_grid_high_diameter(sg, orc, gpu):
    hg = sg.grid_graph(256, 256, 0.1, seed=1, weights=None)
    for s, (lv, st) in zip([0, 33000], run_bfs(sg, hg, [0, 33000])):
        assert np.array_equal(lv, orc.bfs(hg.row, hg.col, s))
        assert st["rounds"] == lv.max() + 1 and st["rounds"] > 300


def test_bfs_rmat22_full_size(sg, orc, gpu):
    """Config 2(a) graph at full size: oracle equality for 2 sources, the BFS
    level characterisation for 8 more."""
    hg = sg.rmat_graph(22, 16, symmetrize=False, seed=1)
    srcs = sg.pick_sources(hg.row, 10, seed=3)
    outs = run_bfs(sg, hg, srcs)
    for i, (s, (lv, st)) in enumerate(zip(srcs, outs)):
        if i < 2:
            assert np.array_equal(lv, orc.bfs(hg.row, hg.col, int(s)))
        check_bfs_properties(hg, lv, int(s))
        assert st["reached"] == int((lv >= 0).sum())


def test_bfs_engine_reuse_and_two_engines(sg, orc, gpu, rmat16_sym):
    hg = rmat16_sym
    g = sg.Graph(hg)
    e1, e2 = sg.Engine(g), sg.Engine(g)
    for s in (5, 0, 5):
        e1.bfs(s)
        e2.bfs(s + 1)
        assert np.array_equal(e1.gather(), orc.bfs(hg.row, hg.col, s))
        assert np.array_equal(e2.gather(), orc.bfs(hg.row, hg.col, s + 1))


def test_bfs_state_reuse_across_kinds(sg, orc, gpu):
    """A BFS that follows another BFS reuses the state that query left
    (frontier bitmaps and counters zeroed, the idle visited filter reset) and
    expands a narrow source in one block; one that follows other kernels, or
    starts from a wide source, resets everything. Every order gives the
    oracle's levels."""
    hg = sg.rmat_graph(15, 16, symmetrize=True, seed=5, weights=(1, 64))
    deg = np.diff(hg.row.astype(np.int64))
    wide = int(np.argmax(deg))
    assert deg[wide] > 4096  # above what one block expands alone
    narrow = [int(s) for s in sg.pick_sources(hg.row, 6, seed=8)]
    for relabel in (False, True):
        e = sg.Engine(sg.Graph(hg, relabel=relabel))
        steps = [("bfs", narrow[0]), ("bfs", narrow[1]), ("sssp", narrow[2]), ("bfs", narrow[3]),
                 ("pr", 0), ("bfs", narrow[4]), ("bfs", wide), ("bfs", narrow[5]), ("multi", 0),
                 ("bfs", narrow[0]), ("bfs", narrow[0])]
        for kind, s in steps:
            if kind == "bfs":
                e.bfs(s)
                assert np.array_equal(e.gather(np.int32), orc.bfs(hg.row, hg.col, s)), (relabel, s)
                assert e.stats()["status"] == 0
            elif kind == "sssp":
                e.sssp(s)
            elif kind == "pr":
                e.pagerank(0.85, 1e-4, 3, sg.PR_PULL_TD)
            else:
                e.bfs_multi(narrow[:3], 64)
            e.sync()


def test_bfs_round_limit_after_reuse(sg, orc, gpu):
    """max_iteration = 1 stops after round 0: the source at 0, its
    out-neighbours at 1 (written at the start of the next round when block 0
    expanded the source alone), every other vertex -1."""
    hg = sg.rmat_graph(13, 16, symmetrize=False, seed=2)
    srcs = [int(s) for s in sg.pick_sources(hg.row, 4, seed=1)]
    e = sg.Engine(sg.Graph(hg), max_iteration=1)
    for s in srcs + srcs:
        e.bfs(s)
        ref = orc.bfs(hg.row, hg.col, s)
        want = np.where((ref >= 0) & (ref <= 1), ref, -1)
        assert np.array_equal(e.gather(np.int32), want), s


def test_two_engines_interleaved_without_sync(sg, orc, gpu, rmat16_sym):
    """Two engines (two streams) launching persistent kernels back to back
    without host synchronisation: the launches are ordered device-wide, so
    no run waits at a barrier for blocks the other holds."""
    hg = rmat16_sym
    g = sg.Graph(hg)
    e1, e2 = sg.Engine(g), sg.Engine(g)
    srcs = [int(s) for s in sg.pick_sources(hg.row, 12, seed=7)]
    for i, s in enumerate(srcs):
        e1.bfs(s)
        if i % 3 == 2:
            e2.pagerank(0.85, 1e-4, 5, sg.PR_PULL_TD)
        else:
            e2.bfs(srcs[-1 - i])
    assert e1.stats()["status"] == 0 and e2.stats()["status"] == 0
    assert np.array_equal(e1.gather(np.int32), orc.bfs(hg.row, hg.col, srcs[-1]))
    e2.bfs(srcs[0])
    assert np.array_equal(e2.gather(np.int32), orc.bfs(hg.row, hg.col, srcs[0]))


def test_options_outside_the_implementation_are_rejected(sg, gpu, rmat16_sym):
    """LoadBalancing::NONE exists for BFS and SSSP only; large_degree is the
    split the graph was built with; unknown modes are invalid."""
    e = sg.Engine(sg.Graph(rmat16_sym))
    for kw, code in (({"lb_push": 7}, -1), ({"lb_pull": -2}, -1), ({"large_degree": 100}, -8)):
        with pytest.raises(sg.SepGraphError) as ei:
            e.set_options(**kw)
        assert ei.value.code == code, kw
    e.set_options(lb_pull=0)
    with pytest.raises(sg.SepGraphError) as ei:
        e.pagerank(0.85, 1e-4, 3, sg.PR_PULL_TD)
    assert ei.value.code == -8
    with pytest.raises(sg.SepGraphError) as ei:
        e.bc_dependency([0, 1], 2)
    assert ei.value.code == -8


@pytest.mark.parametrize("block_size", [256, 1024])
def test_bfs_hub_lists_staged_in_lds(sg, orc, gpu, block_size):
    """stage_hubs=1: push rounds stage the neighbour lists of high-degree
    frontier vertices in LDS per workgroup; levels stay bit-exact."""
    for hg in (sg.rmat_graph(18, 16, symmetrize=False, seed=21), sg.rmat_graph(16, 16, symmetrize=True, seed=22)):
        e = sg.Engine(sg.Graph(hg), stage_hubs=1, block_size=block_size)
        for s in sg.pick_sources(hg.row, 6, seed=23):
            e.bfs(int(s))
            assert np.array_equal(e.gather(np.int32), orc.bfs(hg.row, hg.col, int(s))), (hg.name, int(s))
            assert e.stats()["status"] == 0
