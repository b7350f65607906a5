"""Betweenness centrality on the GPU through the C ABI against the FP64
Brandes oracle (BetweennessCentralityHost, samples/hybrid_bc/hybrid_bc_host.cpp:16-92):
value[v] = 0.5 * sum over sources s of delta_s(v).

Tolerance: per vertex |gpu - ref| <= 1e-5 * |ref| + 1e-6 (FP32 output of
FP64 accumulations). Batches of 1..64 sources per launch must agree with
each other and with the oracle; at full size (undirected RMAT-20, 256
sources) the check is the identity sum_v delta_s(v) = sum over reached t of
(d(s, t) - 1), with distances from the GPU BFS (itself bit-exact against the
oracle in the BFS tests)."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu


def close(x, ref):
    ref = np.asarray(ref, np.float64)
    err = np.abs(x.astype(np.float64) - ref)
    return bool(np.all(err <= 1e-5 * np.abs(ref) + 1e-6)), float(err.max(initial=0.0))


def run_bc(sg, hg, sources, batch, **opts):
    g = sg.Graph(hg)
    e = sg.Engine(g, **opts)
    e.bc(sources, batch)
    return e.gather(np.float32), e.stats(), e


SMALL = {
    "path": lambda: G.path(33),
    "cycle": lambda: G.cycle(41),
    "star": lambda: G.star(25),
    "grid": lambda: G.grid(9, 13),
    "complete": lambda: G.complete(10),
    "und_random": lambda: G.random_graph(300, 900, 4, undirected=True),
    "dir_random": lambda: G.random_graph(400, 2400, 3),
    "two_parts": lambda: G.csr_from_pairs(60, [(i, i + 1) for i in range(29)] + [(30 + i, 31 + i) for i in range(29)], True),
}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("batch", [1, 5, 64])
def test_bc_small(sg, orc, gpu, name, batch):
    row, col = SMALL[name]()
    n = len(row) - 1
    und = name != "dir_random"
    hg = sg.HostGraph(n, row, col, None, undirected=und)
    srcs = np.arange(0, n, max(1, n // 20), dtype=np.uint32)[:23]
    x, st, _ = run_bc(sg, hg, srcs, batch)
    ok, err = close(x, orc.bc(row, col, srcs))
    assert ok, (name, batch, err)
    assert st["iterations"] == -(-len(srcs) // batch)


def test_bc_closed_forms(sg, gpu):
    # undirected path, all sources: BC(v_i) = i (n - 1 - i)
    n = 50
    row, col = G.path(n)
    hg = sg.HostGraph(n, row, col, None, undirected=True)
    x, _, _ = run_bc(sg, hg, np.arange(n), 64)
    i = np.arange(n)
    np.testing.assert_allclose(x, i * (n - 1 - i), rtol=1e-6)
    # star K_{1,k}: centre k (k - 1) / 2, leaves 0
    k = 30
    row, col = G.star(k)
    hg = sg.HostGraph(k + 1, row, col, None, undirected=True)
    x, _, _ = run_bc(sg, hg, np.arange(k + 1), 16)
    assert abs(x[0] - k * (k - 1) / 2) <= 1e-5 * k * k and np.all(x[1:] == 0)
    # complete graph: 0 everywhere
    row, col = G.complete(12)
    hg = sg.HostGraph(12, row, col, None, undirected=True)
    x, _, _ = run_bc(sg, hg, np.arange(12), 64)
    assert np.all(x == 0)
    # single source on a path from 0: 0.5 * delta(i) = 0.5 (n - 1 - i), source 0
    row, col = G.path(n)
    hg = sg.HostGraph(n, row, col, None, undirected=True)
    x, _, _ = run_bc(sg, hg, [0], 1)
    exp = 0.5 * (n - 1 - i)
    exp[0] = 0
    np.testing.assert_allclose(x, exp, rtol=1e-6)


def test_bc_duplicate_sources_and_errors(sg, orc, gpu):
    row, col = G.random_graph(500, 2000, 8, undirected=True)
    hg = sg.HostGraph(500, row, col, None, undirected=True)
    srcs = np.array([3, 3, 7, 3, 499, 7, 0], np.uint32)
    for batch in (1, 3, 64):
        x, _, e = run_bc(sg, hg, srcs, batch)
        ok, err = close(x, orc.bc(row, col, srcs))
        assert ok, (batch, err)
    with pytest.raises(sg.SepGraphError):
        e.bc([500], 64)
    with pytest.raises(sg.SepGraphError):
        e.bc([1, 2], 65)
    with pytest.raises(sg.SepGraphError):
        e.bc(np.zeros(0, np.uint32), 64)


@pytest.mark.parametrize("variant", ["auto", "async_push_dd", "async_pull_dd"])
def test_bc_rmat16_batches(sg, orc, gpu, rmat16_sym, variant):
    hg = rmat16_sym
    srcs = sg.pick_sources(hg.row, 64, seed=3)
    ref = orc.bc(hg.row, hg.col, srcs)
    for batch in (64, 32, 1):
        x, st, _ = run_bc(sg, hg, srcs, batch, force_variant=variant)
        ok, err = close(x, ref)
        assert ok, (variant, batch, err, st)


@pytest.mark.parametrize("variant", ["auto", "async_push_dd", "async_pull_dd"])
@pytest.mark.parametrize("name", ["long_path", "grid", "dir_random", "two_parts", "rmat14_dir", "rmat14_und"])
def test_bc_single_source_directions(sg, orc, gpu, name, variant):
    """One source per launch (the single-source kernel): its sigma and
    dependency level passes pull or push by the levels' edge volumes, and a
    forced push or pull variant fixes every pass's direction. All three must
    match the oracle, including levels past the 64 whose volumes are counted
    (the 150-vertex path) and hub chunk lists (RMAT-14)."""
    if name == "long_path":
        row, col, und = *G.path(150), True
    elif name == "rmat14_dir" or name == "rmat14_und":
        h = sg.rmat_graph(14, 16, symmetrize=name == "rmat14_und", seed=5)
        row, col, und = h.row, h.col, name == "rmat14_und"
    else:
        row, col = SMALL[name]()
        und = name != "dir_random"
    n = len(row) - 1
    hg = sg.HostGraph(n, row, col, None, undirected=und)
    srcs = np.asarray(sg.pick_sources(hg.row, 12, seed=7), np.uint32)
    x, st, _ = run_bc(sg, hg, srcs, 1, force_variant=variant)
    ok, err = close(x, orc.bc(row, col, srcs))
    assert ok, (name, variant, err)
    assert st["iterations"] == len(srcs)


def test_bc_rmat18_directed(sg, orc, gpu):
    hg = sg.rmat_graph(18, 16, symmetrize=False, seed=1)
    srcs = sg.pick_sources(hg.row, 40, seed=3)
    x, _, _ = run_bc(sg, hg, srcs, 64)
    ok, err = close(x, orc.bc(hg.row, hg.col, srcs))
    assert ok, err


def test_bc_rmat20_undirected_full_size(sg, orc, gpu):
    """Config: undirected, deduplicated RMAT-20, 256 seeded sources, 64 per
    launch. The first 64 sources (a whole launch) are checked against the
    oracle; all 256 through the dependency-sum identity; batch 32 must agree
    with batch 64 (one source per launch against 64: test_configs_gpu)."""
    hg = sg.rmat_graph(20, 16, symmetrize=True, seed=1)
    srcs = sg.pick_sources(hg.row, 256, seed=3)
    g = sg.Graph(hg)
    e = sg.Engine(g)
    e.bc(srcs[:64], 64)
    ok, err = close(e.gather(np.float32), orc.bc(hg.row, hg.col, srcs[:64]))
    assert ok, err
    e.bc(srcs, 64)
    x64 = e.gather(np.float32)
    st = e.stats()
    assert st["iterations"] == 4 and st["status"] == 0
    e.bc(srcs, 32)
    x32 = e.gather(np.float32)
    ok, err = close(x32, x64.astype(np.float64))
    assert ok, err
    # sum_v bc(v) = 0.5 * sum_s sum_{t reached, t != s} (d(s,t) - 1)
    tot = 0.0
    for s in srcs:
        e.bfs(int(s))
        lv = e.gather(np.int32)
        r = lv[lv > 0].astype(np.float64)
        tot += 0.5 * (r - 1).sum()
    assert abs(float(x64.astype(np.float64).sum()) - tot) <= 1e-5 * tot
