"""Serving: continuous batching of BFS / SSSP / BC-dependency queries through
the C ABI, in both forms -- the serving kernel with one query slot per XCD
(SERVE_SLOTS) and bit-packed batches (SERVE_BATCHES). Every result must equal
the same query run alone through the single-query entry points (bit-exact for
BFS levels and SSSP distances, 1e-5 relative for BC dependencies, whose FP64
sums may associate differently in a different batch or slot), whatever
launch it landed in; admission must hold queries back while no result row is
free."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wgraph(sg):
    return sg.rmat_graph(14, 16, symmetrize=True, seed=1, weights=(1, 64))


def alone(sg, e, kind, s):
    if kind == sg.QUERY_BFS:
        e.bfs(int(s))
        return e.gather(np.int32)
    if kind == sg.QUERY_SSSP:
        e.sssp(int(s))
        return e.gather(np.uint32)
    e.bc_dependency([int(s)], 1)
    return e.gather(np.float32)


def same(sg, kind, got, ref):
    if kind == sg.QUERY_BC_DEP:
        return bool(np.all(np.abs(got.astype(np.float64) - ref) <= 1e-5 * np.abs(ref) + 1e-6))
    return np.array_equal(got, ref)


MODES = [0, 1]  # SERVE_SLOTS, SERVE_BATCHES


@pytest.mark.parametrize("mode", MODES)
def test_server_mix_equals_alone(sg, gpu, wgraph, mode):
    print(f"[server] mix, mode {mode}", flush=True)
    g = sg.Graph(wgraph)
    srv = sg.Server(g, max_rows=512, mode=mode)
    rng = np.random.default_rng(4)
    srcs = sg.pick_sources(wgraph.row, 150, seed=5)
    kinds = rng.choice([0, 1, 2], size=150, p=[0.6, 0.3, 0.1])
    tickets = [None] * 150

    def client(lo, hi):
        for i in range(lo, hi):
            tickets[i] = srv.submit(int(kinds[i]), int(srcs[i]))

    th = [threading.Thread(target=client, args=(i * 50, i * 50 + 50)) for i in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    results = []
    for i in range(150):
        srv.wait(tickets[i], 60000)
        results.append(srv.result(tickets[i], int(kinds[i])))
        q, tot = srv.times(tickets[i])
        assert 0 <= q <= tot
        srv.release(tickets[i])
    st = srv.stats()
    assert st["completed"] == 150 and st["failed"] == 0
    assert st["mean_batch"] > 1.0, st  # concurrent arrivals were batched
    e = sg.Engine(g)
    for i in range(150):
        ref = alone(sg, e, int(kinds[i]), srcs[i])
        assert same(sg, int(kinds[i]), results[i], ref), (i, int(kinds[i]), int(srcs[i]))
    srv.close()


@pytest.mark.parametrize("mode", MODES)
def test_server_admission_by_rows(sg, gpu, wgraph, mode):
    print(f"[server] admission, mode {mode}", flush=True)
    g = sg.Graph(wgraph)
    n = wgraph.n
    srv = sg.Server(g, memory_budget=3 * n * 4, max_rows=0, mode=mode)
    assert srv.stats()["rows"] == 3
    srcs = sg.pick_sources(wgraph.row, 8, seed=6)
    t = [srv.submit(sg.QUERY_BFS, int(s)) for s in srcs]
    for i in range(3):
        srv.wait(t[i], 60000)
    # the remaining five wait for a free row
    with pytest.raises(sg.SepGraphError):
        srv.wait(t[3], 300)
    assert srv.stats()["queued"] == 5
    got = []
    for i in range(8):
        srv.wait(t[i], 60000)
        got.append(srv.result(t[i], sg.QUERY_BFS))
        srv.release(t[i])
    with pytest.raises(sg.SepGraphError):
        srv.submit(sg.QUERY_BFS, n)
    with pytest.raises(sg.SepGraphError):
        srv.submit(7, 0)
    srv.close()
    e = sg.Engine(g)  # lone runs once the server is idle
    for i in range(8):
        assert np.array_equal(got[i], alone(sg, e, sg.QUERY_BFS, srcs[i]))


@pytest.mark.parametrize("packed_min", [1, 65])
def test_server_batch_forms(sg, gpu, wgraph, packed_min):
    """packed_min 1: every BFS / SSSP batch runs as one bit-packed launch;
    65: every batch runs as single whole-GPU queries. Results equal the lone
    entry points either way."""
    g = sg.Graph(wgraph)
    srv = sg.Server(g, max_rows=256, bfs_packed_min=packed_min, sssp_packed_min=packed_min, mode=sg.SERVE_BATCHES)
    srcs = sg.pick_sources(wgraph.row, 48, seed=11)
    kinds = [sg.QUERY_BFS, sg.QUERY_SSSP] * 24
    tickets = [srv.submit(k, int(s)) for k, s in zip(kinds, srcs)]
    results = []
    for t, k in zip(tickets, kinds):
        srv.wait(t, 60000)
        results.append(srv.result(t, k))
        srv.release(t)
    assert srv.stats()["failed"] == 0
    srv.close()
    # lone runs only once the server is idle: persistent kernels of two
    # engines on different streams must not overlap on one GPU
    e = sg.Engine(g)
    for got, k, s in zip(results, kinds, srcs):
        assert same(sg, k, got, alone(sg, e, k, s)), (packed_min, k, int(s))
    with pytest.raises(sg.SepGraphError):
        sg.Server(g, bfs_packed_min=66)
    with pytest.raises(sg.SepGraphError):
        sg.Server(g, mode=2)


@pytest.mark.parametrize("mode", MODES)
def test_server_replay(sg, gpu, wgraph, mode):
    print(f"[server] replay, mode {mode}", flush=True)
    g = sg.Graph(wgraph)
    srv = sg.Server(g, max_rows=1024, mode=mode)
    rng = np.random.default_rng(8)
    count = 2000
    kinds = rng.choice([0, 1, 2], size=count, p=[0.6, 0.3, 0.1])
    srcs = sg.pick_sources(wgraph.row, count, seed=9) if count <= wgraph.n else None
    arr = np.cumsum(rng.exponential(1.0 / 20000.0, count))
    lat = srv.replay(kinds, srcs, arr)
    assert np.all(lat >= 0)
    st = srv.stats()
    assert st["completed"] == count and st["free_rows"] == st["rows"]
    assert st["mean_batch"] > 1.0
    srv.close()


@pytest.mark.parametrize("prio", [0, 2])
def test_server_two_graphs(sg, gpu, wgraph, prio):
    """One serving kernel over an RMAT graph and a grid resident together:
    queries on either graph, mixed in one arrival stream, equal lone runs on
    their own graph; with prio = 2 the grid's SSSP queries run the two-level
    priority schedule (sssp_prio_graphs), the RMAT graph's do not."""
    print(f"[server] two graphs, sssp_prio_graphs {prio}", flush=True)
    grid = sg.grid_graph(48, 40, p_delete=0.1, seed=3, weights=(1, 1000))
    ga, gb = sg.Graph(wgraph), sg.Graph(grid)
    srv = sg.Server([ga, gb], max_rows=128, sssp_prio_graphs=prio)
    rng = np.random.default_rng(12)
    sa = sg.pick_sources(wgraph.row, 80, seed=13)
    sb = sg.pick_sources(grid.row, 80, seed=14)
    which = rng.integers(0, 2, size=80)
    kinds = rng.choice([0, 1, 2], size=80, p=[0.6, 0.3, 0.1])
    plan = []
    ia = ib = 0
    for w, k in zip(which, kinds):
        if w == 0:
            plan.append((0, int(k), int(sa[ia])))
            ia += 1
        else:
            plan.append((1, int(k), int(sb[ib])))
            ib += 1
    tickets = [srv.submit(k, s, graph=gi) for gi, k, s in plan]
    results = []
    for t, (gi, k, s) in zip(tickets, plan):
        srv.wait(t, 60000)
        results.append(srv.result(t, k))
        srv.release(t)
    assert srv.stats()["failed"] == 0
    srv.close()
    engines = [sg.Engine(ga), sg.Engine(gb)]
    for got, (gi, k, s) in zip(results, plan):
        assert len(got) == (wgraph.n, grid.n)[gi]
        assert same(sg, k, got, alone(sg, engines[gi], k, s)), (gi, k, s)
    with pytest.raises(sg.SepGraphError):
        sg.Server([ga, gb], mode=sg.SERVE_BATCHES)
    with pytest.raises(sg.SepGraphError):
        sg.Server([ga, gb], sssp_prio_graphs=4)  # no graph 2
    with pytest.raises(sg.SepGraphError):
        sg.Server(ga, mode=sg.SERVE_BATCHES, sssp_prio_graphs=1)
