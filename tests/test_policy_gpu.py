"""The TRAVERSAL direction rule, re-derived round by round from the device's
own trace.

The host restatement below follows PolicyDecisionMaker::GetInitPolicy /
GetNextPolicy (include/framework/hybrid_policy.h:218-291) with the reference's
bookkeeping: total_workload starts at nedges * edge_factor as a uint32
(framework.cuh:131) and loses each push round's input out-degree sum in uint32
arithmetic (framework.cuh:562); after a pull round the output workload is
forced to 1 (framework.cuh:790), so only the beta test can keep a run in pull.

The counts fed to the rule are this engine's, which are exact: a vertex enters
a BFS frontier once (bit claims), while the reference's default non-atomic BFS
(hybrid_bfs.cu:128-134) counts duplicate queue entries. The BFS rows are
therefore also checked against the oracle's level sizes and level out-degree
sums, which a duplicate entry would inflate.
"""
import numpy as np
import pytest

U32 = 0xFFFFFFFF


def restated_directions(rows, n, m, alpha=15.0, beta=18.0, edge_factor=1.0):
    """Push/pull choice of every round from its input (count, scout)."""
    total = int(m * edge_factor) & U32
    pull_prev = False
    out = []
    for r, row in enumerate(rows):
        count, scout = row["in_nodes"], row["in_edges"]
        if r == 0:
            pull = False  # GetInitPolicy: ASYNC_PUSH_DD
        elif pull_prev and count > n / beta:
            pull = True
        else:
            out_wl = 1 if pull_prev else (scout & U32)
            pull = out_wl > total / alpha
        if not pull:
            total = (total - (scout & U32)) & U32
        out.append("async_pull_dd" if pull else "async_push_dd")
        pull_prev = pull
    return out


def level_profile(hg, lv):
    deg = np.diff(hg.row.astype(np.int64))
    depth = int(lv.max()) + 1
    sizes = np.bincount(lv[lv >= 0], minlength=depth)
    sums = np.bincount(lv[lv >= 0], weights=deg[lv >= 0], minlength=depth).astype(np.int64)
    return sizes, sums


def traversal_rows(e):
    t = e.trace()
    assert t[-1]["variant"] == "end"
    return t[:-1]


@pytest.mark.gpu
@pytest.mark.parametrize("opts", [{}, {"alpha": 4.0, "beta": 6.0}, {"edge_factor": 3.0}, {"block_size": 256}])
def test_bfs_directions_follow_the_rule(sg, orc, gpu, opts):
    hg = sg.rmat_graph(17, 16, symmetrize=False, seed=4)
    e = sg.Engine(sg.Graph(hg), **opts)
    kw = {k: v for k, v in opts.items() if k in ("alpha", "beta", "edge_factor")}
    switched = 0
    for s in sg.pick_sources(hg.row, 12, seed=9):
        s = int(s)
        e.bfs(s)
        lv = e.gather(np.int32)
        assert np.array_equal(lv, orc.bfs(hg.row, hg.col, s))
        rows = traversal_rows(e)
        sizes, sums = level_profile(hg, lv)
        assert len(rows) == len(sizes)
        assert [r["in_nodes"] for r in rows] == sizes.tolist(), s
        assert [r["in_edges"] for r in rows] == sums.tolist(), s
        want = restated_directions(rows, hg.n, hg.m, **kw)
        assert [r["variant"] for r in rows] == want, s
        switched += "async_pull_dd" in want
    assert switched > 0


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["async_push_dd", "auto"])
def test_bfs_duplicate_heads_counted_once(sg, orc, gpu, variant):
    """A vertex whose (raw, unsorted) out-list repeats a head in non-adjacent
    lanes of one same-word run (heads 5, 6, 5, 7, 5, 6): each head enters the
    next frontier once, and SSSP queues each lowered vertex once."""
    n = 100
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(np.bincount([0, 1, 1, 1, 1, 1, 1], minlength=n))
    col = np.array([1, 5, 6, 5, 7, 5, 6], np.uint32)
    w = np.array([1, 4, 2, 3, 1, 9, 2], np.uint32)
    hg = sg.HostGraph(n, row, col, w)
    e = sg.Engine(sg.Graph(hg, relabel=False), force_variant=variant)
    e.bfs(0)
    assert np.array_equal(e.gather(np.int32), orc.bfs(row, col, 0))
    assert [r["in_nodes"] for r in traversal_rows(e)] == [1, 1, 3]
    e.sssp(0)
    assert np.array_equal(e.gather(np.uint32), orc.sssp_dijkstra(row, col, w, 0))
    assert [r["in_nodes"] for r in traversal_rows(e)] == [1, 1, 3]


@pytest.mark.gpu
def test_sssp_directions_follow_the_rule(sg, orc, gpu):
    hg = sg.rmat_graph(16, 16, symmetrize=False, seed=4, weights=(1, 64))
    e = sg.Engine(sg.Graph(hg))
    seen_pull = False
    for s in sg.pick_sources(hg.row, 8, seed=9):
        s = int(s)
        e.sssp(s)
        assert np.array_equal(e.gather(np.uint32), orc.sssp_dijkstra(hg.row, hg.col, hg.w, s))
        rows = traversal_rows(e)
        want = restated_directions(rows, hg.n, hg.m)
        assert [r["variant"] for r in rows] == want, s
        seen_pull |= "async_pull_dd" in want
    assert seen_pull


def test_restatement_pins_reference_cases():
    """The restatement on hand-made counters (CPU): round 0 pushes; a scout
    above total/alpha pulls; after a pull only count > n/beta keeps pulling;
    the workload decrement wraps in uint32."""
    n, m = 1800, 1000
    rows = [{"in_nodes": 1, "in_edges": 10}, {"in_nodes": 5, "in_edges": 80},
            {"in_nodes": 101, "in_edges": 5}, {"in_nodes": 99, "in_edges": 5},
            {"in_nodes": 3, "in_edges": 7}]
    # total 1000 -> 990 after round 0; round 1: 80 > 990/15=66 -> pull;
    # round 2: 101 > 1800/18 = 100 -> pull; round 3: 99 <= 100, out_wl 1 -> push;
    # round 4: 7 <= (990 - 5)/15 -> push
    assert restated_directions(rows, n, m) == ["async_push_dd", "async_pull_dd", "async_pull_dd",
                                               "async_push_dd", "async_push_dd"]
    # uint32 wrap: a push round whose scout exceeds what is left makes the
    # remaining workload huge, so the next round cannot pull on scout
    rows = [{"in_nodes": 1, "in_edges": 50}, {"in_nodes": 9, "in_edges": 60}]
    assert restated_directions(rows, 100, 40) == ["async_push_dd", "async_push_dd"]
