"""PageRank on the GPU through the C ABI against the FP64 oracle
(power iteration with uniform teleport and uniform dangling redistribution),
for both schedules: topology-driven pull (the reference's RunSyncPullTD,
include/framework/algo_variants.cuh:753-823) and data-driven residual
propagation (RunAsyncPushDD with the PageRank app, samples/hybrid_pr/hybrid_pr.cu:66-110).

Tolerances: ranks are FP32, so the bar is relative L1 error <= 1e-6 against
the FP64 oracle. The pull schedule stops on the same rule as the oracle
(first iteration whose L1 change is < tol) and is compared with the oracle run
for the same number of iterations. The data-driven schedule converges by a
different route and is compared with an oracle converged to 1e-12."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu
D = 0.85


def rel_l1(x, ref):
    return float(np.abs(x.astype(np.float64) - ref).sum() / np.abs(ref).sum())


def power_step(row, col, x, d=D):
    """One FP64 power-iteration step F(x) (numpy, any size)."""
    n = len(row) - 1
    x = x.astype(np.float64)
    deg = np.diff(row.astype(np.int64))
    contrib = np.where(deg > 0, x / np.maximum(deg, 1), 0.0)
    src = np.repeat(np.arange(n), deg)
    y = np.bincount(col, weights=contrib[src], minlength=n)
    dang = x[deg == 0].sum()
    return (1 - d) / n + d * (y + dang / n)


def run_pr(sg, hg, mode, tol=1e-6, max_iter=0, **opts):
    g = sg.Graph(hg)
    e = sg.Engine(g, **opts)
    e.pagerank(D, tol, max_iter, mode)
    x = e.gather(np.float32)
    st = e.stats()
    return x, st, e


def check_td(orc, hg, x, st, tol=1e-6):
    ref, it, res = orc.pagerank(hg.row, hg.col, D, tol, 100000)
    # same stopping rule; the ranks are FP32, so the L1 change near the
    # threshold carries rounding noise of about n ulp(x), which can move the
    # stop by a few iterations where the change decays slowly (bipartite
    # graphs oscillate, shrinking by the damping factor per iteration)
    assert abs(st["iterations"] - it) <= 3, (st["iterations"], it, res)
    same, _, _ = orc.pagerank(hg.row, hg.col, D, 0.0, st["iterations"])
    assert rel_l1(x, same) <= 1e-6
    if st["iterations"] == it:
        assert rel_l1(x, ref) <= 1e-6
    assert st["residual"] < tol
    assert abs(float(x.astype(np.float64).sum()) - 1.0) < 1e-5


SMALL = {
    "path": lambda: G.path(40),
    "cycle_dir": lambda: G.cycle(50, undirected=False),
    "star": lambda: G.star(30),
    "star_out": lambda: G.star(30, undirected=False),  # leaves are dangling
    "complete": lambda: G.complete(12),
    "dir_random": lambda: G.random_graph(600, 4000, 3),
    "sparse_dir": lambda: G.random_graph(3000, 2500, 5),  # many dangling and isolated vertices
    "und_random": lambda: G.random_graph(500, 1500, 4, undirected=True),
    "grid": lambda: G.grid(19, 21),
}


@pytest.mark.parametrize("name", sorted(SMALL))
def test_pr_pull_td_small(sg, orc, gpu, name):
    row, col = SMALL[name]()
    hg = sg.HostGraph(len(row) - 1, row, col, None, undirected=False)
    x, st, _ = run_pr(sg, hg, sg.PR_PULL_TD)
    check_td(orc, hg, x, st)


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("variant", ["auto", "async_push_dd", "async_pull_dd"])
def test_pr_dd_small(sg, orc, gpu, name, variant):
    row, col = SMALL[name]()
    hg = sg.HostGraph(len(row) - 1, row, col, None, undirected=False)
    x, st, _ = run_pr(sg, hg, sg.PR_PUSH_DD, tol=1e-9, force_variant=variant)
    ref, _, _ = orc.pagerank(hg.row, hg.col, D, 1e-13, 100000)
    assert rel_l1(x, ref) <= 1e-6, (name, variant, st)
    if variant == "async_push_dd":
        assert st["pull_rounds"] == 0
    if variant == "async_pull_dd":
        assert st["push_rounds"] == 0


def test_pr_closed_forms(sg, gpu):
    # directed cycle and complete graph: uniform 1/n
    for row, col in (G.cycle(64, undirected=False), G.complete(9)):
        n = len(row) - 1
        hg = sg.HostGraph(n, row, col, None)
        for mode in (sg.PR_PULL_TD, sg.PR_PUSH_DD):
            x, _, _ = run_pr(sg, hg, mode, tol=1e-9)
            np.testing.assert_allclose(x, 1.0 / n, rtol=1e-6)
    # star K_{1,k}: y = (1-d)/n + d k x, x = (1-d)/n + d y / k
    k = 20
    row, col = G.star(k)
    n = k + 1
    A = np.array([[1.0, -D * k], [-D / k, 1.0]])
    yc, xl = np.linalg.solve(A, [(1 - D) / n, (1 - D) / n])
    hg = sg.HostGraph(n, row, col, None)
    for mode in (sg.PR_PULL_TD, sg.PR_PUSH_DD):
        x, _, _ = run_pr(sg, hg, mode, tol=1e-10)
        np.testing.assert_allclose(x[0], yc, rtol=1e-6)
        np.testing.assert_allclose(x[1:], xl, rtol=1e-6)
    # edgeless graph: every vertex dangling, ranks 1/n
    n = 1000
    hg = sg.HostGraph(n, np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32), None)
    for mode in (sg.PR_PULL_TD, sg.PR_PUSH_DD):
        x, _, _ = run_pr(sg, hg, mode)
        np.testing.assert_allclose(x, 1.0 / n, rtol=1e-6)


def test_pr_max_iter_and_errors(sg, orc, gpu):
    row, col = G.random_graph(2000, 12000, 9)
    hg = sg.HostGraph(len(row) - 1, row, col, None)
    x, st, e = run_pr(sg, hg, sg.PR_PULL_TD, tol=1e-12, max_iter=7)
    assert st["iterations"] == 7
    ref, _, _ = orc.pagerank(row, col, D, 0.0, 7)
    assert rel_l1(x, ref) <= 1e-6
    for bad in ((1.0, 1e-6), (-0.1, 1e-6), (0.85, 0.0)):
        with pytest.raises(sg.SepGraphError):
            e.pagerank(bad[0], bad[1], 0, sg.PR_PULL_TD)
    with pytest.raises(sg.SepGraphError):
        e.pagerank(D, 1e-6, 0, 7)


@pytest.mark.parametrize("bs", [256, 512, 1024])
def test_pr_rmat16_block_sizes(sg, orc, gpu, bs):
    hg = sg.rmat_graph(16, 16, symmetrize=False, seed=1)
    x, st, e = run_pr(sg, hg, sg.PR_PULL_TD, block_size=bs)
    check_td(orc, hg, x, st)
    # repeated runs give identical bits
    e.pagerank(D, 1e-6, 0, sg.PR_PULL_TD)
    assert np.array_equal(e.gather(np.float32), x)
    e.pagerank(D, 1e-8, 0, sg.PR_PUSH_DD)
    xd = e.gather(np.float32)
    ref, _, _ = orc.pagerank(hg.row, hg.col, D, 1e-13, 100000)
    assert rel_l1(xd, ref) <= 1e-6
    # the ITERATIVE scheme (LDS bitmap copy, sampled priority) at this block size
    e.pagerank(D, 1e-8, 0, sg.PR_ITERATIVE)
    assert rel_l1(e.gather(np.float32), ref) <= 1e-6
    e2 = sg.Engine(sg.Graph(hg), block_size=bs, force_variant="async_push_td", priority=2, cut_threshold=0.3)
    e2.pagerank(D, 1e-8, 0, sg.PR_ITERATIVE)
    assert rel_l1(e2.gather(np.float32), ref) <= 1e-6


def test_pr_rmat20_directed(sg, orc, gpu):
    hg = sg.rmat_graph(20, 16, symmetrize=False, seed=1)
    x, st, e = run_pr(sg, hg, sg.PR_PULL_TD)
    check_td(orc, hg, x, st)
    e.pagerank(D, 1e-8, 0, sg.PR_PUSH_DD)
    xd = e.gather(np.float32)
    sd = e.stats()
    ref, _, _ = orc.pagerank(hg.row, hg.col, D, 1e-13, 100000)
    assert rel_l1(xd, ref) <= 1e-6, sd
    assert sd["push_rounds"] + sd["pull_rounds"] == sd["rounds"]


def test_pr_rmat23_directed_full_size(sg, gpu):
    """PageRank on directed RMAT-23 (n = 8,388,608): both schedules. At this
    size the check is the fixed-point equation itself: one FP64 power step
    from the GPU's ranks moves them by less than the stopping tolerance
    (plus FP32 storage rounding), ranks sum to 1 and are >= (1-d)/n, and the
    two schedules agree."""
    hg = sg.rmat_graph(23, 16, symmetrize=False, seed=1)
    n = hg.n
    x, st, e = run_pr(sg, hg, sg.PR_PULL_TD)
    assert st["residual"] < 1e-6 and 5 <= st["iterations"] <= 200, st
    xd64 = x.astype(np.float64)
    assert abs(xd64.sum() - 1.0) < 1e-5
    assert xd64.min() >= (1 - D) / n * (1 - 1e-6)
    step = power_step(hg.row, hg.col, x)
    assert np.abs(step - xd64).sum() < 1e-6 + 2 ** -23
    e.pagerank(D, 1e-8, 0, sg.PR_PUSH_DD)
    xd = e.gather(np.float32)
    assert np.abs(power_step(hg.row, hg.col, xd) - xd.astype(np.float64)).sum() < 1e-6
    # the pull result is within its stopping error of the fixed point
    assert rel_l1(x, xd.astype(np.float64)) < 1e-5


# ---------------------------------------------------- ITERATIVE scheme
ITER_ORDER = ["sync_pull_td", "async_push_td", "async_push_dd"]


def iterative_decisions_ok(trace):
    """Re-derive every round's variant from the trace with the reference's
    rule (hybrid_policy.h:130-264): rounds 0..2 inspect the variant set in
    order; later rounds take the first minimum of the predicted times -- the
    last measured time of each topology-driven variant, and for
    ASYNC_PUSH_DD its last round's time per input edge times the current
    worklist's out-degree sum (the next round's input edges). Times are the
    barrier-clock round durations in the trace. A choice that differs from
    the restatement must be a tie within float rounding."""
    rows = [r for r in trace if r["variant"] in ITER_ORDER]
    t_prev = 0.0
    times, last, last_in, tpw = {}, None, 0.0, 0.0
    for k, r in enumerate(rows):
        if k < 3:
            assert r["variant"] == ITER_ORDER[k], (k, r)
        else:
            if last == "async_push_dd":
                tpw = np.float32(times[last]) / np.float32(last_in)
            pred = [np.float32(times["sync_pull_td"]), np.float32(times["async_push_td"]),
                    np.float32(tpw) * np.float32(r["in_edges"])]
            best = int(np.argmin(pred))
            got = ITER_ORDER.index(r["variant"])
            if got != best:
                assert abs(pred[got] - pred[best]) <= 1e-5 * max(abs(pred[best]), 1e-30), (k, pred, r)
        dur = r["time_us"] - t_prev
        t_prev = r["time_us"]
        assert dur > 0, (k, r)
        times[r["variant"]] = dur
        last = r["variant"]
        last_in = r["in_edges"] if last == "async_push_dd" else None
    return rows


@pytest.mark.parametrize("name", sorted(SMALL))
def test_pr_iterative_small(sg, orc, gpu, name):
    row, col = SMALL[name]()
    hg = sg.HostGraph(len(row) - 1, row, col, None, undirected=False)
    x, st, e = run_pr(sg, hg, sg.PR_ITERATIVE, tol=1e-9)
    ref, _, _ = orc.pagerank(hg.row, hg.col, D, 1e-13, 100000)
    assert rel_l1(x, ref) <= 1e-6, (name, st)
    rows = e.trace(4096)
    for k, r in enumerate(rows[: min(3, st["rounds"])]):
        assert r["variant"] == ITER_ORDER[k]


@pytest.mark.parametrize("directed", [True, False])
def test_pr_iterative_rmat18_policy(sg, orc, gpu, directed):
    """Each round's variant follows the ITERATIVE rule given the round times
    the device measured, and the ranks match the FP64 oracle."""
    hg = sg.rmat_graph(18, 16, symmetrize=not directed, seed=2)
    x, st, e = run_pr(sg, hg, sg.PR_ITERATIVE, tol=1e-8)
    ref, _, _ = orc.pagerank(hg.row, hg.col, D, 1e-13, 100000)
    assert rel_l1(x, ref) <= 1e-6, st
    rows = iterative_decisions_ok(e.trace(4096))
    assert len(rows) == st["rounds"] >= 4
    assert st["pull_rounds"] == sum(r["variant"] == "sync_pull_td" for r in rows)
    # identical inputs, identical results whatever the timings chose
    e.pagerank(D, 1e-8, 0, sg.PR_ITERATIVE)
    assert rel_l1(e.gather(np.float32), ref) <= 1e-6


@pytest.mark.parametrize("variant", ITER_ORDER)
def test_pr_iterative_forced(sg, orc, gpu, variant):
    hg = sg.rmat_graph(16, 16, symmetrize=False, seed=1)
    x, st, e = run_pr(sg, hg, sg.PR_ITERATIVE, tol=1e-8, force_variant=variant)
    ref, _, _ = orc.pagerank(hg.row, hg.col, D, 1e-13, 100000)
    assert rel_l1(x, ref) <= 1e-6, st
    rows = [r for r in e.trace(4096) if r["variant"] != "end"]  # the closing row has no variant
    assert len(rows) == st["rounds"] and all(r["variant"] == variant for r in rows)


def test_pr_iterative_rejects_other_variants(sg, gpu):
    row, col = G.random_graph(500, 3000, 3)
    hg = sg.HostGraph(len(row) - 1, row, col, None)
    e = sg.Engine(sg.Graph(hg), force_variant="async_pull_dd")
    with pytest.raises(sg.SepGraphError) as ei:
        e.pagerank(D, 1e-6, 0, sg.PR_ITERATIVE)
    assert ei.value.code == -8


@pytest.mark.parametrize("name", ["dir_random", "sparse_dir", "star_out", "grid", "rmat16"])
@pytest.mark.parametrize("cut", [0.2, 0.6])
def test_pr_iterative_sampling_priority(sg, orc, gpu, name, cut):
    """Sample-based priority (SetSampleBasedPriority): every ASYNC_PUSH_TD
    round pushes only residuals at or above a sampled quantile; residuals held
    back are picked up by later rounds, so the ranks still converge to the
    oracle's."""
    if name == "rmat16":
        hg = sg.rmat_graph(16, 16, symmetrize=False, seed=1)
    else:
        row, col = SMALL[name]()
        hg = sg.HostGraph(len(row) - 1, row, col, None, undirected=False)
    ref, _, _ = orc.pagerank(hg.row, hg.col, D, 1e-13, 100000)
    for variant in ("async_push_td", "auto"):
        x, st, _ = run_pr(sg, hg, sg.PR_ITERATIVE, tol=1e-9, force_variant=variant, priority=2, cut_threshold=cut)
        assert rel_l1(x, ref) <= 1e-6, (name, cut, variant, st)
        assert st["status"] == 0


def test_pr_sampling_priority_rejects_bad_cut(sg, gpu):
    row, col = G.random_graph(500, 3000, 3)
    hg = sg.HostGraph(len(row) - 1, row, col, None)
    for cut in (0.0, 1.0):
        e = sg.Engine(sg.Graph(hg), priority=2, cut_threshold=cut)
        with pytest.raises(sg.SepGraphError) as ei:
            e.pagerank(D, 1e-6, 0, sg.PR_ITERATIVE)
        assert ei.value.code == -2


def test_pr_sampling_priority_holds_back(sg, orc, gpu):
    """The priority actually filters: with the top 40 % of sampled residuals
    pushed per round, forced ASYNC_PUSH_TD activates fewer vertices in its
    second round than without priority, and needs more rounds to converge."""
    hg = sg.rmat_graph(16, 16, symmetrize=False, seed=1)
    _, plain, e0 = run_pr(sg, hg, sg.PR_ITERATIVE, tol=1e-9, force_variant="async_push_td")
    t0 = [r for r in e0.trace(4096) if r["variant"] == "async_push_td"]
    _, prio, e1 = run_pr(sg, hg, sg.PR_ITERATIVE, tol=1e-9, force_variant="async_push_td", priority=2,
                         cut_threshold=0.4)
    t1 = [r for r in e1.trace(4096) if r["variant"] == "async_push_td"]
    assert t1[1]["in_nodes"] < t0[1]["in_nodes"], (t0[:3], t1[:3])
    assert prio["rounds"] > plain["rounds"], (plain, prio)
