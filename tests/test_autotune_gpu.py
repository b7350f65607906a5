"""The autotuner times real kernels and picks knobs that never change
results: after tuning, BFS levels are still bit-exact and the tuned
configuration is no slower than the default by more than timing noise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_rmat16(sg, orc, gpu, rmat16_sym):
    hg = sg.rmat_graph(16, 16, symmetrize=True, seed=1, weights=(1, 64))
    g = sg.Graph(hg)
    e = sg.Engine(g)
    srcs = sg.pick_sources(hg.row, 6, seed=3)
    r = sg.autotune(e, srcs, latency_budget_ms=10.0)
    assert r["bfs_block_size"] in (256, 512, 1024) and r["sssp_block_size"] in (256, 512, 1024)
    assert r["bfs_alpha"] > 0 and r["bfs_beta"] > 0
    assert 0 < r["bfs_ms"] <= r["bfs_default_ms"] * 1.0001
    assert 0 < r["sssp_ms"] <= r["sssp_default_ms"] * 1.0001
    for k in (sg.QUERY_BFS, sg.QUERY_BC_DEP):
        assert r["max_batch"][k] in (4, 8, 16, 32, 64) and r["batch_qps"][k] > 0
    # SSSP: whole-grid queries back to back (1, 2, 4) or one bit-packed
    # launch (8 and up), within the latency budget
    assert r["max_batch"][sg.QUERY_SSSP] in (1, 2, 4, 8, 16, 32, 64) and r["batch_qps"][sg.QUERY_SSSP] > 0
    assert r["batch_ms"][sg.QUERY_SSSP] <= 10.0 or r["max_batch"][sg.QUERY_SSSP] == 1
    # the engine is left on the tuned BFS options; levels unchanged
    for s in srcs[:3]:
        e.bfs(int(s))
        assert np.array_equal(e.gather(np.int32), orc.bfs(hg.row, hg.col, int(s)))
