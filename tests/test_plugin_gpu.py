"""Apps written against the reference's plugin API (sepgraph::api::AppBase,
include/framework/variants/api.cuh:22-120) run through the facade's Engine
on the generic engine (include/sepgraph_app.hpp). The driver
tests/cpp/plugin_apps.hip holds a BFS app with the hooks of
samples/hybrid_bfs/hybrid_bfs.cu:29-170, an SSSP app with those of
samples/hybrid_sssp/hybrid_sssp.cu:30-140, and a connected-components app
that the library has no kernel for; it is built by __graft_entry__.build()."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bin", "plugin_apps")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/bin/plugin_apps is not built: run __graft_entry__.build()")
    return EXE


def run_app(app, path, out, source=0, undirected=False, variant="auto"):
    r = subprocess.run([_exe(), app, path, out, str(source), "1" if undirected else "0", variant],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.fromfile(out, np.uint32), r.stdout


def test_plugin_driver_built_and_checks_arguments():
    r = subprocess.run([_exe()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["auto", "push", "pull"])
def test_plugin_bfs_matches_oracle(sg, orc, gpu, tmp_path, variant):
    hg = sg.rmat_graph(14, 16, symmetrize=False, seed=2)
    path = str(tmp_path / "g.gr")
    sg.write_graph(hg, path, "gr")
    for s in sg.pick_sources(hg.row, 3, seed=4):
        lv, log = run_app("bfs", path, str(tmp_path / "lv.bin"), int(s), variant=variant)
        ref = orc.bfs(hg.row, hg.col, int(s))
        assert np.array_equal(lv.view(np.int32), ref), (variant, int(s))
        if variant == "auto":
            assert "pull 0" not in log  # the alpha/beta rule switched on RMAT


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["auto", "push"])
def test_plugin_sssp_matches_oracle(sg, orc, gpu, tmp_path, variant):
    hg = sg.rmat_graph(13, 16, symmetrize=False, seed=3, weights=(1, 64))
    path = str(tmp_path / "g.gr")
    sg.write_graph(hg, path, "gr")
    for s in sg.pick_sources(hg.row, 3, seed=5):
        d, _ = run_app("sssp", path, str(tmp_path / "d.bin"), int(s), variant=variant)
        assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s))), (variant, int(s))


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["auto", "push", "pull"])
def test_plugin_connected_components(sg, gpu, tmp_path, variant):
    """A new app through the same API: every vertex gets the smallest id of
    its component (checked against scipy's components)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    rng = np.random.default_rng(7)
    n, m = 5000, 6000  # sparse enough to leave many components
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    row, col = sg.build_csr(n, s, d, symmetrize=True)
    hg = sg.HostGraph(n, row, col, undirected=True)
    path = str(tmp_path / "g.gr")
    sg.write_graph(hg, path, "gr")
    lab, _ = run_app("cc", path, str(tmp_path / "cc.bin"), 0, undirected=True, variant=variant)
    a = csr_matrix((np.ones(len(col)), col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    ncomp, comp = connected_components(a, directed=False)
    smallest = np.full(ncomp, n, np.int64)
    np.minimum.at(smallest, comp, np.arange(n))
    assert np.array_equal(lab.astype(np.int64), smallest[comp])
    assert ncomp > 100
