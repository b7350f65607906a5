"""The hybrid_* command-line drivers (sep-graph_amd/bin, built from
sep-graph_amd/samples): the reference's flag names and defaults
(include/utils/app_skeleton.h:42-87 plus the per-sample flags), its run-info
JSON keys (include/framework/framework.cuh:205-257), its --output formats
(samples/hybrid_*/hybrid_*_host.cpp) and --check."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sep-graph_amd", "bin")

ENGINE_KEYS = {
    "time_input_active_node", "time_output_active_node", "time_input_workload", "time_output_workload",
    "time_queue2bitmap", "time_bitmap2queue", "time_rebuild_worklist", "time_priority_sample",
    "time_sort_worklist", "time_unique_worklist", "time_kernel", "time_total", "time_per_round",
    "num_iteration", "priority_type", "fused_kernel", "max_iteration_reached", "device", "dataset",
    "nnodes", "nedges", "algo_type",
}


def drv(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} is not built: run __graft_entry__.build()")
    return p


def run(name, *args, timeout=120):
    return subprocess.run([drv(name), *args], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("name", ["hybrid_bfs", "hybrid_sssp", "hybrid_pr", "hybrid_bc"])
def test_flags_parse_like_gflags(name, sg, tmp_path):
    """--help lists the reference's flags; an unknown flag or a bad value is
    an error; --stats prints the graph summary and exits before any device
    work (so this runs without a GPU)."""
    r = run(name, "--help")
    assert r.returncode == 0
    for flag in ("max_iteration", "alpha", "beta", "lb_push", "lb_pull", "graphfile", "format", "undirected",
                 "json", "check", "trace", "output", "gen_graph", "gen_weights", "block_size", "edge_factor"):
        assert f"-{flag} (" in r.stdout, (name, flag)
    assert run(name, "--no_such_flag=1").returncode == 1
    assert run(name, "--alpha=abc").returncode == 1
    hg = sg.rmat_graph(8, 8, symmetrize=True, seed=1)
    path = str(tmp_path / "g.gr")
    sg.write_graph(hg, path, "gr")
    r = run(name, f"--graphfile={path}", "--stats")
    assert r.returncode == 0, r.stderr
    assert f"The graph has {hg.n} vertices, and {hg.m} edges" in r.stdout
    # --name value, -name, --noname forms
    r = run(name, "--graphfile", path, "-stats", "--noverbose")
    assert r.returncode == 0 and "The graph has" in r.stdout


def test_per_sample_flags_exist():
    assert "-non_atomic (" in run("hybrid_bfs", "--help").stdout
    assert "-source_node (" in run("hybrid_sssp", "--help").stdout
    out = run("hybrid_pr", "--help").stdout
    for f in ("error", "cut_threshold", "top_ranks", "print_ranks", "norm", "variant", "sparse"):
        assert f"-{f} (" in out
    assert "-src (" in run("hybrid_bc", "--help").stdout


def _graph(sg, tmp_path, scale=11, weights=None, sym=False):
    hg = sg.rmat_graph(scale, 16, symmetrize=sym, seed=3, weights=weights)
    path = str(tmp_path / "g.gr")
    sg.write_graph(hg, path, "gr")
    return hg, path


def _levels_file(path, n):
    out = np.full(n, 0xFFFFFFFF, np.uint64)
    for line in open(path):
        i, v = line.split()
        out[int(i)] = 0xFFFFFFFF if v == "INF" else int(v)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--sparse"], ["--variant=async_pull_dd"], ["--alpha=0"]])
def test_hybrid_bfs_driver(sg, orc, gpu, tmp_path, extra):
    hg, path = _graph(sg, tmp_path)
    js, out = str(tmp_path / "r.json"), str(tmp_path / "lv.txt")
    src = int(sg.pick_sources(hg.row, 1, seed=2)[0])
    r = run("hybrid_bfs", f"--graphfile={path}", f"--source_node={src}", "--check", "--trace", f"--json={js}",
            f"--output={out}", *extra)
    assert r.returncode == 0, r.stderr
    assert "total errors: 0" in r.stdout and "Overall: Test passed" in r.stdout
    ref = orc.bfs(hg.row, hg.col, src).astype(np.int64)
    got = _levels_file(out, hg.n)
    assert np.array_equal(np.where(ref < 0, 0xFFFFFFFF, ref), got)
    j = json.load(open(js))
    assert ENGINE_KEYS <= set(j) and j["non_atomic"] == "YES"
    assert j["nnodes"] == hg.n and j["nedges"] == hg.m and j["algo_type"] == "TRAVERSAL_SCHEME"
    assert j["num_iteration"] == int(ref.max()) + 1 == len(j["round"])
    # round 0 pushes (GetInitPolicy) unless a variant is forced
    first = "ASYNC_PULL_DD" if extra == ["--variant=async_pull_dd"] else "ASYNC_PUSH_DD"
    assert j["round"][0].startswith(f"Round: 0 Policy to execute: {first}")
    if extra == ["--sparse"]:
        assert j["force_variant"] == "ASYNC_PUSH_DD" and j["force_push_load_balancing"] == "NONE"
    else:
        assert j.get("force_pull_load_balancing") == "NONE"
    # source clamped into [0, n - 1] like the reference driver
    r = run("hybrid_bfs", f"--graphfile={path}", "--source_node=99999999", "--check")
    assert r.returncode == 0 and "total errors: 0" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--sparse"], ["--prio_delta=40"], ["--lb_push=none", "--lb_pull=none"]])
def test_hybrid_sssp_driver(sg, orc, gpu, tmp_path, extra):
    hg, path = _graph(sg, tmp_path, weights=(1, 64))
    js, out = str(tmp_path / "r.json"), str(tmp_path / "d.txt")
    src = int(sg.pick_sources(hg.row, 1, seed=2)[0])
    r = run("hybrid_sssp", f"--graphfile={path}", f"--source_node={src}", "--check", f"--json={js}",
            f"--output={out}", *extra)
    assert r.returncode == 0, r.stderr
    assert "total errors: 0" in r.stdout
    ref = orc.sssp_dijkstra(hg.row, hg.col, hg.w, src).astype(np.uint64)
    assert np.array_equal(_levels_file(out, hg.n), ref)
    j = json.load(open(js))
    assert ENGINE_KEYS <= set(j)
    if extra in (["--sparse"], ["--prio_delta=40"]):
        assert j["priority_type"] == "low_high"


@pytest.mark.gpu
def test_hybrid_sssp_generated_weights(sg, orc, gpu, tmp_path):
    """--gen_weights on an unweighted file: the reference drivers' weight
    sequence (traversal.h:173-187)."""
    hg, path = _graph(sg, tmp_path)
    out = str(tmp_path / "d.txt")
    r = run("hybrid_sssp", f"--graphfile={path}", "--gen_weights", "--gen_weight_range=50", "--source_node=3",
            f"--output={out}", "--check")
    assert r.returncode == 0 and "total errors: 0" in r.stdout, r.stdout + r.stderr
    w = sg.gen_weights_std(hg.m, 50)
    ref = orc.sssp_dijkstra(hg.row, hg.col, w, 3).astype(np.uint64)
    assert np.array_equal(_levels_file(out, hg.n), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--variant=sync_pull_td"], ["--variant=async_push_dd"]])
def test_hybrid_pr_driver(sg, orc, gpu, tmp_path, extra):
    hg, path = _graph(sg, tmp_path)
    js, out = str(tmp_path / "r.json"), str(tmp_path / "pr.txt")
    r = run("hybrid_pr", f"--graphfile={path}", "--error=1e-7", "--check", f"--json={js}", f"--output={out}",
            "--print_ranks", "--norm", *extra)
    assert r.returncode == 0, r.stderr
    assert "total errors: 0" in r.stdout
    j = json.load(open(js))
    assert ENGINE_KEYS <= set(j) and j["algo_type"] == "ITERATIVE_SCHEME"
    assert abs(j["error_tolerance"] - 1e-7) < 1e-12
    lines = open(out).read().splitlines()
    assert lines[0].startswith("ALPHA") and lines[1] == f"RANKS 1--{hg.n} of {hg.n}"
    ref, _, _ = orc.pagerank(hg.row, hg.col, 0.85, 1e-10, 10000)
    nodes = [int(x.split()[0]) for x in lines[2:]]
    vals = np.array([float(x.split()[1]) for x in lines[2:]])
    assert sorted(nodes) == list(range(hg.n))
    assert np.all(np.diff(vals) <= 0)  # highest rank first
    assert np.abs(vals - ref[nodes]).sum() / ref.sum() < 1e-5


@pytest.mark.gpu
def test_hybrid_bc_driver(sg, orc, gpu, tmp_path):
    hg, path = _graph(sg, tmp_path, sym=True)
    js, out = str(tmp_path / "r.json"), str(tmp_path / "bc.txt")
    r = run("hybrid_bc", f"--graphfile={path}", "--undirected", "--source_node=5", "--check", f"--json={js}",
            f"--output={out}")
    assert r.returncode == 0, r.stderr
    assert "Overall: Test passed" in r.stdout
    got = np.array([float(x.split()[1]) for x in open(out)])
    want, _ = orc.bc_single(hg.row, hg.col, 5)
    assert np.all(np.abs(got - want) <= 1e-5 * np.abs(want) + 1e-4)
    assert ENGINE_KEYS <= set(json.load(open(js)))


@pytest.mark.gpu
def test_generated_graphs(sg, orc, gpu, tmp_path):
    """--gen_graph with the random (0), chain (2) and clique (3) methods. The
    500-vertex chain needs 500 rounds, past the reference's default
    --max_iteration of 200 (which stops the run there, as framework.cuh:359-364
    does), so the cap is raised."""
    for method, nn in ((0, 3000), (2, 500), (3, 60)):
        r = run("hybrid_bfs", "--gen_graph", f"--gen_method={method}", f"--gen_nnodes={nn}", "--gen_factor=6",
                "--max_iteration=1000", "--check")
        assert r.returncode == 0 and "total errors: 0" in r.stdout, (method, r.stdout, r.stderr)
