"""The C++ facade (include/sepgraph.hpp): a driver written like the
reference's samples (Engine<...> / EngineOptions / LoadGraph / InitGraph /
Start / GatherValue) compiles with g++ against the C ABI, and on the GPU its
results match the oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_sample.cpp")
LIBDIR = os.path.join(ROOT, "sep-graph_amd", "lib")


def build(tmp_path):
    exe = str(tmp_path / "facade_sample")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIBDIR, "-lsepgraph", f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_facade_compiles(tmp_path):
    build(tmp_path)


def test_facade_loader_rejects_bad_file(tmp_path):
    """The facade's ReadGraph reports a malformed file as an error (exit 2)
    before any device work."""
    exe = build(tmp_path)
    bad = tmp_path / "bad.gr"
    bad.write_bytes(b"\x02" + b"\x00" * 31)
    r = subprocess.run([exe, str(bad), "gr", str(tmp_path / "o"), "0", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown file version" in r.stderr


@pytest.mark.gpu
def test_facade_sample_matches_oracle(sg, orc, gpu, tmp_path):
    hg = sg.rmat_graph(12, 16, symmetrize=True, seed=1, weights=(1, 64))
    path = str(tmp_path / "g.gr")
    sg.write_graph(hg, path, "gr")
    exe = build(tmp_path)
    src = 7
    out = str(tmp_path / "out")
    r = subprocess.run([exe, path, "gr", out, str(src), "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lv = np.fromfile(out + ".bfs.bin", np.int32)
    assert np.array_equal(lv, orc.bfs(hg.row, hg.col, src))
    d = np.fromfile(out + ".sssp.bin", np.uint32)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, hg.w, src))
    pr = np.fromfile(out + ".pr.bin", np.float32).astype(np.float64)
    ref, _, _ = orc.pagerank(hg.row, hg.col, 0.85, 1e-6, 10000)
    assert np.abs(pr - ref).sum() / ref.sum() <= 1e-6
    bc = np.fromfile(out + ".bc.bin", np.float32).astype(np.float64)
    hd, _ = orc.bc_single(hg.row, hg.col, src)
    assert np.all(np.abs(bc - hd) <= 1e-5 * np.abs(hd) + 1e-6)


@pytest.mark.gpu
def test_facade_metis_with_generated_weights(sg, orc, gpu, tmp_path):
    """An unweighted METIS file with --gen_weights: SSSP runs on the
    reference drivers' weight sequence."""
    hg = sg.rmat_graph(10, 8, symmetrize=True, seed=3)
    path = str(tmp_path / "g.graph")
    sg.write_graph(hg, path, "metis")
    exe = build(tmp_path)
    out = str(tmp_path / "out")
    r = subprocess.run([exe, path, "metis", out, "3", "1", "100"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    w = sg.gen_weights_std(hg.m, 100)
    d = np.fromfile(out + ".sssp.bin", np.uint32)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.row, hg.col, w, 3))
