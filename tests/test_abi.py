"""The ctypes mirrors of the C structs have the C layout (sizes and field
offsets as compiled by gcc from include/sepgraph.h)."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "sg_engine_options": ("EngineOptions", ["alpha", "block_size", "grid_blocks", "stage_hubs", "reserved"]),
    "sg_stats": ("Stats", ["kernel_ms", "residual", "status", "reserved"]),
    "sg_round_trace": ("RoundTrace", ["in_edges", "time_us"]),
    "sg_server_config": ("ServerConfig", ["max_rows", "max_batch", "bfs_packed_min", "sssp_packed_min", "mode", "sssp_prio_graphs", "reserved"]),
    "sg_server_stats": ("ServerStats", ["device_bytes", "mean_batch", "rows", "reserved"]),
    "sg_tune_config": ("TuneConfig", ["tune_bfs", "reserved"]),
    "sg_tune_result": ("TuneResult", ["sssp_alpha", "sssp_priority", "bfs_ms", "max_batch", "batch_ms", "reserved"]),
}


def test_struct_layouts(sg, tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sepgraph.h"', "int main(void) {"]
    for cname, (_, fields) in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, c = ln.split()
            got[(a, b)] = int(c)
    for cname, (pyname, fields) in STRUCTS.items():
        cls = getattr(sg, pyname)
        assert C.sizeof(cls) == got[(cname, "size")], cname
        for f in fields:
            assert getattr(cls, f).offset == got[(cname, f)], (cname, f)
