"""Benchmark of the hybrid traversal hot path on one MI355X.

Workload (N=1): plain single-source BFS from 64 seeded non-isolated sources on
a directed RMAT scale-22 graph, edge factor 16, (a,b,c,d) = (0.57,0.19,0.19,0.05),
self-loops and duplicates removed, uint32 weights in [1,64). One step = the 64
queries back to back on one HIP stream; inputs are resident in HBM before the
timed region. value = aggregate GTEPS = sum of m_r (out-degree of reached
vertices, Graph500 convention) over all timed queries / wall time.

roofline: algorithmic bytes of the BFS kernel per launch, 8*n_r + 4*m_r + 4*n,
over its average launch duration from HIP events recorded on the launch stream.
cpu_baseline: the oracle's single-threaded FIFO BFS (the reference's BFSHost)
on a bounded sample of the same sources, timed on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sep-graph_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 counter summary
    (profiles/<name>.txt, a 'traffic_bytes_per_launch = X' line), or None."""
    path = os.path.join(ROOT, "profiles", name + ".txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_launch"):
            try:
                return float(line.split("=", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(hg, sources, gpu_levels, budget_s):
    """Oracle BFS (single thread) over as many sources as fit in budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    deg = np.diff(hg.row.astype(np.int64))
    tot_edges, tot_t, done, exact = 0, 0.0, 0, 0
    for i, s in enumerate(sources):
        t0 = time.perf_counter()
        lv = pyoracle.bfs(hg.row, hg.col, int(s))
        dt = time.perf_counter() - t0
        tot_t += dt
        tot_edges += int(deg[lv >= 0].sum())
        done += 1
        if i in gpu_levels:
            exact += int(np.array_equal(lv, gpu_levels[i]))
        if tot_t > budget_s:
            break
    return {
        "value": tot_edges / tot_t / 1e9,
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} of {len(sources)} BFS sources on the same graph, oracle FIFO BFS "
                  f"(restating BFSHost), {tot_t:.1f} s",
        "bit_exact_vs_gpu": f"{exact}/{min(done, len(gpu_levels))}",
        "host": host_info(),
    }


def host_info():
    """The host CPU the baseline ran on: model name and hardware threads."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "threads": os.cpu_count(), "threads_usable": len(os.sched_getaffinity(0))}


def cpu_sssp_baseline(hg, sources, gpu_dist, budget_s):
    """Oracle label-correcting SSSP (single thread, the reference's
    SSSPHostNaive) over as many sources as fit in budget_s, in the same
    GTEPS convention as the GPU figure (out-edges of reached vertices)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    deg = np.diff(hg.row.astype(np.int64))
    tot_edges, tot_t, done, exact = 0, 0.0, 0, 0
    for i, s in enumerate(sources):
        t0 = time.perf_counter()
        d = pyoracle.sssp_fifo(hg.row, hg.col, hg.w, int(s))
        tot_t += time.perf_counter() - t0
        tot_edges += int(deg[d != 0xFFFFFFFF].sum())
        done += 1
        if i in gpu_dist:
            exact += int(np.array_equal(d, gpu_dist[i]))
        if tot_t > budget_s:
            break
    return {
        "value": tot_edges / tot_t / 1e9,
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} of {len(sources)} SSSP sources on the same graph, oracle FIFO label-correcting "
                  f"SSSP (restating SSSPHostNaive), {tot_t:.1f} s",
        "bit_exact_vs_gpu": f"{exact}/{min(done, len(gpu_dist))}",
    }


def draw_mix(rng, pools, share, count):
    """count queries of the 60/30/10 BFS/SSSP/BC-dependency mix: graph index
    drawn with probabilities `share`, source uniform over that graph's pool."""
    gi = rng.choice(len(pools), size=count, p=share)
    kinds = rng.choice([0, 1, 2], size=count, p=[0.6, 0.3, 0.1])
    srcs = np.array([pools[k][rng.integers(len(pools[k]))] for k in gi], np.uint32)
    return gi.astype(np.int32), kinds.astype(np.int32), srcs


def serving_bench(sg, graphs, hgs, engines, share, label, rates, count, plain_n, seed, mode=0, prio_graphs=0):
    """Queries/s and latency of the server against plain one-at-a-time
    execution on the same resident graphs: a seeded Poisson stream of the
    60/30/10 BFS/SSSP/BC-dependency mix, sources uniform over non-isolated
    vertices. Plain = the single-query entry points (sg_bfs, sg_sssp,
    sg_bc_dependency with one source) one query at a time, synchronised."""
    rng = np.random.default_rng(seed)
    pools = [sg.pick_sources(h.row, min(1 << 16, h.n // 4), seed=seed + 1 + i) for i, h in enumerate(hgs)]
    gi, kinds, srcs = draw_mix(rng, pools, share, plain_n)
    t0 = time.perf_counter()
    for g_, k, s in zip(gi, kinds, srcs):
        e = engines[g_]
        if k == 0:
            e.bfs(int(s))
        elif k == 1:
            e.sssp(int(s))
        else:
            e.bc_dependency([int(s)], 1)
        e.sync()
    plain_qps = plain_n / (time.perf_counter() - t0)
    log(f"[{label}] plain one-at-a-time: {plain_qps:.1f} queries/s")
    srv = sg.Server(graphs, memory_budget=16 << 30, max_rows=1024, mode=mode, sssp_prio_graphs=prio_graphs)
    # warm the serving kernel (residency probe done at creation; first launch)
    wg, wk, ws = draw_mix(rng, pools, share, 16)
    srv.replay(wk, ws, [0.0] * 16, graphs=wg)
    runs = []
    for rate in rates:
        gi, kinds, srcs = draw_mix(rng, pools, share, count)
        arr = np.cumsum(rng.exponential(1.0 / rate, count))
        before = srv.stats()
        t0 = time.perf_counter()
        lat = srv.replay(kinds, srcs, arr, graphs=gi)
        wall = time.perf_counter() - t0
        after = srv.stats()
        nb = after["batches"] - before["batches"]
        run = {
            "offered_qps": rate,
            "achieved_qps": round(count / wall, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "launches": int(nb),
            "queries": count,
        }
        if len(graphs) > 1:
            for j in range(len(graphs)):
                sel = gi == j
                if sel.any():
                    run[f"p99_ms_graph{j}"] = round(float(np.percentile(lat[sel], 99)), 3)
        log(f"[{label}] {rate:.0f} q/s offered: {run}")
        runs.append(run)
        if run["achieved_qps"] < 0.8 * rate and run["p99_ms"] > 200.0:
            break  # saturated: higher rates only queue longer
    st = srv.stats()
    srv.close()
    best = max(r["achieved_qps"] for r in runs)
    ok50 = [r["achieved_qps"] for r in runs if r["p99_ms"] <= 50.0]
    best50 = max(ok50) if ok50 else None
    return {
        "workload": label,
        "mix": "60% BFS / 30% SSSP / 10% BC-dependency, Poisson arrivals",
        "mode": ["slots (serving kernel, one query slot per XCD)", "bit-packed batches"][mode],
        "graph_share": [round(float(x), 4) for x in share],
        "plain_qps": round(plain_qps, 1),
        "plain_queries": plain_n,
        "served_qps": best,
        "speedup_vs_plain": round(best / plain_qps, 2),
        "served_qps_p99_under_50ms": best50,
        "speedup_vs_plain_p99_under_50ms": None if best50 is None else round(best50 / plain_qps, 2),
        "runs": runs,
        "queries_per_rate": count,
        "device_bytes": int(st["device_bytes"]),
    }


def lone_times(sg, e, hg, k, base):
    """Mean milliseconds of single BFS, SSSP (with and without the near/far
    priority schedule) and BC-dependency queries from k seeded non-isolated
    sources, each synchronised, after one untimed query of each kind; the
    engine is left on the options `base`."""
    src = [int(s) for s in sg.pick_sources(hg.row, k, seed=77)]
    runs = [
        ("bfs_ms", lambda s: e.bfs(s), {}),
        ("sssp_ms", lambda s: e.sssp(s), {}),
        ("sssp_priority_ms", lambda s: e.sssp(s), {"priority": 1}),
        ("bc_dependency_ms", lambda s: e.bc_dependency([s], 1), {}),
    ]
    out = {}
    for name, run, opts in runs:
        e.set_options(**opts)
        run(src[0])
        e.sync()
        t0 = time.perf_counter()
        for s in src:
            run(s)
            e.sync()
        out[name] = round((time.perf_counter() - t0) * 1e3 / len(src), 3)
    e.set_options(**base)
    return out


def config5(sg, args, rates):
    """Serving on a directed RMAT-24 (edge factor 16, weights in [1,64)) and
    the 2048x2048 grid (10% of edges deleted, weights in [1,1000)), resident
    together in one server; each query's graph is drawn in proportion to the
    graphs' non-isolated vertex counts, so sources are uniform over the
    non-isolated vertices of both."""
    t0 = time.time()
    h24 = sg.rmat_graph(24, 16, symmetrize=False, seed=1, weights=(1, 64), weight_seed=2)
    hgrid = sg.grid_graph(2048, 2048, p_delete=0.1, seed=1, weights=(1, 1000), weight_seed=2)
    g24, ggrid = sg.Graph(h24), sg.Graph(hgrid)
    log(f"[config5] RMAT-24 n={h24.n} m={h24.m}, grid n={hgrid.n} m={hgrid.m}, resident in {time.time() - t0:.1f}s "
        f"({(g24.info()[2] + ggrid.info()[2]) / 1e9:.2f} GB)")
    live = [int((np.diff(h.row.astype(np.int64)) > 0).sum()) for h in (h24, hgrid)]
    share = np.array(live, np.float64) / sum(live)
    # the grid's SSSP runs the two-level priority schedule, the reference's
    # setting for road-network-like graphs (hybrid_sssp --sparse); RMAT-24's
    # does not (slower there), in the server and in the plain baseline alike
    opts = [{}, {"priority": 1}]
    engines = [sg.Engine(g24, **opts[0]), sg.Engine(ggrid, **opts[1])]
    lone = {"rmat24-directed": lone_times(sg, engines[0], h24, 8, opts[0]),
            "grid2048": lone_times(sg, engines[1], hgrid, 8, opts[1])}
    log(f"[config5] one query at a time: {lone}")
    out = serving_bench(sg, [g24, ggrid], [h24, hgrid], engines, share, "rmat24-directed+grid2048", rates,
                        args.config5_count, args.serve_plain, seed=40, prio_graphs=0b10)
    out["sssp_priority"] = {"rmat24-directed": False, "grid2048": True}
    out["lone"] = lone
    for e in engines:
        e.close()
    g24.close()
    ggrid.close()
    return out


def kernel_sweep(sg, e, hg, sources, reps):
    """Each traversal kernel `reps` times on the benchmark graph, one kind
    after another (k_bfs, k_sssp, k_pr_pull, k_pr_dd, k_bc1, k_bc, k_msbfs,
    k_mssssp), for per-kernel rocprofv3 statistics and counters."""
    src = [int(s) for s in sources]
    steps = [
        ("k_bfs", lambda i: e.bfs(src[i % len(src)])),
        ("k_sssp", lambda i: e.sssp(src[i % len(src)])),
        ("k_pr_pull", lambda i: e.pagerank(0.85, 1e-6, 0, sg.PR_PULL_TD)),
        ("k_pr_dd", lambda i: e.pagerank(0.85, 1e-6, 0, sg.PR_PUSH_DD)),
        ("k_bc1", lambda i: e.bc_dependency([src[i % len(src)]], 1)),
        ("k_bc", lambda i: e.bc(src[:64], 64)),
        ("k_msbfs", lambda i: e.bfs_multi(src[:64], 64)),
        ("k_mssssp", lambda i: e.sssp_multi(src[:64], 64)),
    ]
    for name, run in steps:
        t0 = time.perf_counter()
        for i in range(reps):
            run(i)
            e.sync()
        log(f"{name}: {reps} runs in {(time.perf_counter() - t0) * 1e3:.1f} ms")
    print(json.dumps({"kernels": [n for n, _ in steps], "reps": reps, "n": hg.n, "m": hg.m}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sssp-sources", type=int, default=16)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-rates", type=str, default="500,1000,1500,2000,2250,2500,2750,3000,5000",
                    help="Poisson arrival rates (queries/s) for the RMAT serving run (stops once saturated)")
    ap.add_argument("--config5-rates", type=str, default="50,80,120,160,200,1000",
                    help="Poisson arrival rates (queries/s) for the RMAT-24 + grid serving run (stops once "
                         "saturated; its configuration names 1k-50k q/s, above what one GPU serves)")
    ap.add_argument("--serve-count", type=int, default=2000, help="queries per arrival rate")
    ap.add_argument("--config5-count", type=int, default=2000, help="queries per arrival rate, RMAT-24 + grid")
    ap.add_argument("--serve-plain", type=int, default=200, help="queries timed one at a time (plain baseline)")
    ap.add_argument("--config5", type=int, default=1,
                    help="also serve RMAT-24 and the 2048x2048 grid resident together (0: skip)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--kernels", type=int, default=0,
                    help="only run every traversal kernel this many times on the benchmark graph "
                         "(for rocprofv3 kernel statistics and counters), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the engine runs on a single MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import sepgraph as sg

    if not torch.cuda.is_available() or sg.lib().sg_device_count() < 1:
        print("bench.py needs a visible MI355X", file=sys.stderr)
        sys.exit(1)
    if args.copy_mb:
        # A copy of known size: its FETCH_SIZE / WRITE_SIZE readings calibrate
        # the counters used for roofline.traffic.
        nbytes = args.copy_mb << 20
        x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(4):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nbytes, "launches": 4}))
        return

    t0 = time.time()
    hg = sg.rmat_graph(args.scale, 16, symmetrize=False, seed=1, weights=(1, 64), weight_seed=2)
    log(f"graph {hg.name}: n={hg.n} m={hg.m} built in {time.time() - t0:.1f}s")
    t0 = time.time()
    g = sg.Graph(hg)
    log(f"uploaded (CSR+CSC+weights, {g.info()[2] / 1e9:.2f} GB) in {time.time() - t0:.1f}s")
    sources = sg.pick_sources(hg.row, args.sources, seed=3)
    stream = torch.cuda.current_stream()
    e = sg.Engine(g, stream=stream.cuda_stream)
    if args.kernels:
        kernel_sweep(sg, e, hg, sources, args.kernels)
        return

    # Warm-up pass: per-source reach statistics (deterministic), kept levels for
    # the CPU-baseline cross-check.
    n_r, m_r, keep = [], [], {}
    for i, s in enumerate(sources):
        e.bfs(int(s))
        st = e.stats()
        if st["status"] != 0:
            raise RuntimeError(f"BFS status {st['status']} for source {s}")
        n_r.append(st["reached"])
        m_r.append(st["reached_edges"])
        if i < 8:
            keep[i] = e.gather(np.int32)
    for _ in range(max(0, args.warmup - 1)):
        for s in sources:
            e.bfs(int(s))
    torch.cuda.synchronize()
    log(f"warm-up done; mean n_r={np.mean(n_r):.0f} m_r={np.mean(m_r):.0f}")

    nq = len(sources)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(nq * args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    k = 0
    for _ in range(args.steps):
        for s in sources:
            evs[k][0].record(stream)
            e.bfs(int(s))
            evs[k][1].record(stream)
            k += 1
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    e.sync()
    durs_ms = np.array([a.elapsed_time(b) for a, b in evs])
    total_edges = args.steps * float(np.sum(m_r))
    gteps = total_edges / wall / 1e9
    bytes_per = 8.0 * np.array(n_r, np.float64) + 4.0 * np.array(m_r, np.float64) + 4.0 * hg.n
    alg_bytes = float(np.tile(bytes_per, args.steps).sum())
    achieved = alg_bytes / (durs_ms.sum() / 1e3) / 1e9
    log(f"timed {nq * args.steps} BFS in {wall * 1e3:.1f} ms: {gteps:.1f} GTEPS, "
        f"kernel mean {durs_ms.mean() * 1e3:.1f} us, achieved {achieved:.0f} GB/s")

    # SSSP on the same graph and sources (a subset: one query is ~20x a BFS)
    ns = min(args.sssp_sources, nq)
    sssp_m, sssp_ms = 0.0, 0.0
    sssp_keep = {}
    for i, s in enumerate(sources[:ns]):
        e.sssp(int(s))
        st = e.stats()
        if st["status"] != 0:
            raise RuntimeError(f"SSSP status {st['status']} for source {s}")
        sssp_m += st["reached_edges"]
        sssp_ms += st["kernel_ms"]
        if i < 2 and not args.no_cpu:
            sssp_keep[i] = e.gather(np.uint32)  # checked against the CPU baseline below
    sssp_gteps = sssp_m / (sssp_ms / 1e3) / 1e9 if sssp_ms else 0.0
    log(f"SSSP: {ns} queries, mean {sssp_ms / max(ns, 1):.2f} ms, {sssp_gteps:.2f} GTEPS")

    serving = None
    rates = [float(r) for r in args.serve_rates.split(",")]
    if not args.no_serving:
        serving = serving_bench(sg, [g], [hg], [e], [1.0], f"rmat{args.scale}-directed", rates,
                                args.serve_count, args.serve_plain, seed=4)
    serving5 = None
    if args.config5:
        serving5 = config5(sg, args, [float(r) for r in args.config5_rates.split(",")])

    cpu = None
    cpu_sssp = None
    if not args.no_cpu:
        cpu = cpu_baseline(hg, sources, keep, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.4f} GTEPS ({cpu['sample']}), exact {cpu['bit_exact_vs_gpu']}")
        if ns:
            cpu_sssp = cpu_sssp_baseline(hg, sources[:ns], sssp_keep, args.cpu_seconds / 2)
            log(f"cpu SSSP baseline {cpu_sssp['value']:.4f} GTEPS ({cpu_sssp['sample']}), "
                f"exact {cpu_sssp['bit_exact_vs_gpu']}")

    out = {
        "metric": f"BFS GTEPS, RMAT-{args.scale} directed, {nq} sources (bit-exact levels)",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic RMAT (seeded, generated in-process)",
        "config": {
            "workload": f"bfs-rmat{args.scale}-ef16-directed-{nq}src",
            "scale": args.scale,
            "edge_factor": 16,
            "n": hg.n,
            "m": hg.m,
            "sources": nq,
            "mean_reached_vertices": float(np.mean(n_r)),
            "mean_reached_edges": float(np.mean(m_r)),
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_bfs (persistent hybrid BFS, one launch per query)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(f"k_bfs_rmat{args.scale}_traffic"),
            "mean_launch_us": round(float(durs_ms.mean()) * 1e3, 2),
        },
        "cpu_baseline": cpu,
        "sssp": {"gteps": round(sssp_gteps, 3), "queries": ns, "mean_ms": round(sssp_ms / max(ns, 1), 3),
                 "cpu_baseline": cpu_sssp},
        "serving": serving,
        "serving_config5": serving5,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
