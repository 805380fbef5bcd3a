#!/usr/bin/env python3
"""Benchmark of the EverythingGraph data-parallel path on one MI355X.

Workload (one step): a full PageRank, 20 iterations, FP32 iterates, r0 = 1/n,
d = 0.85, over an undirected R-MAT scale-24 graph (edge factor 16, a,b,c =
0.57,0.19,0.19, seed 1; 2^24 vertices, 536,870,912 stored entries) on the CSR
of the symmetric graph, inputs resident in HBM, with the kernels the
autotuner picks (the binned two-phase kernels with exact fixed-point row
sums, or the pull kernel with compensated FP32 sums).
`value` is algorithmic GB/s = iterations * B_PR / step time with
B_PR = 4m + 4(n+1) + 16n bytes per iteration.

Beside it, on the same resident graph: direction-optimising BFS GTEPS
(harmonic mean over 64 seeded roots), the batched forms (256 BFS sources as
multi-source sweeps; 64 personalized PageRank vectors as one SpMM, 30
iterations) against the plain per-query path, the live roofline of the
dominant kernel, and the CPU oracle timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PR/SpMV GB/s + BFS GTEPS @RMAT-24; batched queries/s & p99 vs plain path"
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_calibration(nbytes: int) -> float:
    """One device-to-device copy of nbytes through the HIP runtime (a plain
    copy kernel of known size, the reference point for the FETCH_SIZE /
    WRITE_SIZE counters).  Returns its wall time in ms."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    t0 = time.perf_counter()
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    ms = 1e3 * (time.perf_counter() - t0)
    hip.hipFree(a)
    hip.hipFree(b)
    return ms


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the rocprofv3
    FETCH_SIZE / WRITE_SIZE figures kept in profiles/traffic_rmat24.txt
    (KiB per dispatch and their copy-calibrated factors), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_rmat24.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 6 and p[0] == kernel and p[1] == workload:
                return int((float(p[2]) * float(p[4]) + float(p[3]) * float(p[5])) * 1024)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--bfs-roots", type=int, default=64)
    ap.add_argument("--batch-bfs", type=int, default=256)
    ap.add_argument("--batch-ppr", type=int, default=64)
    ap.add_argument("--cpu-iterations", type=int, default=2,
                    help="least number of CPU oracle iterations")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU work the oracle sample is sized to")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--serve-qps", type=float, action="append", default=None,
                    help="offered Poisson rate(s) of the serving replay (default 50, 100, "
                         "200 and a burst)")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-sustained-qps", type=float, default=300.0,
                    help="offered rate of the sustained-overload replay (0: skip)")
    ap.add_argument("--serve-sustained-requests", type=int, default=1500)
    ap.add_argument("--serve-ppr-grid", type=int, default=0,
                    help="SpMM workgroups while serving (0: library default)")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-configs", action="store_true",
                    help="skip the grid (ER 2^24 / 2^28, 16x16) and RMAT-22 BFS/WCC figures")
    ap.add_argument("--skip-c5", action="store_true",
                    help="skip the R-MAT 26 serving replay")
    ap.add_argument("--c5-requests", type=int, default=150,
                    help="requests per R-MAT 26 serving replay")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in kernel shapes (plus any --tune)")
    ap.add_argument("--tune", action="append", default=[], metavar="PARAM=VALUE",
                    help="force a kernel shape, e.g. pull_hot=262144")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import everythinggraph_amd as eg

    dev = eg.device_info()
    log(f"device: {dev.name.decode()} CUs={dev.compute_units} gfx{dev.gcn_arch} "
        f"HBM={dev.hbm_bytes / 2**30:.0f} GiB")
    if args.copy_calib:
        ms = copy_calibration(args.copy_calib << 20)
        log(f"copy calibration: {args.copy_calib} MiB in {ms:.3f} ms")
    scale, ef = args.scale, 16
    n = 1 << scale
    t0 = time.perf_counter()
    g = eg.Graph(n, eg.MODE_ADJ_SORT, undirected=True, symmetric=True)
    g.generate(eg.rmat_params(scale, ef, seed=1))
    info = g.info
    m = info.m
    log(f"graph: R-MAT {scale} EF{ef} -u: n={n} m_in={info.m_in} m={m} "
        f"build {info.build_ms:.1f} ms device (wall {time.perf_counter() - t0:.1f} s)")
    csr_bytes = 2 * 8 * info.m_in + 4 * m + 4 * (n + 1) + 8 * n

    # ---- kernel shapes: autotuned on this graph before anything is timed ----
    t0 = time.perf_counter()
    if not args.no_autotune:
        tune = g.autotune()
        log("autotune: " + ", ".join(f"{p}={v} {ms:.3f} ms{'*' if c else ''}"
                                     for p, v, ms, c in tune)
            + f" ({time.perf_counter() - t0:.1f} s)")
    for kv in args.tune:
        p, v = kv.split("=")
        g.tune_set(p, int(v))
    tuned = {p: g.tune_get(p) for p in ("pr_binned", "pr_xcd", "pr_scatter", "pr_hot", "pr_persist",
                                        "pull_batch", "pull_hot", "pull_persist", "spmm_depth",
                                        "ms_batch", "ms_push_div")}

    # ---- PageRank: warmup + timed steps ----
    iters = args.iterations
    b_pr = 4 * m + 4 * (n + 1) + 16 * n
    for _ in range(args.warmup):
        g.pagerank(0.85, iters, download=False)
    g.sync()
    g.prof_reset()
    g.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r, dh, lh = g.pagerank(0.85, iters, download=False)
    g.sync()
    t_steps = time.perf_counter() - t0
    prof = g.prof_read()
    g.prof_enable(False)
    ms_per_step = 1e3 * t_steps / args.steps
    value = iters * b_pr / (ms_per_step * 1e-3) / 1e9

    def avg_ms(name):
        launches, tot = prof.get(name, (0, 0.0))
        return tot / launches if launches else 0.0

    # The dominant work of one iteration: the binned kernels (scatter +
    # accumulate + finish, three launches) or the pull kernel (one launch).
    binned = "pr_scatter" in prof
    if binned:
        parts = {k: avg_ms(k) for k in ("pr_scatter", "pr_accum", "pr_finish")}
        kernel = "k_pb_scatter + k_pb_accum + k_pb_finish (one binned iteration)"
        tname = ("pr_scatter", "pr_accum")
    else:
        parts = {"pr_pull": avg_ms("pr_pull")}
        kernel = "k_pull<true,false> (pr_pull)"
        tname = ("pr_pull",)
    k_ms = sum(parts.values())
    achieved = b_pr / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    traffic = [measured_traffic(t, f"rmat{scale}_ef{ef}_u_s") for t in tname]
    traffic = None if any(t is None for t in traffic) else sum(traffic)
    log(f"pagerank: {ms_per_step:.3f} ms/step ({iters} it), {value:.1f} GB/s; iteration "
        + " + ".join(f"{k} {v * 1e3:.1f} us" for k, v in parts.items())
        + f" = {achieved:.1f} GB/s; final L1 delta {lh[-1]:.3e}")

    # ---- BFS over 64 seeded roots ----
    deg, _ = g.degrees()
    roots = eg.sample_roots(deg, args.bfs_roots, seed=2)
    g.bfs(int(roots[0]), download=False)
    teps, walls, levels = [], [], []
    for rt in roots:
        w0 = time.perf_counter()
        _, _, st = g.bfs(int(rt), download=False, parents=False)
        walls.append(time.perf_counter() - w0)
        teps.append(st.edges_traversed / (st.device_ms * 1e-3))
        levels.append(st.levels)
    gteps = len(teps) / sum(1.0 / t for t in teps) / 1e9
    plain_bfs_qps = len(walls) / sum(walls)
    log(f"bfs: harmonic-mean {gteps:.2f} GTEPS, plain {plain_bfs_qps:.1f} queries/s, "
        f"levels {min(levels)}..{max(levels)}")

    extra = {}
    # SpMV over the same graph (unweighted, w = 1): y = A^T x, signed x
    x = np.random.default_rng(5).uniform(-1, 1, n).astype(np.float32)
    y_gpu = g.spmv(x)
    g.prof_reset()
    g.prof_enable(True)
    for _ in range(5):
        g.spmv(x)
    sp_prof = g.prof_read()
    g.prof_enable(False)
    sp_ms = sum(t for k, (_, t) in sp_prof.items() if k.startswith("spmv_")) / 5
    b_spmv1 = 4 * m + 4 * (n + 1) + 8 * n
    extra["spmv_rmat24"] = {"ms": round(sp_ms, 4), "gbs": round(b_spmv1 / (sp_ms * 1e-3) / 1e9, 1),
                            "bytes": b_spmv1, "weights": "unit"}
    log(f"spmv: {sp_ms * 1e3:.1f} us = {b_spmv1 / (sp_ms * 1e-3) / 1e9:.1f} GB/s of "
        f"4m + 4(n+1) + 8n")
    if not args.skip_batch:
        broots = eg.sample_roots(deg, args.batch_bfs, seed=2)
        g.bfs_batch(broots, download=False)  # state of this batch width allocated untimed
        w0 = time.perf_counter()
        _, ms = g.bfs_batch(broots, download=False)
        wb = time.perf_counter() - w0
        batched_bfs_qps = len(broots) / wb
        bb_info = g.bfs_batch_info()
        rng = np.random.default_rng(3)
        seeds = [eg.sample_roots(deg, int(rng.integers(1, 9)), seed=1000 + q)
                 for q in range(args.batch_ppr)]
        w0 = time.perf_counter()
        _, pms = g.ppr_batch(seeds, 0.15, 30, download=False)
        wp = time.perf_counter() - w0
        batched_ppr_qps = len(seeds) / wp
        # plain per-query PPR: 30 iterations of the single-vector pull kernel
        g.ppr(seeds[0], 0.15, tol=0.0, max_iter=2, download=False)
        w0 = time.perf_counter()
        for sd in seeds[:4]:
            g.ppr(sd, 0.15, tol=0.0, max_iter=30, download=False)
        plain_ppr_qps = 4 / (time.perf_counter() - w0)
        extra.update({
            "batched_bfs_qps": round(batched_bfs_qps, 2),
            "plain_bfs_qps": round(plain_bfs_qps, 2),
            "batched_bfs_speedup": round(batched_bfs_qps / plain_bfs_qps, 2),
            "batched_bfs_info": {k: (round(v, 3) if isinstance(v, float) else v)
                                 for k, v in bb_info.items()},
            "batched_ppr_qps": round(batched_ppr_qps, 2),
            "plain_ppr_qps": round(plain_ppr_qps, 2),
            "batched_ppr_speedup": round(batched_ppr_qps / plain_ppr_qps, 2),
        })
        log(f"batched: MS-BFS {batched_bfs_qps:.1f} q/s (plain {plain_bfs_qps:.1f}; "
            f"{bb_info['traversals']} traversal(s) of {bb_info['lanes']} lanes, "
            f"{bb_info['levels']} levels, {bb_info['host_syncs']} host waits); "
            f"PPR SpMM {batched_ppr_qps:.1f} q/s (plain {plain_ppr_qps:.1f})")

    if not args.skip_serve:
        # serving: the same Poisson stream (80% BFS, 20% PPR with 1-8 seeds,
        # tolerance 1e-6 L1, seed 4) through continuous batching and through
        # the plain one-query-at-a-time path
        # offered rates below and past the plain path's saturation (about 190
        # queries/s at R-MAT 24), then a burst: every request at once
        rates = args.serve_qps or [50.0, 100.0, 200.0, SERVE_BURST_QPS]
        serving = []
        for qps in rates:
            srv = {}
            for mode in ("batched", "plain"):
                w0 = time.perf_counter()
                r = g.serve_replay(qps, args.serve_requests, bfs_fraction=0.8,
                                   ppr_max_seeds=8, seed=4, plain=(mode == "plain"),
                                   ppr_grid=args.serve_ppr_grid)
                srv[mode] = r
                log(f"serve {mode}: offered {qps:.0f} q/s, {r['completed']} done "
                    f"({r['bfs']} BFS / {r['ppr']} PPR) at {r['qps']:.1f} q/s; p50 "
                    f"{r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms (BFS p50 "
                    f"{r['bfs_p50_ms']:.1f} p99 {r['bfs_p99_ms']:.1f}; PPR p50 "
                    f"{r['ppr_p50_ms']:.1f} p99 {r['ppr_p99_ms']:.1f}); levels "
                    f"{r['stats']['bfs_levels']} PPR steps {r['stats']['ppr_steps']}; "
                    f"wall {time.perf_counter() - w0:.1f} s")
            b, p = srv["batched"], srv["plain"]
            serving.append({
                "offered_qps": qps, "requests": args.serve_requests,
                "mix": "80% BFS / 20% PPR (1-8 seeds, tol 1e-6 L1), Poisson, seed 4",
                **{f"{k}_{f}": round(v[f], 3) for k, v in srv.items()
                   for f in ("qps", "p50_ms", "p99_ms", "bfs_p50_ms", "bfs_p99_ms",
                             "ppr_p50_ms", "ppr_p99_ms")},
                "batched_qps_vs_plain": round(b["qps"] / max(p["qps"], 1e-9), 2),
                "plain_p99_over_batched_p99": round(p["p99_ms"] / max(b["p99_ms"], 1e-9), 2),
            })
        extra["serving"] = serving
        burst = [x for x in serving if x["offered_qps"] >= SERVE_BURST_QPS]
        if burst:
            b = burst[-1]
            extra["serving_saturation"] = {
                "offered": "burst (every request submitted at once)",
                "requests": args.serve_requests,
                "engine_qps": b["batched_qps"], "plain_qps": b["plain_qps"],
                "engine_over_plain": round(b["batched_qps"] / max(b["plain_qps"], 1e-9), 2),
            }
            log(f"serving saturation: engine {b['batched_qps']:.1f} q/s, plain "
                f"{b['plain_qps']:.1f} q/s (burst of {args.serve_requests})")
        # sustained overload: an offered rate past the plain path's saturation
        # and below the engine's, for long enough that a backlog shows
        if args.serve_sustained_qps > 0:
            qps, nreq = args.serve_sustained_qps, args.serve_sustained_requests
            row = {"offered_qps": qps, "requests": nreq}
            for mode in ("batched", "plain"):
                r = g.serve_replay(qps, nreq, bfs_fraction=0.8, ppr_max_seeds=8, seed=5,
                                   plain=(mode == "plain"), ppr_grid=args.serve_ppr_grid)
                row[mode] = {f: round(r[f], 3) for f in
                             ("qps", "p50_ms", "p99_ms", "bfs_p99_ms", "ppr_p99_ms")}
                log(f"serve {mode}: sustained {qps:.0f} q/s offered, {r['completed']} done at "
                    f"{r['qps']:.1f} q/s; p50 {r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms")
            extra["serving_sustained"] = row

    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import eg_oracle_py as O

        ptr, col, _ = g.csr()
        odeg = np.diff(ptr).astype(np.uint32)
        threads = O.L().ego_num_threads()
        # a bounded sample of about --cpu-seconds of CPU work: a short probe
        # sizes the iteration count, then that many iterations are timed
        c0 = time.perf_counter()
        O.pagerank_pull(n, ptr, col, odeg, 0.85, 2)
        probe = (time.perf_counter() - c0) / 2
        cpu_iters = max(args.cpu_iterations, min(400, int(args.cpu_seconds / max(probe, 1e-3))))
        c0 = time.perf_counter()
        O.pagerank_pull(n, ptr, col, odeg, 0.85, cpu_iters)
        ct = time.perf_counter() - c0
        cpu_gbs = cpu_iters * b_pr / ct / 1e9
        cpu = {"value": round(cpu_gbs, 3), "unit": "GB/s", "cores": int(threads), "kind": "port",
               "sample": f"{cpu_iters} FP64 pull-PageRank iterations of the oracle "
                         f"(OpenMP) over the same R-MAT {scale} graph, {ct:.1f} s"}
        log(f"cpu oracle: {cpu_gbs:.2f} GB/s on {threads} threads ({ct:.1f} s)")
        # error of the device results against the FP64 oracle on this graph:
        # the timed PageRank (same iteration count, r0 = 1/n) and the SpMV
        r_gpu, _, _ = g.pagerank(0.85, iters)
        r_ref, _, _ = O.pagerank_pull(n, ptr, col, odeg, 0.85, iters)
        d = np.abs(r_gpu.astype(np.float64) - r_ref)
        y_ref = O.spmv_pull(n, ptr, col, None, x)
        y_abs = O.spmv_pull(n, ptr, col, None, np.abs(x))
        e = np.abs(y_gpu.astype(np.float64) - y_ref)
        extra["error_vs_fp64"] = {
            "pagerank": {"l1": float(d.sum()), "max_abs": float(d.max()),
                         "iterations": iters},
            "spmv": {"l1": float(e.sum()), "max_abs": float(e.max()),
                     "max_rel_to_abs_sum": float((e / np.maximum(y_abs, 1e-30)).max())},
        }
        log(f"error vs FP64 oracle: pagerank L1 {d.sum():.3e} max {d.max():.3e}; "
            f"spmv L1 {e.sum():.3e} max {e.max():.3e}")
        del ptr, col, r_gpu, r_ref, d, y_ref, y_abs, e
    del x, y_gpu

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"pagerank_pull_rmat{scale}_ef{ef}_undirected_{iters}it",
            "scale": scale, "edge_factor": ef, "n": n, "m": m, "iterations": iters,
            "layout": "csr (load mode 2, -u -s)", "parallelism": "single-gpu",
            "tuned": tuned,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic, "kernel": kernel, "bytes_per_launch": b_pr,
            "avg_launch_us": round(k_ms * 1e3, 2),
            "per_kernel_us": {k: round(v * 1e3, 2) for k, v in parts.items()},
        },
        "cpu_baseline": cpu,
        "bfs_gteps": round(gteps, 3),
        "bfs_roots": len(roots),
        "csr_build_ms": round(info.build_ms, 3),
        "csr_build_gbs": round(csr_bytes / (info.build_ms * 1e-3) / 1e9, 1),
    }
    out.update(extra)
    # one-time preparations of this graph (built on first use): device build
    # of the layout, the degree-ordered copy, the binned plan, the BFS rows
    prep = g.prep
    out["prep"] = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in prep.items()}
    log("prep: layout {layout_ms:.1f} ms / {layout_bytes} B, ordered copy {order_ms:.1f} ms / "
        "{order_bytes} B, binned plan {binned_plan_ms:.1f} ms / {binned_plan_bytes} B, BFS rows "
        "{bfs_rows_ms:.1f} ms / {bfs_rows_bytes} B".format(**prep))
    g.close()
    if not args.skip_configs:
        out["configs"] = other_configs(eg, args.skip_c5, args.c5_requests, not args.skip_cpu)
    print(json.dumps(out), flush=True)


SERVE_BURST_QPS = 1e5  # offered rate of the burst replay: every request at once


def other_configs(eg, skip_c5=False, c5_requests=150, fp64=True) -> dict:
    """Device figures of the other configurations on fresh graphs: PageRank /
    SpMV on the 2-D grid (directed Erdos-Renyi, 2^24 vertices, 2^28 weighted
    edges, 16 x 16 cells), file ingest, BFS + WCC + SSSP on undirected R-MAT
    22 (CSR) and serving on directed R-MAT 26."""
    res = {}
    n, m = 1 << 24, 1 << 28
    g = eg.Graph(n, eg.MODE_GRID_SORT, weighted=True, P=16)
    g.generate(eg.er_params(m, seed=1, weight_seed=1))
    build_ms = g.info.build_ms
    b_pr = 4 * m + 4 * (n + 1) + 16 * n
    b_spmv = 8 * m + 4 * (n + 1) + 8 * n
    tune = g.autotune(eg.TUNE_GRID)
    band = g.tune_get("grid_band")
    binned = g.tune_get("pr_binned")
    log("grid autotune: " + ", ".join(f"{p}={v} {ms:.3f} ms{'*' if c else ''}"
                                      for p, v, ms, c in tune))
    g.pagerank(0.85, 2, layout="grid", download=False)
    x = np.random.default_rng(1).uniform(-1, 1, n).astype(np.float32)
    g.spmv(x, layout="grid")
    g.prof_reset()
    g.prof_enable(True)
    _, _, lh = g.pagerank(0.85, 20, layout="grid", download=False)
    pr_prof = g.prof_read()
    g.prof_reset()
    for _ in range(5):
        g.spmv(x, layout="grid")
    sp_prof = g.prof_read()
    g.prof_enable(False)
    # device time of the PageRank iterations (grid kernel, or the binned
    # scatter + accumulate + finish) and of the SpMV passes
    pr_ms = sum(t for k, (_, t) in pr_prof.items() if k.startswith("pr_")) / 20
    sp_ms = sum(t for k, (_, t) in sp_prof.items() if k.startswith("spmv_")) / 5
    res["grid_er24"] = {
        "n": n, "m": m, "P": 16, "weighted": True, "grid_band": band,
        "kernels": "binned" if binned == 1 else "grid pull", "build_ms": round(build_ms, 2),
        "pr_iter_ms": round(pr_ms, 4), "pr_gbs": round(b_pr / (pr_ms * 1e-3) / 1e9, 1),
        "spmv_ms": round(sp_ms, 4), "spmv_gbs": round(b_spmv / (sp_ms * 1e-3) / 1e9, 1),
        "pr_final_l1_delta": float(lh[-1]),
    }
    log(f"grid ER 2^24/2^28 16x16 ({res['grid_er24']['kernels']}): build {build_ms:.1f} ms; "
        f"PR {pr_ms * 1e3:.1f} us/iter = {b_pr / (pr_ms * 1e-3) / 1e9:.1f} GB/s; "
        f"SpMV {sp_ms * 1e3:.1f} us = {b_spmv / (sp_ms * 1e-3) / 1e9:.1f} GB/s")
    if fp64:
        # error against the FP64 oracle over the same grid: 20 PageRank
        # iterations (r0 = 1/n) and the SpMV of this x
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import eg_oracle_py as O

        cp, gs, gd, gw = g.grid()
        odeg, _ = g.degrees()
        r, _, _ = g.pagerank(0.85, 20, layout="grid")
        d = np.abs(r.astype(np.float64) - O.pagerank_grid(n, 16, cp, gs, gd, odeg, 0.85, 20))
        y = g.spmv(x, layout="grid")
        e = np.abs(y.astype(np.float64) - O.spmv_grid(n, 16, cp, gs, gd, gw, x))
        yabs = O.spmv_grid(n, 16, cp, gs, gd, gw, np.abs(x))
        res["grid_er24"]["error_vs_fp64"] = {
            "pagerank": {"l1": float(d.sum()), "max_abs": float(d.max()), "iterations": 20},
            "spmv": {"l1": float(e.sum()), "max_abs": float(e.max()),
                     "max_rel_to_abs_sum": float((e / np.maximum(yabs, 1e-30)).max())},
        }
        log(f"grid error vs FP64 oracle: pagerank L1 {d.sum():.3e} max {d.max():.3e}; "
            f"spmv max {e.max():.3e} ({(e / np.maximum(yabs, 1e-30)).max():.2e} of sum |a x|)")
        del cp, gs, gd, gw, r, d, y, e, yabs
    g.close()

    # file ingest: the R-MAT 24 records (2.15 GB) read back from a file in
    # the page cache through the pinned two-slot pipeline, CSR build included
    import tempfile
    n = 1 << 24
    src, dst, _ = eg.generate_edges(eg.rmat_params(24, 16, seed=1), n)
    rec = np.empty((src.size, 2), np.uint32)
    rec[:, 0], rec[:, 1] = src, dst
    nbytes = rec.nbytes
    del src, dst
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "rmat24.bin")
        rec.tofile(path)
        del rec
        g = eg.Graph(n, eg.MODE_ADJ_SORT, undirected=True, symmetric=True)
        g.load_file(path)  # page cache warm, device allocations settled
        w0 = time.perf_counter()
        g.load_file(path)
        t = time.perf_counter() - w0
        build_ms = g.info.build_ms
        g.close()
    res["file_ingest_rmat24"] = {
        "bytes": nbytes, "load_s": round(t, 4), "gbs": round(nbytes / t / 1e9, 2),
        "csr_build_ms": round(build_ms, 2),
    }
    log(f"file ingest R-MAT 24 ({nbytes / 1e9:.2f} GB): load_file {t * 1e3:.1f} ms = "
        f"{nbytes / t / 1e9:.2f} GB/s including the {build_ms:.1f} ms CSR build")

    n = 1 << 22
    g = eg.Graph(n, eg.MODE_ADJ_SORT, undirected=True)
    g.generate(eg.rmat_params(22, 16, seed=1))
    build_ms = g.info.build_ms
    deg, _ = g.degrees()
    roots = eg.sample_roots(deg, 64, seed=2)
    g.bfs(int(roots[0]), download=False)
    teps = []
    for rt in roots:
        _, _, st = g.bfs(int(rt), download=False, parents=False)
        teps.append(st.edges_traversed / (st.device_ms * 1e-3))
    gteps = len(teps) / sum(1.0 / t for t in teps) / 1e9
    g.wcc()
    g.prof_reset()
    g.prof_enable(True)
    w0 = time.perf_counter()
    _, rounds = g.wcc()
    wcc_wall = time.perf_counter() - w0
    prof = g.prof_read()
    g.prof_enable(False)
    wl, wt = prof.get("wcc_relax", (0, 0.0))
    g.sssp(int(roots[0]), download=False)
    sssp_ms, sssp_rounds = [], []
    for rt in roots[:8]:
        _, _, st = g.sssp(int(rt), download=False, parents=False)
        sssp_ms.append(st.device_ms)
        sssp_rounds.append(st.iterations)
    res["rmat22_u"] = {
        "n": n, "m": g.info.m, "csr_build_ms": round(build_ms, 2), "bfs_gteps": round(gteps, 2),
        "wcc_rounds": rounds, "wcc_relax_ms": round(wt, 3), "wcc_wall_ms": round(wcc_wall * 1e3, 2),
        "sssp_ms": round(float(np.mean(sssp_ms)), 3), "sssp_rounds": float(np.mean(sssp_rounds)),
        "sssp_weights": "(u % 10) + (v % 10)",
    }
    log(f"RMAT-22 -u SSSP (reference weights): {np.mean(sssp_ms):.2f} ms, "
        f"{np.mean(sssp_rounds):.1f} rounds (mean of 8 roots)")
    log(f"RMAT-22 -u: CSR+CSC build {build_ms:.1f} ms; BFS {gteps:.1f} GTEPS; WCC {rounds} rounds, "
        f"relax {wt:.2f} ms, call {wcc_wall * 1e3:.1f} ms (incl. label download)")
    g.close()

    if not skip_c5:
        # serving at R-MAT 26 (directed, CSR + CSC resident): the Poisson
        # stream (80% BFS, 20% PPR with 1-8 seeds, tolerance 1e-6 L1) through
        # continuous batching and through the plain one-query-at-a-time path
        n = 1 << 26
        g = eg.Graph(n, eg.MODE_ADJ_SORT)
        g.generate(eg.rmat_params(26, 16, seed=1))
        t0 = time.perf_counter()
        tune = g.autotune()  # the shapes both modes run with
        log("RMAT-26 autotune: " + ", ".join(f"{p}={v} {ms:.3f} ms{'*' if c else ''}"
                                             for p, v, ms, c in tune)
            + f" ({time.perf_counter() - t0:.1f} s)")
        srv = {"n": n, "m": g.info.m, "device_gb": round(g.info.device_bytes / 1e9, 2),
               "requests": c5_requests,
               "tuned": {p: g.tune_get(p) for p in ("pr_binned", "pr_xcd", "spmm_depth",
                                                    "ms_batch")}}
        for qps in (10.0, 20.0):
            row = {}
            for mode in ("batched", "plain"):
                r = g.serve_replay(qps, c5_requests, bfs_fraction=0.8, ppr_max_seeds=8, seed=4,
                                   plain=(mode == "plain"))
                row[mode] = {k: round(r[k], 2) for k in
                             ("qps", "p50_ms", "p99_ms", "bfs_p50_ms", "bfs_p99_ms",
                              "ppr_p50_ms", "ppr_p99_ms")}
                row[mode]["completed"] = r["completed"]
                log(f"RMAT-26 serve {mode}: offered {qps:.0f} q/s, {r['completed']} done at "
                    f"{r['qps']:.1f} q/s; p50 {r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms "
                    f"(BFS p99 {r['bfs_p99_ms']:.1f}; PPR p50 {r['ppr_p50_ms']:.1f} "
                    f"p99 {r['ppr_p99_ms']:.1f})")
            srv[f"qps{int(qps)}"] = row
        res["serve_rmat26"] = srv
        g.close()
    return res


if __name__ == "__main__":
    main()
