#!/usr/bin/env python3
"""Benchmark: plain single-source hybrid BFS on one MI355X (GTEPS).

Workload (default): RMAT scale 25, edge factor 36 (2^25 vertices, 1,207,959,552
samples, a=0.57 b=0.19 c=0.19, seed 4), self-loops/duplicates removed,
degree-reordered, column-stripe tiles W=4096 S=16; the graph is generated and
laid out in HBM by libgphi before timing.  One step = one hybrid BFS from the
next seeded source (out-degree > 0), results left in HBM.

Prints one JSON line.  `value` = sum of edges traversed (sum of out-degrees of
reached vertices) over the timed BFS runs / wall time of those runs.  The
`roofline` object is the dominant kernel (the dense tile sweep): algorithmic
bytes per launch over its mean HIP-event duration in the timed region.  The
`cpu_baseline` is the oracle's OpenMP restatement of bfs_simd's hybrid BFS on
the same graph, timed on this host for a bounded number of sources; its levels
are also compared with the GPU's (full-size parity).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS hybrid BFS (1 GPU); batched BFS queries/s + p99 vs plain, exact levels"


def sweep_bytes(n: int, e_act: float, t_act: float, f_next: float) -> float:
    """Algorithmic bytes of one dense tile sweep: 8 B per active edge (head+tail),
    4 B per active tile, frontier and visited bitmaps read, parent write per
    discovered vertex."""
    return 8 * e_act + 4 * t_act + 2 * ((n + 7) // 8) + 4 * f_next


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="timed BFS runs")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c4", help="workload key (c1, c2, c4)")
    ap.add_argument("--W", type=int, default=None, help="tile width")
    ap.add_argument("--S", type=int, default=None, help="stripe length")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="budget of the CPU baseline (0 disables it)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copies": 5,
                          "ok": bool(torch.equal(src, dst))}))
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "bench.py measures one GPU; every box here has one GPU, so "
                          "--gpus N>1 is not supported"}))
        return 2

    import bfs_simd_benchmarks_amd as G
    from bfs_simd_benchmarks_amd.workloads import WORKLOADS, pick_sources

    wl = WORKLOADS[args.config]
    ctx = G.Context(0)
    t0 = time.perf_counter()
    g = wl.generate(ctx, W=args.W, S=args.S)
    gen_s = time.perf_counter() - t0
    print(f"[bench] generated {wl.name}: n={g.n} m={g.m} W={g.W} S={g.S} in {gen_s:.1f}s "
          f"({g.device_bytes / 2**30:.1f} GiB in HBM)", file=sys.stderr, flush=True)
    deg = g.download(csr=False, tiles=False)["deg"]
    srcs = pick_sources(deg, args.warmup + args.steps, wl.seed)

    for s in srcs[: args.warmup]:
        g.bfs(int(s), parents=False, levels=False)
    timed = srcs[args.warmup:]
    stats = []
    t0 = time.perf_counter()
    for s in timed:
        _, _, st = g.bfs(int(s), parents=False, levels=False)
        stats.append(st)
    wall = time.perf_counter() - t0
    edges = sum(int(st.edges_traversed) for st in stats)
    gteps = edges / wall / 1e9
    print(f"[bench] {len(timed)} BFS in {wall * 1e3:.1f} ms wall, "
          f"device {sum(st.total_ms for st in stats):.1f} ms, {gteps:.2f} GTEPS",
          file=sys.stderr, flush=True)

    # roofline of the dense sweep: bytes from the (deterministic) levels of
    # each timed source, durations from the timed runs' HIP events
    dense_ms, dense_bytes, dense_launches = 0.0, 0.0, 0
    step_bytes_total, step_ms_total = 0.0, 0.0
    levels0 = None
    for s, st in zip(timed, stats):
        lv, _, st2 = g.bfs(int(s), parents=False, levels=True)
        if levels0 is None:
            levels0 = lv
        tr = st2.trace()
        ab = g.alg_bytes(lv, st2)
        sw = st.sweep_times()
        for k in range(st.steps):
            f_next = float(tr[k + 1][1]) if k + 1 < st.steps else 0.0
            step_bytes_total += ab[k]
            step_ms_total += float(sw[k]) + float(st.update_times()[k])
            if tr[k][0] == 1:
                # e_act/t_act are folded into ab; recover the sweep share
                other = 3 * ((g.n + 7) // 8) + 8 * f_next
                if k > 0 and tr[k - 1][0] == 0:
                    other += (g.n + 7) // 8
                e_t = ab[k] - other
                dense_bytes += e_t + 2 * ((g.n + 7) // 8) + 4 * f_next
                dense_ms += float(sw[k])
                dense_launches += 1
    per_launch_bytes = dense_bytes / max(dense_launches, 1)
    per_launch_ms = dense_ms / max(dense_launches, 1)
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9 if dense_launches else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "bfs_dense_kernel", "launches": dense_launches,
                "mean_launch_ms": round(per_launch_ms, 4),
                "alg_bytes_per_launch": round(per_launch_bytes)}
    whole = step_bytes_total / (step_ms_total * 1e-3) / 1e9 if step_ms_total else 0.0

    cpu = None
    levels_match = None
    if args.cpu_seconds > 0:
        import oracle as O
        d = g.download()
        toff = g.tile_offsets()
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        runs, cpu_edges, cpu_t = 0, 0, 0.0
        for s in timed:
            lv = np.empty(g.n, np.int32)
            pa = np.empty(g.n, np.int32)
            tr = np.zeros(3 * 4096, np.uint32)
            t1 = time.perf_counter()
            O.lib().orc_bfs_hybrid(g.n, g.m, g.W, g.S, O._p(d["deg"]), O._p(d["cols"]),
                                   O._p(d["heads"]), O._p(d["tails"]), O._p(toff), int(s), 20,
                                   threads, O._p(lv), O._p(pa), O._p(tr), 4096)
            cpu_t += time.perf_counter() - t1
            if runs == 0:
                levels_match = bool(np.array_equal(lv, levels0))
            cpu_edges += int(d["deg"][lv >= 0].astype(np.uint64).sum())
            runs += 1
            if cpu_t >= args.cpu_seconds:
                break
        cpu = {"value": round(cpu_edges / cpu_t / 1e9, 4), "unit": "GTEPS", "cores": threads,
               "kind": "port",
               "sample": f"{runs} single-source hybrid BFS runs (same graph and sources as the "
                         f"GPU) with the oracle's OpenMP restatement of apps/bfs_simd, "
                         f"{cpu_t:.1f} s"}
        print(f"[bench] cpu baseline {cpu['value']} GTEPS on {threads} threads, levels match: "
              f"{levels_match}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": len(timed),
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / max(len(timed), 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"{wl.name} plain single-source hybrid BFS", "vertices": g.n,
                   "edges": g.m, "tile_width": g.W, "stripe_length": g.S, "t_ratio": 20,
                   "sources": len(timed), "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device_ms_per_bfs": round(sum(st.total_ms for st in stats) / max(len(stats), 1), 4),
        "whole_bfs_alg_GBps": round(whole, 1),
        "levels_match_cpu": levels_match,
        "graph_build_s": round(gen_s, 2),
    }
    print(json.dumps(out))
    g.free()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
