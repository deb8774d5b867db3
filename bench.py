#!/usr/bin/env python3
"""Benchmark: plain single-source hybrid BFS on one MI355X (GTEPS).

Workload (default): RMAT scale 25, edge factor 36 (2^25 vertices, 1,207,959,552
samples, a=0.57 b=0.19 c=0.19, seed 4), self-loops/duplicates removed,
degree-reordered, column-stripe tiles W=2048 S=32 (the autotuner's choice);
the graph is generated and
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
import threading
import time
from pathlib import Path

# CPU baseline placement (read by the OpenMP runtime when it loads, so before
# any import that loads it): threads on consecutive physical cores.  With a
# binding policy the runtime pins the calling thread to one core at its first
# parallel region, and threads created later inherit that mask: the serving
# legs put their thread back on the cores the process started with.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
START_CPUS = os.sched_getaffinity(0)

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS hybrid BFS (1 GPU); batched BFS queries/s + p99 vs plain, exact levels"


def sweep_bytes(n: int, e_act: float, t_act: float, f_next: float) -> float:
    """Algorithmic bytes of one dense tile sweep: 8 B per active edge (head+tail),
    4 B per active tile, frontier and visited bitmaps read, parent write per
    discovered vertex."""
    return 8 * e_act + 4 * t_act + 2 * ((n + 7) // 8) + 4 * f_next


PROFILES = Path(__file__).resolve().parent / "profiles"


def cpu_model():
    """(model name, physical cores) of the host, from lscpu."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None, None
    info = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            info[k.strip()] = v.strip()
    phys = None
    try:
        phys = int(info["Core(s) per socket"]) * int(info["Socket(s)"])
    except (KeyError, ValueError):
        pass
    return info.get("Model name"), phys


def measured_traffic(config: str, W: int, S: int):
    """HBM bytes per bfs_dense_kernel launch from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE in KiB per
    dispatch, each scaled by what the counter reads for a device copy of known
    size), or (None, None) when no summary covers it."""
    from bfs_simd_benchmarks_amd.workloads import WORKLOADS
    f = PROFILES / f"bfs_dense_kernel_{WORKLOADS[config].name}.md"
    if not f.exists():
        return None, None
    vals = {}
    for line in f.read_text().splitlines():
        if "=" in line and line.strip().startswith("traffic."):
            k, v = line.strip().split("=", 1)
            vals[k.strip()[len("traffic."):]] = v.strip()
    try:
        if (int(vals["W"]), int(vals["S"])) != (W, S):
            return None, None, None
        b = (float(vals["fetch_kib"]) * float(vals["fetch_factor"]) +
             float(vals["write_kib"]) * float(vals["write_factor"])) * 1024
    except (KeyError, ValueError):
        return None, None, None
    mean_ms = None
    try:
        mean_ms = float(vals["dense_mean_ms"])  # kernel-trace mean of the dense-step dispatches
    except (KeyError, ValueError):
        pass
    return round(b), str(f.relative_to(f.parents[1])), mean_ms


def host_link_probe(mib: int, reps: int = 20) -> int:
    """Device-to-host copy bandwidth, GB/s: pinned and pageable destinations,
    1 and 8 threads each copying its own buffer on its own stream."""
    import torch
    nbytes = mib * (1 << 20)
    res = {}
    for pinned in (True, False):
        for nthreads in (1, 8):
            srcs = [torch.full((nbytes,), i + 1, dtype=torch.uint8, device="cuda") for i in range(nthreads)]
            dsts = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=pinned) for _ in range(nthreads)]
            streams = [torch.cuda.Stream() for _ in range(nthreads)]
            torch.cuda.synchronize()

            def work(i, count):
                with torch.cuda.stream(streams[i]):
                    for _ in range(count):
                        dsts[i].copy_(srcs[i], non_blocking=pinned)
                    streams[i].synchronize()
            for i in range(nthreads):
                work(i, 1)
            t0 = time.perf_counter()
            th = [threading.Thread(target=work, args=(i, reps)) for i in range(nthreads)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            dt = time.perf_counter() - t0
            ok = all(int(dsts[i][0]) == i + 1 and int(dsts[i][-1]) == i + 1 for i in range(nthreads))
            key = f"{'pinned' if pinned else 'pageable'}_{nthreads}_threads_GBps"
            res[key] = round(nthreads * reps * nbytes / dt / 1e9, 1)
            res["ok"] = res.get("ok", True) and ok
            print(f"[bench] host link {key}: {res[key]}", file=sys.stderr, flush=True)
            del srcs, dsts
    print(json.dumps({"host_link_probe_MiB": mib, "copies_per_thread": reps, **res}))
    return 0


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
    ap.add_argument("--batches", default="64,256",
                    help="batched multi-source BFS sizes to measure ('' disables)")
    ap.add_argument("--batch-launches", type=int, default=30,
                    help="timed launches per batch size (p50/p99 over them)")
    ap.add_argument("--other-paths", type=int, default=1,
                    help="also time PageRank, hop relaxation and weighted SSSP on the same graph (0 skips)")
    ap.add_argument("--config-legs", type=int, default=1,
                    help="also run the other configurations briefly: pokec-shape BFS and PageRank, "
                         "weighted rmat16 SSSP, the serving mix at two loads (0 skips)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    ap.add_argument("--host-link-probe", type=int, default=0, metavar="MIB",
                    help="only time device-to-host copies of MIB MiB into pinned and pageable host "
                         "memory, from 1 and 8 threads with a stream each (the host link that "
                         "answers copied out of HBM cross)")
    ap.add_argument("--serving", action="store_true",
                    help="run the serving mix instead: 8 resident RMAT graphs (scales 18-24), a "
                         "Poisson stream of BFS/SSSP queries, batched vs one-at-a-time")
    ap.add_argument("--qps", default="1000,4000,16000,64000",
                    help="offered loads in queries/s, comma-separated (serving)")
    ap.add_argument("--queries", type=int, default=20_000, help="queries in each load's stream (serving)")
    args = ap.parse_args()
    if args.serving:
        return serving(args)
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

    if args.host_link_probe:
        return host_link_probe(args.host_link_probe)

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
    # the timed BFS run with per-level event timing off (their kernels back to
    # back, total device time only); the roofline pass below re-runs them with it on
    ctx.set_level_timing(False)
    t0 = time.perf_counter()
    for s in timed:
        _, _, st = g.bfs(int(s), parents=False, levels=False)
        stats.append(st)
    wall = time.perf_counter() - t0
    ctx.set_level_timing(True)
    edges = sum(int(st.edges_traversed) for st in stats)
    gteps = edges / wall / 1e9
    print(f"[bench] {len(timed)} BFS in {wall * 1e3:.1f} ms wall, "
          f"device {sum(st.total_ms for st in stats):.1f} ms, {gteps:.2f} GTEPS",
          file=sys.stderr, flush=True)

    # roofline of the dense sweep: bytes from the (deterministic) levels of
    # each timed source, durations from the HIP events of a second run of it
    dense_ms, dense_bytes, dense_launches, select_ms = 0.0, 0.0, 0, 0.0
    step_bytes_total, step_ms_total = 0.0, 0.0
    levels0 = None
    for s, st in zip(timed, stats):
        lv, _, st2 = g.bfs(int(s), parents=False, levels=True)
        if levels0 is None:
            levels0 = lv
        tr = st2.trace()
        ab = g.alg_bytes(lv, st2)
        sw, sel, up = st2.sweep_times(), st2.select_times(), st2.update_times()
        for k in range(st2.steps):
            f_next = float(tr[k + 1][1]) if k + 1 < st2.steps else 0.0
            step_bytes_total += ab[k]
            step_ms_total += float(sw[k]) + float(sel[k]) + float(up[k])
            select_ms += float(sel[k])
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
    traffic, traffic_src, prof_ms = measured_traffic(args.config, g.W, g.S)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "bfs_dense_kernel", "launches": dense_launches,
                "mean_launch_ms": round(per_launch_ms, 4),
                "alg_bytes_per_launch": round(per_launch_bytes),
                "chunk_select_ms_per_launch": round(select_ms / max(dense_launches, 1), 4)}
    if traffic_src:
        roofline["traffic_source"] = traffic_src
    if prof_ms:
        # the same bytes over the committed rocprofv3 kernel-trace mean (that
        # run's dense-step dispatches); `frac` above, from HIP events in this
        # run, is the headline
        roofline["profile_mean_launch_ms"] = prof_ms
        roofline["frac_profile"] = round(per_launch_bytes / (prof_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    whole = step_bytes_total / (step_ms_total * 1e-3) / 1e9 if step_ms_total else 0.0

    # the oracle (CPU) on the same graph: baseline timing and row checks
    host = None
    if args.cpu_seconds > 0:
        import oracle as O
        d = g.download()
        host = {"deg": d["deg"], "cols": d["cols"], "heads": d["heads"], "tails": d["tails"],
                "toff": g.tile_offsets()}
        del d
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))

    def oracle_bfs(s):
        lv = np.empty(g.n, np.int32)
        pa = np.empty(g.n, np.int32)
        tr = np.zeros(3 * 4096, np.uint32)
        O.lib().orc_bfs_hybrid(g.n, g.m, g.W, g.S, O._p(host["deg"]), O._p(host["cols"]),
                               O._p(host["heads"]), O._p(host["tails"]), O._p(host["toff"]), int(s),
                               20, threads, O._p(lv), O._p(pa), O._p(tr), 4096)
        return lv

    # batched multi-source BFS vs repeated plain runs (same GPU, same graph);
    # both rates from host wall clock: plain = timed BFS / their wall time
    plain_qps = len(timed) / wall
    plain_lat = np.array([st.total_ms for st in stats])
    batch_out = {}
    for k in [int(x) for x in args.batches.split(",") if x.strip()]:
        bsrcs = pick_sources(deg, k, wl.seed + 1000 + k)
        g.bfs_batch(bsrcs, levels=False)                      # warm-up
        lat, bst = [], None
        for _ in range(args.batch_launches):
            t1 = time.perf_counter()
            _, bst = g.bfs_batch(bsrcs, levels=False)
            lat.append(time.perf_counter() - t1)
        lat = np.array(lat)
        bwall = float(lat.mean())
        # exactness: rows against the CPU oracle (when it runs) and every
        # checked row against that source's single-source GPU levels
        blv, _ = g.bfs_batch(bsrcs, levels=True)
        rows = list(range(0, k, max(1, k // 4)))
        exact_gpu = True
        for i in rows:
            lv_i, _, _ = g.bfs(int(bsrcs[i]), parents=False, levels=True)
            exact_gpu &= bool(np.array_equal(blv[i], np.where(lv_i >= 0, lv_i, 255).astype(np.uint8)))
        exact_oracle = None
        if host is not None:
            exact_oracle = all(np.array_equal(blv[i], np.where((r := oracle_bfs(bsrcs[i])) >= 0, r, 255)
                                              .astype(np.uint8)) for i in rows[:2])
        del blv
        batch_out[str(k)] = {
            "queries_per_s": round(k / bwall, 1),
            "vs_plain": round((k / bwall) / plain_qps, 2),
            "launches": len(lat),
            "latency_ms_p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "latency_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "device_ms": round(bst.total_ms, 3), "steps": int(bst.steps),
            "gteps": round(bst.edges_traversed / bwall / 1e9, 1),
            "rows_equal_single_source_gpu": exact_gpu, "rows_checked_gpu": len(rows),
            "rows_equal_oracle": exact_oracle, "rows_checked_oracle": 2 if host is not None else 0}
        print(f"[bench] batch {k}: {k / bwall:.0f} queries/s ({bwall * 1e3:.1f} ms per launch, "
              f"p99 {batch_out[str(k)]['latency_ms_p99']} ms), {batch_out[str(k)]['vs_plain']}x plain, "
              f"exact gpu={exact_gpu} oracle={exact_oracle}", file=sys.stderr, flush=True)

    cpu = None
    levels_match = None
    if host is not None:
        # up to 10 single-source runs within the time budget, the timed sources in order
        per_run, runs_edges, cpu_t = [], 0, 0.0
        for s in timed[:10]:
            t1 = time.perf_counter()
            lv = oracle_bfs(s)
            dt = time.perf_counter() - t1
            cpu_t += dt
            if not per_run:
                levels_match = bool(np.array_equal(lv, levels0))
            e = int(host["deg"][lv >= 0].astype(np.uint64).sum())
            per_run.append(e / dt / 1e9)
            runs_edges += e
            if cpu_t >= args.cpu_seconds:
                break
        per_run = np.array(per_run)
        model, phys = cpu_model()
        cpu = {"value": round(runs_edges / cpu_t / 1e9, 4), "unit": "GTEPS", "cores": threads,
               "kind": "port",
               "mean": round(float(per_run.mean()), 4), "median": round(float(np.median(per_run)), 4),
               "best": round(float(per_run.max()), 4), "runs": len(per_run),
               "cpu_model": model, "physical_cores": phys,
               "omp": {"threads": threads, "proc_bind": os.environ.get("OMP_PROC_BIND"),
                       "places": os.environ.get("OMP_PLACES")},
               "simd": f"avx512 dense sweep" if O.lib().orc_simd_width() == 512 else "one lane at a time",
               "sample": f"{len(per_run)} single-source hybrid BFS runs (the GPU's graph and first "
                         f"timed sources) with the oracle's OpenMP restatement of apps/bfs_simd "
                         f"(dense sweep bfs.cpp:181-245 in AVX-512 when the host has it), "
                         f"{cpu_t:.1f} s; value = edges / total time, mean/median/best per run"}
        print(f"[bench] cpu baseline {cpu['value']} GTEPS on {threads} threads ({model}), "
              f"levels match: {levels_match}", file=sys.stderr, flush=True)

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
        "plain_queries_per_s": round(plain_qps, 1),
        "plain_device_ms_p50": round(float(np.percentile(plain_lat, 50)), 3),
        "plain_device_ms_p99": round(float(np.percentile(plain_lat, 99)), 3),
        "batched": batch_out,
        "graph_build_s": round(gen_s, 2),
    }
    if args.other_paths:
        out["other_paths"] = other_paths(G, ctx, g, wl, deg)
    g.free()
    ctx.close()
    if args.config_legs:
        out["configs"] = config_legs(G)
    print(json.dumps(out))
    return 0


def other_paths(G, ctx, g, wl, deg) -> dict:
    """The same tiled sweep under the reference's other apps, on the same
    graph: PageRank (20 FP32 iterations, apps/pageRank_simd) and hop
    relaxation (apps/bellman-ford_simd) on the row-major tiles those apps
    read, and weighted hybrid SSSP (apps/sssp_simd) on the graph with uint32
    weights in [1,255].  Bytes per unit follow the byte model: PageRank
    8M + 4T + 16N per iteration (T = non-empty tiles), a hop sweep 8M + 8N."""
    from bfs_simd_benchmarks_amd.workloads import pick_sources
    res = {}
    W, S = g.W, g.S
    g.retile(W, S, G.TILES_ROW_MAJOR)
    offs = g.tile_offsets().astype(np.int64)
    sizes = np.diff(np.append(offs, g.m))
    t_nonempty = int(np.count_nonzero(sizes))
    del offs, sizes
    g.pagerank(1, 0.85, 1.0 / g.n)                       # warm-up
    _, pr_ms = g.pagerank(20, 0.85, 1.0 / g.n)
    it_ms = pr_ms / 20
    pr_bytes = 8 * g.m + 4 * t_nonempty + 16 * g.n
    res["pagerank"] = {"ms_per_iteration": round(it_ms, 4), "iterations": 20,
                       "alg_GBps": round(pr_bytes / (it_ms * 1e-3) / 1e9, 1),
                       "frac_of_hbm_peak": round(pr_bytes / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "nonempty_tiles": t_nonempty}
    _, sweeps, bf_ms = g.bellman_ford(0, 99)
    bf_bytes = 8 * g.m + 8 * g.n
    res["bellman_ford"] = {"sweeps": int(sweeps), "ms": round(bf_ms, 3),
                           "alg_GBps": round(bf_bytes * sweeps / (bf_ms * 1e-3) / 1e9, 1)}
    g.retile(W, S, G.TILES_COL_STRIPE)
    print(f"[bench] pagerank {it_ms:.3f} ms/iteration, bellman-ford {sweeps} sweeps in {bf_ms:.2f} ms",
          file=sys.stderr, flush=True)
    import dataclasses
    gw = dataclasses.replace(wl, weight_seed=wl.seed + 100).generate(ctx, weighted=True)
    try:
        srcs = pick_sources(deg, 5, wl.seed + 77)
        gw.sssp(int(srcs[0]))
        ms, edges = 0.0, 0
        for s_ in srcs[1:]:
            _, st = gw.sssp(int(s_))
            ms += st.total_ms
            # the vertices SSSP reaches are the ones BFS reaches from the same source
            edges += g.bfs(int(s_), parents=False, levels=False)[2].edges_traversed
        res["sssp"] = {"ms_per_query": round(ms / (len(srcs) - 1), 3),
                       "gteps": round(edges / (ms * 1e-3) / 1e9, 1), "sources": len(srcs) - 1}
        print(f"[bench] sssp {res['sssp']['ms_per_query']} ms per query", file=sys.stderr, flush=True)
        # 16 sources in one multi-source traversal, rows checked against the
        # single-source runs of two of them
        bsrcs = pick_sources(deg, 16, wl.seed + 78)
        gw.sssp_batch(bsrcs, dist=False)                      # warm-up (state)
        t0 = time.perf_counter()
        _, bst = gw.sssp_batch(bsrcs, dist=False)
        bwall = time.perf_counter() - t0
        bd, _ = gw.sssp_batch(bsrcs)
        exact = all(np.array_equal(bd[i], gw.sssp(int(bsrcs[i]))[0]) for i in (0, 15))
        res["sssp_batch"] = {"sources": 16, "ms_per_launch": round(bwall * 1e3, 3),
                             "queries_per_s": round(16 / bwall, 1),
                             "vs_single": round(16 / bwall / (1e3 / res["sssp"]["ms_per_query"]), 2),
                             "steps": int(bst.steps), "rows_equal_single_source": bool(exact)}
        print(f"[bench] sssp batch 16: {res['sssp_batch']}", file=sys.stderr, flush=True)
        del bd
    finally:
        gw.free()
    return res


def config_legs(G) -> dict:
    """The other benchmark configurations, each briefly, on the same GPU:
    pokec shape (64 seeded sources one at a time, then 20 PageRank
    iterations), weighted rmat16 (SSSP from its 16 seeded sources) and the
    serving mix at two offered loads, 1,000 q/s and 6,000 q/s (below and above
    what the one-at-a-time server keeps up with; 6,000 queries each, batched
    vs one-at-a-time; the full load sweep is `--serving`)."""
    from bfs_simd_benchmarks_amd.workloads import WORKLOADS, pick_sources
    res = {}
    with G.Context(0) as ctx:
        g = WORKLOADS["c2"].generate(ctx)
        try:
            deg = g.download(csr=False, tiles=False)["deg"]
            srcs = pick_sources(deg, 64, WORKLOADS["c2"].seed)
            g.bfs(int(srcs[0]), parents=False, levels=False)
            ctx.set_level_timing(False)
            t0 = time.perf_counter()
            edges = sum(int(g.bfs(int(s), parents=False, levels=False)[2].edges_traversed) for s in srcs)
            wall = time.perf_counter() - t0
            ctx.set_level_timing(True)
            g.pagerank(1, 0.85, 1.0 / g.n)
            _, pr_ms = g.pagerank(20, 0.85, 1.0 / g.n)
            res["pokec_shape"] = {"vertices": g.n, "edges": g.m, "bfs_sources": len(srcs),
                                  "bfs_ms_per_query": round(wall * 1e3 / len(srcs), 3),
                                  "bfs_gteps": round(edges / wall / 1e9, 2),
                                  "pagerank_iterations": 20, "pagerank_ms_per_iteration": round(pr_ms / 20, 4)}
        finally:
            g.free()
        g = WORKLOADS["c3"].generate(ctx, weighted=True)
        try:
            deg = g.download(csr=False, tiles=False)["deg"]
            srcs = pick_sources(deg, 16, 3)
            g.sssp(int(srcs[0]))
            t0 = time.perf_counter()
            for s in srcs:
                g.sssp(int(s))
            wall = time.perf_counter() - t0
            res["rmat16_weighted_sssp"] = {"sources": len(srcs), "ms_per_query": round(wall * 1e3 / len(srcs), 3)}
        finally:
            g.free()
    print(f"[bench] configs: {res}", file=sys.stderr, flush=True)
    res["serving_mix"] = serving_loads(G, [1000.0, 6000.0], 6000)
    return res


def serving_loads(G, loads, queries) -> dict:
    """Serving mix: 8 resident graphs (RMAT scales 18,19,20,21,22,22,23,24, edge
    factor 16, seeds 10-17, every other graph with uint32 weights in [1,255]);
    at each offered load, a Poisson stream of `queries` queries (seed 5): 80%
    BFS on a uniform random graph, 20% SSSP on a uniform random weighted graph,
    uniform random source.  The same streams run against a server that batches
    the BFS queries pending on a graph (continuous batching: up to 256 sources
    per traversal, later arrivals join it between levels) and against one that
    runs every query alone (max_batch 1).  Past saturation the answered rate
    stays below the offered one and latency grows with the backlog."""
    from bfs_simd_benchmarks_amd.workloads import SERVING_MIX, serving_stream

    # the server's lane threads (and the producer below) inherit this thread's
    # CPU mask: not the one core an OpenMP region may have pinned it to
    try:
        os.sched_setaffinity(0, START_CPUS)
    except OSError:
        pass

    results = {f"{q:g}": {} for q in loads}
    sizes = []
    for mode, max_batch in (("batched", 256), ("one_at_a_time", 1)):
        srv = G.Server(max_batch=max_batch)
        sizes = []
        t0 = time.perf_counter()
        for w in SERVING_MIX:
            ctx = G.Context(0)
            g = w.generate(ctx, W=1024, S=16, weighted=bool(w.weight_seed))
            sizes.append((g.n, g.m))
            srv.add_graph(g)
        build_s = time.perf_counter() - t0
        for qps in loads:
            arrivals, is_sssp, gsel, sources = serving_stream(queries, qps, 5)
            srcs = sources([n for n, _ in sizes])
            launches0 = sum(srv.stats(i)["launches"] for i in range(len(SERVING_MIX)))
            tickets = [0] * queries
            start = time.perf_counter()

            def producer():
                for q in range(queries):
                    dt = arrivals[q] - (time.perf_counter() - start)
                    if dt > 0:
                        time.sleep(dt)
                    tickets[q] = srv.submit(int(gsel[q]), G.QUERY_SSSP if is_sssp[q] else G.QUERY_BFS,
                                            int(srcs[q]))
            th = threading.Thread(target=producer)
            th.start()
            lat, batch, joined = np.empty(queries), np.empty(queries), np.empty(queries)
            for q in range(queries):
                while tickets[q] == 0:          # not submitted yet
                    time.sleep(0.0005)
                info = srv.wait(tickets[q])
                lat[q], batch[q], joined[q] = info.latency_us, info.batch, info.joined_level
                if (q + 1) % 10000 == 0:
                    print(f"[serving] {mode} @ {qps:g} q/s: {q + 1} answered after "
                          f"{time.perf_counter() - start:.1f} s", file=sys.stderr, flush=True)
            th.join()
            wall = time.perf_counter() - start
            bfs = ~is_sssp
            launches = sum(srv.stats(i)["launches"] for i in range(len(SERVING_MIX))) - launches0
            r = {"answered_qps": round(queries / wall, 1),
                 "latency_ms_p50": round(float(np.percentile(lat, 50)) / 1e3, 3),
                 "latency_ms_p99": round(float(np.percentile(lat, 99)) / 1e3, 3),
                 "bfs_latency_ms_p50": round(float(np.percentile(lat[bfs], 50)) / 1e3, 3),
                 "bfs_latency_ms_p99": round(float(np.percentile(lat[bfs], 99)) / 1e3, 3),
                 "sssp_latency_ms_p99": round(float(np.percentile(lat[~bfs], 99)) / 1e3, 3),
                 "mean_bfs_batch": round(float(batch[bfs].mean()), 2),
                 "bfs_joined_mid_traversal": round(float((joined[bfs] > 0).mean()), 3),
                 "launches": int(launches), "graph_build_s": round(build_s, 2)}
            results[f"{qps:g}"][mode] = r
            print(f"[serving] {mode} @ {qps:g} q/s: {r}", file=sys.stderr, flush=True)
        srv.close()
    return {"queries_per_load": queries, "bfs_share": 0.8,
            "graphs": [f"{w.name} n={n} m={m}" + (" weighted" if w.weight_seed else "")
                       for w, (n, m) in zip(SERVING_MIX, sizes)],
            "loads": results}


def sssp_group_costs(G) -> dict:
    """What sharing costs on the serving mix's weighted graphs: wall ms of one
    SSSP alone (mean and median over 24 uniform random sources, seed 3) and of
    one multi-source SSSP traversal of k = 2, 4, 8, 16 of those sources (mean
    over 4 groups)."""
    from bfs_simd_benchmarks_amd.workloads import SERVING_MIX
    rng = np.random.default_rng(3)
    out = {}
    for w in SERVING_MIX:
        if not w.weight_seed:
            continue
        ctx = G.Context(0)
        g = w.generate(ctx, W=1024, S=16, weighted=True)
        try:
            srcs = rng.integers(0, g.n, 64)
            g.sssp(int(srcs[0]))
            g.sssp_batch(srcs[:16], dist=False)           # warm-up (state)
            alone = []
            for s_ in srcs[:24]:
                t0 = time.perf_counter()
                g.sssp(int(s_))
                alone.append((time.perf_counter() - t0) * 1e3)
            r = {"alone_mean": round(float(np.mean(alone)), 3), "alone_median": round(float(np.median(alone)), 3)}
            for k in (2, 4, 8, 16):
                tk = []
                for i in range(4):
                    t0 = time.perf_counter()
                    g.sssp_batch(srcs[16 * i:16 * i + k], dist=False)
                    tk.append((time.perf_counter() - t0) * 1e3)
                r[f"group_{k}"] = round(float(np.mean(tk)), 3)
            out[w.name] = r
            print(f"[serving] sssp {w.name}: {r}", file=sys.stderr, flush=True)
        finally:
            g.free()
    return out


def serving(args) -> int:
    """The serving mix over the offered loads of --qps (serving_loads), after
    the SSSP group costs of its weighted graphs (sssp_group_costs)."""
    import bfs_simd_benchmarks_amd as G
    loads = [float(x) for x in str(args.qps).split(",") if x]
    costs = sssp_group_costs(G)
    out = {"metric": "serving mix: answered queries/s and p50/p99 latency per offered load, "
                     "batched vs one-at-a-time",
           "sssp_ms_alone_and_per_group": costs,
           **serving_loads(G, loads, args.queries)}
    print(json.dumps(out))
    return 0

if __name__ == "__main__":
    sys.exit(main())
