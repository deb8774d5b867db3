"""(W, S) autotuner for the tiled sweep, and the binary cache around it.

The reference picks the tile width W and stripe length S by hand: the
candidate range of W comes from tools/tile-size-computer/tile-size-computer.cpp
(restated in ``tile_size_bounds``), and tools/tile_size/process.sh:19-21 and
tools/stripe_length/process.sh:19-21 re-run the app once per value, each run
re-reading the text files made for that value.  Here the graph stays resident:
each candidate is laid out on the device from the CSR (gphi_graph_retile),
timed over the same sources, and the fastest layout is kept (and written to a
binary cache file with Graph.save when asked).

Usage::

    python -m bfs_simd_benchmarks_amd.autotune --workload c4 --sources 8 \\
        --save rmat25.gphi
"""
from __future__ import annotations

import argparse
import json
import sys
from dataclasses import dataclass, field

import numpy as np

U32_MAX = 4294967295


def _side(n: int, w: int) -> int:
    return n // w + (1 if n % w else 0)


def tile_size_bounds(n: int, m: int) -> tuple[int, int]:
    """(min_tile_size, max_tile_size) of tools/tile-size-computer/tile-size-computer.cpp:21-56.

    min: from 64, doubled until SIDE^2 <= 2^32-1, then doubled until
    floor(M / SIDE^2) >= 8 (the average edges per tile; M is an unsigned int
    there).  max: doubled from min while SIDE > 64, then divided by 4, i.e.
    the largest power-of-two multiple of min with SIDE > 64 (or min/2 when
    min already has SIDE <= 64, as the tool's loop gives)."""
    if n <= 0:
        raise ValueError("graph has no vertices")
    m &= U32_MAX
    if m < 8:
        raise ValueError("fewer than 8 edges: the tool's average-edge loop never ends")
    t = 64
    while True:                                   # :26-31
        tiles = _side(n, t) ** 2
        t *= 2
        if tiles <= U32_MAX:
            break
    t //= 2
    while True:                                   # :33-38
        avg = m // (_side(n, t) ** 2)
        t *= 2
        if avg >= 8:
            break
    t //= 2
    lo = t
    while True:                                   # :41-50
        side = _side(n, t)
        t *= 2
        if side <= 64:
            break
    hi = t // 4
    return lo, hi


def candidate_widths(n: int, m: int, below: int = 3, above: int = 2) -> list[int]:
    """Powers of two around the tool's lower bound: W_min/2^below .. W_min*2^above,
    kept to SIDE^2 <= 2^32 (the tile key is 32 bits) and W <= max(W_max, W_min)."""
    lo, hi = tile_size_bounds(n, m) if m >= 8 else (64, 64)
    out = []
    w = max(64, lo >> below)
    while w <= min(lo << above, max(hi, lo)):
        if _side(n, w) ** 2 <= 1 << 32:
            out.append(w)
        w *= 2
    return out or [lo]


@dataclass
class Trial:
    W: int
    S: int
    ms_per_bfs: float
    gteps: float
    retile_s: float


@dataclass
class Result:
    best: Trial
    trials: list[Trial] = field(default_factory=list)

    def table(self) -> list[dict]:
        return [t.__dict__ for t in self.trials]


def autotune(g, sources, widths=None, stripes=(8, 16, 32), reps: int = 2, verbose: bool = False) -> Result:
    """Time hybrid BFS from ``sources`` under every (W, S) candidate and leave
    ``g`` tiled with the fastest.  The time is the device time of the level
    loop (gphi_stats.total_ms), the reference's timed region
    (apps/bfs_simd/bfs.cpp:805-913); levels are checked to stay identical."""
    import time
    if widths is None:
        widths = candidate_widths(g.n, g.m)
    sources = [int(s) for s in sources]
    ref = None
    trials = []
    for W in widths:
        for S in stripes:
            t0 = time.perf_counter()
            g.retile(W, S)
            rt = time.perf_counter() - t0
            g.bfs(sources[0], parents=False, levels=False)          # warm-up (state allocation)
            ms, edges = [], 0
            lv0 = None
            for r in range(reps):
                for i, s in enumerate(sources):
                    lv, _, st = g.bfs(s, parents=False, levels=(r == 0 and i == 0))
                    ms.append(st.total_ms)
                    edges += st.edges_traversed
                    if lv is not None:
                        lv0 = lv
            if ref is None:
                ref = lv0
            elif not np.array_equal(ref, lv0):
                raise RuntimeError(f"levels changed under W={W} S={S}")
            tot = float(np.sum(ms))
            tr = Trial(W, S, tot / len(ms), edges / (tot * 1e-3) / 1e9, rt)
            trials.append(tr)
            if verbose:
                print(f"[autotune] W={W:6d} S={S:3d}: {tr.ms_per_bfs:8.3f} ms/BFS "
                      f"{tr.gteps:7.1f} GTEPS (retile {rt:.2f} s)", flush=True)
    best = min(trials, key=lambda t: t.ms_per_bfs)
    g.retile(best.W, best.S)
    return Result(best, trials)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    src = ap.add_mutually_exclusive_group(required=True)
    src.add_argument("--workload", help="synthetic workload name (c1, c2, c4)")
    src.add_argument("--text", help="prefix of the reference's text files (col-S-coo-tiled-W + untiled)")
    src.add_argument("--cache", help="a binary cache file")
    ap.add_argument("--W", type=int, default=1024, help="W of the text files (--text)")
    ap.add_argument("--S", type=int, default=16, help="S of the text files (--text)")
    ap.add_argument("--widths", default="", help="comma list of W (default: around the tool's bounds)")
    ap.add_argument("--stripes", default="8,16,32")
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--save", help="write the tuned graph to this cache file")
    args = ap.parse_args(argv)

    import bfs_simd_benchmarks_amd as G
    from bfs_simd_benchmarks_amd.workloads import WORKLOADS, pick_sources
    with G.Context(args.device) as ctx:
        if args.workload:
            g = WORKLOADS[args.workload].generate(ctx)
        elif args.text:
            g = ctx.load_text(args.text, args.W, args.S)
        else:
            g = ctx.load_binary(args.cache)
        try:
            lo, hi = tile_size_bounds(g.n, g.m)
            print(f"[autotune] n={g.n} m={g.m} tool bounds W in [{lo}, {hi}]", flush=True)
            deg = g.download(csr=False, tiles=False)["deg"]
            srcs = pick_sources(deg, args.sources, 12345)
            widths = [int(x) for x in args.widths.split(",") if x] or None
            stripes = [int(x) for x in args.stripes.split(",") if x]
            res = autotune(g, srcs, widths, stripes, args.reps, verbose=True)
            if args.save:
                g.save(args.save)
            print(json.dumps({"best": res.best.__dict__, "bounds": [lo, hi], "trials": res.table()}))
        finally:
            g.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
