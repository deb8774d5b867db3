"""Workload definitions: the synthetic inputs the benchmarks and tests run on.

The reference ships no generator (its inputs came from PBBS rMat and Galois,
tools/adj2snap/adj2snap.cpp:26, tools/galois_data/process.sh:21-34), so these
specs are this project's own: RMAT a=0.57 b=0.19 c=0.19 d=0.05, self-loops
and duplicates removed, ids degree-reordered with the reference tool's rule
(tools/vertex_id_reorder_to_degrees/bfs.cpp:50-83).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MASK64 = (1 << 64) - 1


@dataclass(frozen=True)
class Workload:
    name: str
    scale: int
    nverts: int
    seed: int
    samples: int
    target_unique: int = 0
    perm_seed: int = 0
    W: int = 1024
    S: int = 16
    weight_seed: int = 0

    def generate(self, ctx, W: int | None = None, S: int | None = None, weighted: bool = False,
                 tile_order: int = 0, reorder: int = 1):
        """reorder: 1 degree order (the workloads' own), 2 BFS order, 0 none."""
        return ctx.generate(scale=self.scale, nverts=self.nverts, seed=self.seed,
                            samples=self.samples, target_unique=self.target_unique,
                            perm_seed=self.perm_seed, W=W or self.W, S=S or self.S,
                            reorder=reorder, weight_seed=self.weight_seed if weighted else 0,
                            tile_order=tile_order)


WORKLOADS = {
    # scale 16, edge factor 16, seed 1, W=1024 S=16, source 0
    "c1": Workload("rmat16-ef16", 16, 1 << 16, 1, 16 << 16),
    # pokec shape: 1,632,803 V / 30,622,564 unique edges, seed 2, ids permuted then degree-reordered
    "c2": Workload("pokec-shape", 21, 1_632_803, 2, 1 << 27, target_unique=30_622_564, perm_seed=2),
    # c1 graph + uint32 weights uniform in [1,255], seed 3
    "c3": Workload("rmat16-ef16-weighted", 16, 1 << 16, 1, 16 << 16, weight_seed=3),
    # scale 25, edge factor 36 (~1.2e9 samples), seed 4; W=2048 S=32, the
    # autotuner's choice over W in 512..4096 and S in 8, 16, 32
    "c4": Workload("rmat25-ef36", 25, 1 << 25, 4, 36 << 25, W=2048, S=32),
}

# serving mix: scales 18..24, edge factor 16, seeds 10..17, every other graph weighted
SERVING_MIX = [Workload(f"rmat{s}-ef16-s{seed}", s, 1 << s, seed, 16 << s,
                        weight_seed=(seed + 100) if i % 2 else 0)
               for i, (s, seed) in enumerate(zip((18, 19, 20, 21, 22, 22, 23, 24), range(10, 18)))]


def serving_stream(queries: int, qps: float, seed: int = 5):
    """The serving mix's query stream: Poisson arrivals at `qps` offered
    queries/s; 20% SSSP on a uniform random weighted graph, 80% BFS on a
    uniform random graph; a uniform random source.  Returns (arrival times in
    s, is_sssp bool[], graph index[], source[] given each graph's vertex count
    through `sources(nverts)`)."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, queries))
    weighted_ids = [i for i, w in enumerate(SERVING_MIX) if w.weight_seed]
    is_sssp = rng.random(queries) < 0.2
    gsel = np.where(is_sssp, rng.choice(weighted_ids, queries),
                    rng.integers(0, len(SERVING_MIX), queries))
    unif = rng.random(queries)

    def sources(nverts):
        return (unif * np.asarray([nverts[i] for i in gsel])).astype(np.int64)
    return arrivals, is_sssp, gsel, sources


def _sm64(state: int) -> tuple[int, int]:
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


def pick_sources(deg: np.ndarray, count: int, seed: int) -> np.ndarray:
    """The first `count` distinct vertices with out-degree > 0 of a seeded
    splitmix64 stream (v = x mod N)."""
    n = len(deg)
    st = (seed * 0x2545F4914F6CDD1D) & MASK64
    out, seen = [], set()
    for _ in range(1000 * count + 10000):
        st, x = _sm64(st)
        v = x % n
        if deg[v] > 0 and v not in seen:
            seen.add(v)
            out.append(v)
            if len(out) == count:
                break
    return np.array(out, np.uint32)
