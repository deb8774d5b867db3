"""The oracle against the reference's own outputs, frozen as fixtures by
tests/golden/make_golden.py (runs anywhere; no reference tree needed).

Pins: the oracle's format writers (byte digests of every reference tool
output), its BFS levels (apps/sssp_naive hop-distance dump), its hop
relaxation (apps/bellman-ford_simd's printed dist[0..19]) and its FP64
PageRank (apps/pageRank_naive's one-iteration rank dump)."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import refrun as R

HERE = Path(__file__).parent / "golden"
GOLD = json.loads((HERE / "reference_outputs.json").read_text())


def graph_inputs(name):
    if name == "c1":
        n, h, t = O.config_edges("c1")
        return n, h, t, O.CONFIGS["c1"]["W"], O.CONFIGS["c1"]["S"], O.weights(len(h), 3)
    spec = {"tiny": R.TINY, "tiny_odd": R.TINY_ODD}[name]
    h, t = R.tiny_edges(spec)
    return spec["nverts"], h, t, spec["W"], spec["S"], O.weights(len(h), spec["wseed"])


@pytest.fixture(scope="module", params=sorted(GOLD))
def case(request):
    name = request.param
    n, h, t, W, S, w = graph_inputs(name)
    uh, ut, _, deg = O.untiled(n, h, t)
    rank = O.degree_rank(deg)
    g = O.tile_edgelist(n, rank[uh], rank[ut], W, S)
    return name, n, h, t, W, S, w, g


def test_edge_count(case):
    name, n, h, t, *_ = case
    assert (GOLD[name]["n"], GOLD[name]["m"]) == (n, len(h))


def test_oracle_writers_match_reference_digests(case, tmp_path):
    name, n, h, t, W, S, w, g = case
    d = GOLD[name]["digests"]
    o = tmp_path
    uh, ut, _, deg = R.oracle_untiled_files(o / "e", n, h, t)
    assert R.digest(R.files_of(o / "e_untiled")) == d["raw untiled"]
    rank = O.degree_rank(deg)
    O.write_edgelist(o / "e_degreeReordered", n, rank[uh], rank[ut])
    assert R.digest([o / "e_degreeReordered"]) == d["degreeReordered"]
    brank = O.bfs_order_rank(n, deg, ut)
    O.write_edgelist(o / "e_reorder", n, brank[uh], brank[ut])
    assert R.digest([o / "e_reorder"]) == d["bfs_reorder"]
    R.oracle_untiled_files(o / "R", n, g.eh, g.et)
    R.oracle_tiled_files(o / "R", n, g.eh, g.et, W, S)
    R.oracle_untiled_files(o / "Rw", n, g.eh, g.et, w)
    R.oracle_tiled_files(o / "Rw", n, g.eh, g.et, W, S, w)
    for fmt, off in (("untiled", False), (f"coo-tiled-{W}", True), (f"col-{S}-coo-tiled-{W}", True)):
        assert R.digest(R.files_of(o / f"R_{fmt}", off)) == d[fmt], fmt
        assert R.digest(R.files_of(o / f"Rw_{fmt}", off)) == d[f"weighted {fmt}"], fmt


def test_oracle_levels_match_reference_dump(case):
    name, n, *_, g = case
    ref = np.load(HERE / f"{name}_ref_levels0.npz")["levels"]
    lv = O.bfs_levels(n, g.deg, g.cols, 0)
    assert np.array_equal(lv, ref)
    hl, hp, _ = O.bfs_hybrid(g, 0)
    assert np.array_equal(hl, ref)
    assert O.check_parents(n, g.deg, g.cols, 0, hl, hp) == 0


def test_oracle_hops_match_reference_print(case):
    name, n, *_, g = case
    bf = O.bellman_ford_hops(n, g.rh, g.rt, 0, 99)
    assert bf[:20].tolist() == GOLD[name]["bf_dist_0_19"]


def test_oracle_pagerank_matches_reference_dump(case):
    """The reference accumulates in FP32 (omp atomic) and prints "%lf" (6
    decimals): |oracle - ref| <= 1e-6 (printing) + 2^-22 * |rank| (FP32
    accumulation over the in-edges)."""
    name, n, *_, g = case
    ref = np.load(HERE / f"{name}_ref_rank1.npz")["rank"]
    pr = O.pagerank_f64(n, g.rh, g.rt, g.deg, 1, 0.85, 1.0)
    tol = 1e-6 + np.abs(pr) * 2.0**-22 * np.maximum(1, np.bincount(g.rt, minlength=n))
    assert np.all(np.abs(pr - ref) <= tol), float(np.max(np.abs(pr - ref) - tol))
