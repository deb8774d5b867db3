"""The drop-in command lines, end to end on the reference's file layout:
files written by libgphi's writer (byte-identical to the reference tools),
then ./bfs, ./sssp -w, ./page_rank, ./bellman-ford run as the reference is
run, their outputs checked against the reference's own outputs (golden
fixtures) or the oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import bfs_simd_benchmarks_amd as G

pytestmark = pytest.mark.gpu
BIN = Path(G.__file__).resolve().parent / "bin"
GOLD = Path(__file__).parent / "golden"


def run(*args):
    print("case: cli", " ".join(map(str, args)), flush=True)
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.fixture(scope="module")
def c1_files(tmp_path_factory, c1, c3):
    d = tmp_path_factory.mktemp("c1")
    with G.Context(0) as ctx:
        for tag, g, weighted in (("c1_reorder", c1, False), ("c1_weighted_reorder", c3, True)):
            for order, th, tt, tw, off in ((0, g.sh, g.st, g.sw, g.soff), (1, g.rh, g.rt, g.rw, g.roff)):
                gg = ctx.graph_from_arrays(g.n, g.deg, W=g.W, S=g.S, csr_edges=g.cols,
                                           csr_weights=g.cw if weighted else None, tile_heads=th,
                                           tile_tails=tt, tile_weights=tw if weighted else None,
                                           tile_offsets=off, tile_order=order)
                h = gg.to_host()
                h.write(d / tag)
                h.free()
                gg.free()
    return d


def test_bfs_cli(c1_files, c1):
    out = run(BIN / "bfs", c1_files / "c1", 1024, 16, "-o", c1_files / "lv.txt")
    lines = out.strip().splitlines()
    assert len(lines) == 11 and lines[-1].startswith("Average: 1 ")
    lv = np.loadtxt(c1_files / "lv.txt", dtype=np.int64)
    assert np.array_equal(lv, np.load(GOLD / "c1_ref_levels0.npz")["levels"])


def test_bfs_cli_usage_exit():
    r = subprocess.run([str(BIN / "bfs"), "x"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage: ./bfs <data> <tile_width> <stripe_length>" in r.stdout


def test_bfs_cli_missing_input_fails():
    r = subprocess.run([str(BIN / "bfs"), "/nonexistent/g", "1024", "16"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "cannot open" in r.stderr


def test_sssp_weighted_cli(c1_files, c3):
    out = run(BIN / "sssp", c1_files / "c1", 1024, 16, "-w", "-o", c1_files / "d.txt")
    assert "tile_size: 1024" in out and "stripe_length: 16" in out
    d = np.loadtxt(c1_files / "d.txt", dtype=np.uint64)
    assert np.array_equal(d, O.sssp_dijkstra(c3.n, c3.deg, c3.cols, c3.cw, 0))


def test_sssp_unweighted_cli(c1_files):
    run(BIN / "sssp", c1_files / "c1_reorder", 1024, 16, "-o", c1_files / "h.txt")
    h = np.loadtxt(c1_files / "h.txt", dtype=np.int64)
    assert np.array_equal(h, np.load(GOLD / "c1_ref_levels0.npz")["levels"])


def test_pagerank_cli(c1_files, c1):
    run(BIN / "page_rank", c1_files / "c1_reorder", 1024, 16, "-o", c1_files / "r.txt")
    r = np.loadtxt(c1_files / "r.txt", dtype=np.float64)
    ref = np.load(GOLD / "c1_ref_rank1.npz")["rank"]
    tol = 1e-6 + np.abs(ref) * 2.0**-22 * np.maximum(1, np.bincount(c1.rt, minlength=c1.n))
    assert np.all(np.abs(r - ref) <= tol)


def test_pagerank_cli_stripe_too_long(c1_files):
    r = subprocess.run([str(BIN / "page_rank"), str(c1_files / "c1_reorder"), "1024", "65"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 3


def test_bellman_ford_cli(c1_files):
    import json
    out = run(BIN / "bellman-ford", c1_files / "c1_reorder", 1024, 16)
    rows = [ln.split() for ln in out.strip().splitlines()[-20:]]
    want = json.loads((GOLD / "reference_outputs.json").read_text())["c1"]["bf_dist_0_19"]
    assert [int(r[1]) for r in rows] == want
