"""The device-side generator and tool pipeline (gphi_graph_generate) must lay
out exactly the arrays the oracle's restatement of the reference tools
produces: degrees, CSR order, tile order and weights, element for element."""
import numpy as np
import pytest

import oracle as O
from bfs_simd_benchmarks_amd.workloads import WORKLOADS

pytestmark = pytest.mark.gpu


def _cmp(a, b, what):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}"


def test_generate_c1_matches_tool_pipeline(ctx, c1):
    print("case: generate c1", flush=True)
    g = WORKLOADS["c1"].generate(ctx)
    try:
        assert (g.n, g.m) == (c1.n, c1.m)
        d = g.download()
    finally:
        g.free()
    _cmp(d["deg"], c1.deg, "deg")
    _cmp(d["cols"], c1.cols, "csr")
    _cmp(d["heads"], c1.sh, "tile heads")
    _cmp(d["tails"], c1.st, "tile tails")


def test_generate_c3_weights(ctx, c3):
    print("case: generate c3 weighted", flush=True)
    g = WORKLOADS["c3"].generate(ctx, weighted=True)
    try:
        assert g.weighted
        d = g.download()
    finally:
        g.free()
    _cmp(d["cols"], c3.cols, "csr")
    _cmp(d["cw"], c3.cw, "csr weights")
    _cmp(d["tw"], c3.sw, "tile weights")
    assert d["cw"].min() >= 1 and d["cw"].max() <= 255


@pytest.mark.parametrize("W,S,order", [(256, 4, 0), (1000, 3, 0), (1024, 16, 1), (64, 1, 0)])
def test_generate_c1_other_tilings(ctx, W, S, order):
    print(f"case: generate c1 W={W} S={S} order={order}", flush=True)
    n, h, t = O.config_edges("c1")
    ref = O.build_from_edgelist(n, h, t, W, S)
    g = WORKLOADS["c1"].generate(ctx, W=W, S=S, tile_order=order)
    try:
        d = g.download(csr=False)
    finally:
        g.free()
    _cmp(d["heads"], ref.rh if order else ref.sh, "tile heads")
    _cmp(d["tails"], ref.rt if order else ref.st, "tile tails")


@pytest.mark.parametrize("name,weighted", [("c1", False), ("c3", True)])
def test_generate_bfs_order_matches_tool_pipeline(ctx, name, weighted):
    """GPHI_REORDER_BFS: the relabelling of tools/vertex_id_remap, whose oracle
    restatement reproduces that tool's output file byte for byte."""
    print(f"case: generate {name} BFS order", flush=True)
    import bfs_simd_benchmarks_amd as G
    n, h, t = O.config_edges("c1")
    ref = O.build_from_edgelist(n, h, t, 1024, 16, reorder="bfs")
    if weighted:
        ref = O.tile_edgelist(n, ref.eh, ref.et, 1024, 16, O.weights(ref.m, 3))
    g = WORKLOADS[name].generate(ctx, weighted=weighted, reorder=G.REORDER_BFS)
    try:
        assert (g.n, g.m) == (ref.n, ref.m)
        d = g.download()
        lv, _, _ = g.bfs(0)
    finally:
        g.free()
    _cmp(d["deg"], ref.deg, "deg")
    _cmp(d["cols"], ref.cols, "csr")
    _cmp(d["heads"], ref.sh, "tile heads")
    _cmp(d["tails"], ref.st, "tile tails")
    if weighted:
        _cmp(d["cw"], ref.cw, "csr weights")
    # BFS order: new ids rise with the BFS level from vertex 0
    reached = lv[lv >= 0]
    assert np.all(np.diff(reached) >= 0) and np.all(lv[: reached.size] >= 0)


def test_generate_c2_pokec_shape(ctx):
    print("case: generate c2 (first 30,622,564 unique samples, permuted ids)", flush=True)
    ref = O.config_graph("c2")
    g = WORKLOADS["c2"].generate(ctx)
    try:
        assert (g.n, g.m) == (1_632_803, 30_622_564)
        d = g.download()
    finally:
        g.free()
    _cmp(d["deg"], ref.deg, "deg")
    _cmp(d["cols"], ref.cols, "csr")
    _cmp(d["heads"], ref.sh, "tile heads")
    _cmp(d["tails"], ref.st, "tile tails")


def test_retile_c3_matches_tool_pipeline(ctx, c3):
    """Re-tiling a resident graph (the autotuner's step) lays out the same
    tile arrays as running the tiling tools again with the new (W, S)."""
    g = WORKLOADS["c3"].generate(ctx, weighted=True)
    try:
        for W, S, order in [(256, 4, 0), (4096, 2, 1), (1000, 3, 0), (1024, 16, 0)]:
            print(f"case: retile c3 W={W} S={S} order={order}", flush=True)
            ref = O.tile_edgelist(c3.n, c3.eh, c3.et, W, S, c3.ew)
            g.retile(W, S, order)
            assert g.side == ref.n // W + (ref.n % W > 0)
            d = g.download()
            _cmp(d["heads"], ref.rh if order else ref.sh, "tile heads")
            _cmp(d["tails"], ref.rt if order else ref.st, "tile tails")
            _cmp(d["tw"], ref.rw if order else ref.sw, "tile weights")
            if order == 0:
                lv, _, _ = g.bfs(0)
                assert np.array_equal(lv, O.bfs_levels(c3.n, c3.deg, c3.cols, 0))
                dist, _ = g.sssp(0)
                assert np.array_equal(dist, O.sssp_dijkstra(c3.n, c3.deg, c3.cols, c3.cw, 0))
    finally:
        g.free()
