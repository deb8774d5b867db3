"""Multi-source BFS (gphi_bfs_batch): every row must equal that source's own
BFS levels (bit-exact), for 1..256 sources, duplicates and sources without
out-edges."""
import numpy as np
import pytest

import oracle as O
from bfs_simd_benchmarks_amd.workloads import pick_sources

pytestmark = pytest.mark.gpu


def upload(ctx, g):
    return ctx.graph_from_arrays(g.n, g.deg, W=g.W, S=g.S, csr_edges=g.cols, tile_heads=g.sh,
                                 tile_tails=g.st, tile_offsets=g.soff)


def expected(g, srcs):
    rows = []
    for s in srcs:
        lv = O.bfs_levels(g.n, g.deg, g.cols, int(s))
        rows.append(np.where(lv >= 0, lv, 255).astype(np.uint8))
    return np.stack(rows)


@pytest.mark.parametrize("k", [1, 3, 64, 65, 100, 256])
def test_batch_rows_equal_single_source_levels(ctx, c1, k):
    print(f"case: batch c1 k={k}", flush=True)
    srcs = pick_sources(c1.deg, k, 1000 + k)
    gg = upload(ctx, c1)
    try:
        lv, st = gg.bfs_batch(srcs)
    finally:
        gg.free()
    want = expected(c1, srcs)
    bad = np.argwhere(lv != want)
    assert bad.size == 0, f"{len(bad)} mismatches, first (source row, vertex) {bad[:5].tolist()}"
    reached = (want != 255)
    assert st.reached == int(reached.sum())
    assert st.edges_traversed == int((reached * c1.deg[None, :].astype(np.uint64)).sum())


def test_batch_duplicates_and_isolated_sources(ctx, c1):
    iso = int(np.flatnonzero(c1.deg == 0)[0])
    srcs = np.array([0, 0, iso, 5, 5, 0], np.uint32)
    print("case: batch c1 duplicate/isolated sources", flush=True)
    gg = upload(ctx, c1)
    try:
        lv, _ = gg.bfs_batch(srcs)
        lv2, _ = gg.bfs_batch(srcs[:2])          # buffers reused with a smaller batch
    finally:
        gg.free()
    assert np.array_equal(lv, expected(c1, srcs))
    assert np.array_equal(lv2, expected(c1, srcs[:2]))


def test_batch_narrower_after_wider(ctx, c1):
    """Buffers sized for 256 sources, then batches of 64, 100 and 256 again:
    each runs at its own mask width and stays exact."""
    print("case: batch c1 k=256 then 64, 100, 256", flush=True)
    gg = upload(ctx, c1)
    try:
        for k in (256, 64, 100, 256):
            srcs = pick_sources(c1.deg, k, 3000 + k)
            lv, st = gg.bfs_batch(srcs)
            want = expected(c1, srcs)
            assert np.array_equal(lv, want), f"k={k}"
            assert st.reached == int((want != 255).sum()), f"k={k}"
    finally:
        gg.free()


def test_batch_other_tilings(ctx):
    n, h, t = O.config_edges("c1")
    g = O.build_from_edgelist(n, h, t, 1000, 3)
    srcs = pick_sources(g.deg, 64, 7)
    print("case: batch c1 W=1000 S=3", flush=True)
    gg = upload(ctx, g)
    try:
        lv, _ = gg.bfs_batch(srcs)
    finally:
        gg.free()
    assert np.array_equal(lv, expected(g, srcs))


def test_batch_rejects_bad_arguments(ctx, c1):
    import bfs_simd_benchmarks_amd as G
    gg = upload(ctx, c1)
    try:
        for srcs, code in (([], -6), (list(range(257)), -6), ([c1.n], -3)):
            with pytest.raises(G.GphiError) as ei:
                gg.bfs_batch(np.array(srcs, np.uint32))
            assert ei.value.code == code
    finally:
        gg.free()


@pytest.mark.parametrize("k", [64, 200])
def test_batch_hub_and_piece_boundaries(ctx, k):
    """A graph of 1.57M vertices with one tail of 524K in-edges (128 heavy
    pull pieces) and tails at the light/heavy and one/two-piece boundaries:
    every checked row equals that source's own levels."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from test_gpu_traversals import band_graph
    print(f"case: batch band graph k={k}", flush=True)
    n, h, t, deg = band_graph()
    W, S = 4096, 16
    th, tt, _, off = O.coo_tile(n, W, h, t)
    ch, ct, _, coff = O.col_major(n, W, S, th, tt, None, off)
    gg = ctx.graph_from_arrays(n, deg, W=W, S=S, csr_edges=t, tile_heads=ch, tile_tails=ct, tile_offsets=coff)
    srcs = pick_sources(deg, k, 77 + k)
    try:
        lv, _ = gg.bfs_batch(srcs)
    finally:
        gg.free()
    for i in range(0, k, max(1, k // 8)):
        ref = O.bfs_levels(n, deg, t, int(srcs[i]))
        want = np.where(ref >= 0, ref, 255).astype(np.uint8)
        assert np.array_equal(lv[i], want), f"row {i} (source {srcs[i]}) differs"


@pytest.mark.parametrize("k", [64, 100, 256])
def test_batch_sparse_sources_pushed(ctx, c1, k):
    """Sources on a chain hanging off the graph keep one-vertex frontiers
    while the others' union frontier makes the steps dense: they are pushed
    from their frontier, the rest pulled.  Every row equals that source's own
    levels, before and after the chain reaches the graph (one, two and four
    mask words per vertex)."""
    chain = 12
    n = c1.n + chain
    ids = np.arange(c1.n, n, dtype=np.uint32)
    # c1.n -> c1.n+1 -> ... -> last -> vertex 0 (the largest out-degree)
    h = np.concatenate([c1.csr_heads, ids])
    t = np.concatenate([c1.cols, np.append(ids[1:], np.uint32(0))])
    uh, ut, _, deg = O.untiled(n, h, t)
    W, S = c1.W, c1.S
    th, tt, _, off = O.coo_tile(n, W, uh, ut)
    ch, ct, _, coff = O.col_major(n, W, S, th, tt, None, off)
    srcs = np.concatenate([pick_sources(c1.deg, k - 4, 4242), ids[[0, 3, 7, 11]]]).astype(np.uint32)
    print(f"case: batch c1+chain k={k} sparse sources", flush=True)
    gg = ctx.graph_from_arrays(n, deg, W=W, S=S, csr_edges=ut, tile_heads=ch, tile_tails=ct, tile_offsets=coff)
    try:
        lv, st = gg.bfs_batch(srcs)
    finally:
        gg.free()
    assert any(st.dir[i] for i in range(st.steps)), "no dense step"
    for i, s in enumerate(srcs):
        ref = O.bfs_levels(n, deg, ut, int(s))
        want = np.where(ref >= 0, ref, 255).astype(np.uint8)
        assert np.array_equal(lv[i], want), f"row {i} (source {s}) differs at {np.flatnonzero(lv[i] != want)[:8]}"
