"""Failure paths on the GPU: a device allocation that fails part-way through
the traversal state or a re-tiling leaves a graph that later calls either use
correctly or refuse with an error (never one with half-built buffers that a
kernel would run on); a batched BFS whose deepest level is 254 fits its uint8
levels and one of depth 255 is refused."""
import numpy as np
import pytest

import oracle as O
import bfs_simd_benchmarks_amd as G

pytestmark = pytest.mark.gpu


def upload(ctx, g):
    return ctx.graph_from_arrays(g.n, g.deg, W=g.W, S=g.S, csr_edges=g.cols, tile_heads=g.sh,
                                 tile_tails=g.st, tile_offsets=g.soff)


@pytest.mark.parametrize("after", [0, 1, 5, 12])
def test_state_allocation_failure_then_recovery(ctx, c1, after):
    print(f"case: injected state allocation failure after={after}", flush=True)
    gg = upload(ctx, c1)
    try:
        ctx.lib.gphi_debug_fail_alloc(after)
        with pytest.raises(G.GphiError) as ei:
            gg.bfs(0)
        assert ei.value.code == -4
        ctx.lib.gphi_debug_fail_alloc(-1)
        lv, pa, _ = gg.bfs(0)                 # re-allocates the whole state
        assert np.array_equal(lv, O.bfs_levels(c1.n, c1.deg, c1.cols, 0))
        assert O.check_parents(c1.n, c1.deg, c1.cols, 0, lv, pa) == 0
    finally:
        ctx.lib.gphi_debug_fail_alloc(-1)
        gg.free()


@pytest.mark.parametrize("after", [0, 1, 2])
def test_retile_allocation_failure_leaves_no_tiles(ctx, c1, after):
    print(f"case: injected re-tiling allocation failure after={after}", flush=True)
    gg = upload(ctx, c1)
    try:
        ctx.lib.gphi_debug_fail_alloc(after)
        with pytest.raises(G.GphiError):
            gg.retile(512, 8)
        ctx.lib.gphi_debug_fail_alloc(-1)
        with pytest.raises(G.GphiError) as ei:     # no sweep on incomplete tiles
            gg.bfs(0)
        assert ei.value.code == -7
        gg.retile(512, 8)                          # a later re-tiling restores them
        lv, _, _ = gg.bfs(0, parents=False)
        assert np.array_equal(lv, O.bfs_levels(c1.n, c1.deg, c1.cols, 0))
    finally:
        ctx.lib.gphi_debug_fail_alloc(-1)
        gg.free()


@pytest.mark.parametrize("n,ok", [(255, True), (256, False)])
def test_batch_depth_limit(ctx, n, ok):
    # a path 0 -> 1 -> ... -> n-1: source 0 reaches vertex n-1 at level n-1
    print(f"case: batched path graph n={n}", flush=True)
    h = np.arange(n - 1, dtype=np.uint32)
    g = O.tile_edgelist(n, h, h + 1, 64, 2)
    gg = upload(ctx, g)
    try:
        if ok:
            lv, _ = gg.bfs_batch([0, 3])
            assert np.array_equal(lv[0], np.arange(n, dtype=np.uint8))
            assert np.array_equal(lv[1], np.where(np.arange(n) >= 3, np.arange(n) - 3, 255).astype(np.uint8))
        else:
            with pytest.raises(G.GphiError) as ei:
                gg.bfs_batch([0, 3])
            assert ei.value.code == -3
            lv, _ = gg.bfs_batch([3])             # depth 252 still fits
            assert lv[0][n - 1] == n - 4
    finally:
        gg.free()
