"""Multi-source weighted SSSP (gphi_sssp_batch): every row must equal that
source's shortest distances (Dijkstra, bit-exact) and gphi_sssp's, for 1..16
sources, duplicates, a source without out-edges, 1-byte and 32-bit weights."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from bfs_simd_benchmarks_amd.workloads import pick_sources

pytestmark = pytest.mark.gpu


def upload(ctx, g):
    return ctx.graph_from_arrays(g.n, g.deg, W=g.W, S=g.S, csr_edges=g.cols, csr_weights=g.cw,
                                 tile_heads=g.sh, tile_tails=g.st, tile_weights=g.sw, tile_offsets=g.soff)


def check_rows(g, srcs, d):
    for i, s in enumerate(srcs):
        ref = O.sssp_dijkstra(g.n, g.deg, g.cols, g.cw, int(s))
        assert np.array_equal(d[i], ref), f"row {i} (source {s}) differs at {np.flatnonzero(d[i] != ref)[:10]}"


@pytest.mark.parametrize("k", [1, 5, 16])
def test_sssp_batch_rows_equal_dijkstra(ctx, c3, k):
    print(f"case: sssp batch c3 k={k}", flush=True)
    srcs = pick_sources(c3.deg, k, 500 + k)
    gg = upload(ctx, c3)
    try:
        d, st = gg.sssp_batch(srcs)
        one, _ = gg.sssp(int(srcs[-1]))   # the single-source path on the same handle
    finally:
        gg.free()
    check_rows(c3, srcs, d)
    assert np.array_equal(d[-1], one)
    assert st.steps > 0


def test_sssp_batch_duplicates_and_isolated_sources(ctx, c3):
    iso = int(np.flatnonzero(c3.deg == 0)[0])
    srcs = np.array([7, 7, iso, 0, 7], np.uint32)
    print("case: sssp batch c3 duplicate/isolated sources", flush=True)
    gg = upload(ctx, c3)
    try:
        d, _ = gg.sssp_batch(srcs)
        d2, _ = gg.sssp_batch(srcs[:2])   # state reused by a smaller batch
    finally:
        gg.free()
    check_rows(c3, srcs, d)
    check_rows(c3, srcs[:2], d2)
    assert (d[2] != 2147483647).sum() == 1   # only the isolated source itself


def test_sssp_batch_wide_weights(ctx, c3):
    """Weights above 255: the dense steps read the 32-bit weights."""
    scale = np.uint32(4099)
    g = dataclasses.replace(c3, cw=c3.cw * scale, sw=c3.sw * scale, rw=c3.rw * scale)
    assert int(g.cw.max()) > 255
    srcs = pick_sources(g.deg, 16, 77)
    print("case: sssp batch c3 x4099 k=16", flush=True)
    gg = upload(ctx, g)
    try:
        d, _ = gg.sssp_batch(srcs)
    finally:
        gg.free()
    check_rows(g, srcs, d)


def test_sssp_batch_rejects_bad_arguments(ctx, c1, c3):
    import bfs_simd_benchmarks_amd as G
    gg = upload(ctx, c3)
    try:
        for srcs, code in (([], -6), (list(range(17)), -6), ([c3.n], -3)):
            with pytest.raises(G.GphiError) as ei:
                gg.sssp_batch(np.array(srcs, np.uint32))
            assert ei.value.code == code
    finally:
        gg.free()
    gu = ctx.graph_from_arrays(c1.n, c1.deg, W=c1.W, S=c1.S, csr_edges=c1.cols, tile_heads=c1.sh,
                               tile_tails=c1.st, tile_offsets=c1.soff)
    try:
        with pytest.raises(G.GphiError) as ei:
            gu.sssp_batch(np.array([0], np.uint32))
        assert ei.value.code == -7   # no weights
    finally:
        gu.free()
