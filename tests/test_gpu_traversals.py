"""GPU parity of the hot path against the oracle (bit-exact levels/distances,
valid parents, identical direction trace; PageRank within the stated tolerance).

Every call goes through the C ABI of libgphi.so (HIP kernels for gfx950)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def upload(ctx, g, weighted=False, csr=True, tiles=True):
    return ctx.graph_from_arrays(
        g.n, g.deg, W=g.W, S=g.S,
        csr_edges=g.cols if csr else None,
        csr_weights=g.cw if (weighted and csr) else None,
        tile_heads=g.sh if tiles else None, tile_tails=g.st if tiles else None,
        tile_weights=g.sw if (weighted and tiles) else None,
        tile_offsets=g.soff if tiles else None)


def check_bfs(ctx, g, src, t_ratio=20):
    gg = upload(ctx, g)
    try:
        lv, pa, st = gg.bfs(src)
    finally:
        gg.free()
    ref = O.bfs_levels(g.n, g.deg, g.cols, src)
    assert np.array_equal(lv, ref), f"levels differ at {np.flatnonzero(lv != ref)[:10]}"
    bad = O.check_parents(g.n, g.deg, g.cols, src, lv, pa)
    assert bad == 0, f"invalid parent at vertex {bad - 1}"
    _, _, tr = O.bfs_hybrid(g, src, t_ratio)
    assert np.array_equal(st.trace(), tr), (st.trace(), tr)
    assert st.reached == int((ref >= 0).sum())
    assert st.edges_traversed == int(g.deg[ref >= 0].astype(np.uint64).sum())
    return lv, st


def test_bfs_c1_source0(ctx, c1):
    print("case: bfs c1 src=0", flush=True)
    check_bfs(ctx, c1, 0)


@pytest.mark.parametrize("src", [1, 7, 100, 4097, 65535])
def test_bfs_c1_sources(ctx, c1, src):
    print(f"case: bfs c1 src={src}", flush=True)
    check_bfs(ctx, c1, src)


@pytest.mark.parametrize("W,S", [(256, 4), (64, 1), (4096, 16), (1000, 3), (65536, 1), (131072, 1)])
def test_bfs_c1_tilings(ctx, W, S):
    print(f"case: bfs c1 W={W} S={S}", flush=True)
    n, h, t = O.config_edges("c1")
    g = O.build_from_edgelist(n, h, t, W, S)
    check_bfs(ctx, g, 0)


def test_bfs_tiny_graphs(ctx):
    # path 0->1->2->3, an isolated vertex 4, a back edge 3->0 and a duplicate
    for n, edges in [(5, [(0, 1), (1, 2), (2, 3), (3, 0), (1, 2)]),
                     (1, []),
                     (3, [(2, 1)]),
                     (130, [(i, i + 1) for i in range(129)])]:
        h = np.array([e[0] for e in edges], np.uint32)
        t = np.array([e[1] for e in edges], np.uint32)
        g = O.tile_edgelist(n, h, t, 64, 2) if n >= 64 else O.tile_edgelist(n, h, t, 2, 1)
        for src in range(min(n, 4)):
            print(f"case: tiny n={n} src={src}", flush=True)
            check_bfs(ctx, g, src)


def test_bfs_source_without_out_edges(ctx, c1):
    src = int(np.flatnonzero(c1.deg == 0)[0])
    print(f"case: bfs c1 isolated src={src}", flush=True)
    lv, st = check_bfs(ctx, c1, src)
    assert (lv >= 0).sum() == 1


def test_bfs_repeated_calls_reuse_state(ctx, c1):
    gg = upload(ctx, c1)
    try:
        for src in (0, 5, 0):
            lv, pa, _ = gg.bfs(src)
            assert np.array_equal(lv, O.bfs_levels(c1.n, c1.deg, c1.cols, src))
    finally:
        gg.free()


def test_bfs_bad_source_raises(ctx, c1):
    import bfs_simd_benchmarks_amd as G
    gg = upload(ctx, c1)
    try:
        with pytest.raises(G.GphiError) as ei:
            gg.bfs(c1.n)
        assert ei.value.code == -3
    finally:
        gg.free()


@pytest.mark.parametrize("src", [0, 3, 999, 20000])
def test_sssp_c3(ctx, c3, src):
    print(f"case: sssp c3 src={src}", flush=True)
    gg = upload(ctx, c3, weighted=True)
    try:
        d, st = gg.sssp(src)
    finally:
        gg.free()
    ref = O.sssp_dijkstra(c3.n, c3.deg, c3.cols, c3.cw, src)
    assert np.array_equal(d, ref), f"dist differ at {np.flatnonzero(d != ref)[:10]}"


def test_pagerank_c1_vs_f64(ctx, c1):
    print("case: pagerank c1", flush=True)
    gg = ctx.graph_from_arrays(c1.n, c1.deg, W=c1.W, S=c1.S, tile_heads=c1.rh, tile_tails=c1.rt,
                               tile_offsets=c1.roff, tile_order=1)
    try:
        r, ms = gg.pagerank(20, 0.85, 1.0 / c1.n)
    finally:
        gg.free()
    ref = O.pagerank_f64(c1.n, c1.rh, c1.rt, c1.deg, 20, 0.85, 1.0 / c1.n)
    l1 = float(np.abs(r.astype(np.float64) - ref).sum())
    print(f"pagerank L1 vs fp64 = {l1:.3e}")
    assert l1 < 1e-6


def test_pagerank_fixed_summation_order(ctx, c1):
    """Every in-edge sum is added in a fixed order (no atomics), so repeated
    runs, on the same handle and on a view with its own stream, give the same
    FP32 bits; c1's largest in-degree (6310) spans two 4096-edge pieces."""
    print("case: pagerank c1 repeatable", flush=True)
    indeg = np.bincount(c1.cols, minlength=c1.n)
    assert indeg.max() > 4096
    gg = ctx.graph_from_arrays(c1.n, c1.deg, W=c1.W, S=c1.S, csr_edges=c1.cols, tile_heads=c1.sh,
                               tile_tails=c1.st, tile_offsets=c1.soff)
    import bfs_simd_benchmarks_amd as G
    other = G.Context(0)
    v = other.view(gg)
    try:
        a, _ = gg.pagerank(20, 0.85, 1.0 / c1.n)
        b, _ = gg.pagerank(20, 0.85, 1.0 / c1.n)
        c, _ = v.pagerank(20, 0.85, 1.0 / c1.n)
    finally:
        v.free()
        gg.free()
        other.close()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    ref = O.pagerank_f64(c1.n, c1.csr_heads, c1.cols, c1.deg, 20, 0.85, 1.0 / c1.n)
    assert float(np.abs(a.astype(np.float64) - ref).sum()) < 1e-6


def band_graph():
    """N = 3 * 2^19 + 123 vertices, so the PageRank sweep cuts the heads into
    four bands of 2^19 (every window carries far more than M/64 edges).
    Tail 0 gets an in-edge from every third vertex (about 175K per band: heavy
    pairs of many pieces in every band); tails 1-4 get 64, 65, 4096 and 4097
    in-edges from heads 20.. (the light/heavy and one/two-piece boundaries);
    every vertex has 4 random out-edges to tails >= 16."""
    n = 3 * (1 << 19) + 123
    rng = np.random.default_rng(7)
    hs = [np.arange(3, n, 3, dtype=np.int64)]
    ts = [np.zeros(len(hs[0]), np.int64)]
    for t, cnt in ((1, 64), (2, 65), (3, 4096), (4, 4097)):
        hs.append(np.arange(20, 20 + cnt, dtype=np.int64))
        ts.append(np.full(cnt, t, np.int64))
    rh = np.repeat(np.arange(n, dtype=np.int64), 4)
    rt = rng.integers(16, n, len(rh))
    keep = rh != rt
    hs.append(rh[keep])
    ts.append(rt[keep])
    h, t = np.concatenate(hs), np.concatenate(ts)
    order = np.lexsort((t, h))
    h, t = h[order].astype(np.uint32), t[order].astype(np.uint32)
    return n, h, t, np.bincount(h, minlength=n).astype(np.uint32)


def test_pagerank_head_bands(ctx):
    """PageRank over several head bands, with light pairs, one-piece and
    many-piece heavy pairs in each: equal to the FP64 oracle, the same bits
    on a second run (the per-pair piece counters are left at zero), and the
    same bits from the row-major tiles alone with the edges shuffled inside
    every tile (the in-edges are re-sorted by head within each tail)."""
    print("case: pagerank head bands", flush=True)
    n, h, t, deg = band_graph()
    W = 4096
    gg = ctx.graph_from_arrays(n, deg, W=W, S=16, csr_edges=t)
    try:
        a, _ = gg.pagerank(20, 0.85, 1.0 / n)
        b, _ = gg.pagerank(20, 0.85, 1.0 / n)
    finally:
        gg.free()
    ref = O.pagerank_f64(n, h, t, deg, 20, 0.85, 1.0 / n)
    l1 = float(np.abs(a.astype(np.float64) - ref).sum())
    print(f"pagerank head bands L1 vs fp64 = {l1:.3e}")
    assert l1 < 1e-6
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    th, tt, _, off = O.coo_tile(n, W, h, t)
    tile = np.searchsorted(off.astype(np.int64), np.arange(len(th)), side="right") - 1
    shuf = np.lexsort((np.random.default_rng(3).random(len(th)), tile))
    gt = ctx.graph_from_arrays(n, deg, W=W, S=16, tile_heads=th[shuf], tile_tails=tt[shuf], tile_offsets=off,
                               tile_order=1)
    try:
        c, _ = gt.pagerank(20, 0.85, 1.0 / n)
    finally:
        gt.free()
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_bellman_ford_c1(ctx, c1):
    print("case: bellman-ford c1", flush=True)
    gg = ctx.graph_from_arrays(c1.n, c1.deg, W=c1.W, S=c1.S, tile_heads=c1.rh, tile_tails=c1.rt,
                               tile_offsets=c1.roff, tile_order=1)
    try:
        d, run, ms = gg.bellman_ford(0, 99)
    finally:
        gg.free()
    ref = O.bellman_ford_hops(c1.n, c1.rh, c1.rt, 0, 99)
    assert np.array_equal(d, ref)
    lv = O.bfs_levels(c1.n, c1.deg, c1.cols, 0)
    assert np.array_equal(d, np.where(lv >= 0, lv, O.INT_MAX - 2).astype(np.uint32))
    assert run < 99


def test_bellman_ford_unbounded_sweeps(ctx, c1):
    """max_sweeps = 2^32-1 must not size any buffer (regression: a counter
    buffer of (max_sweeps + 4) * 4 bytes wrapped in 32-bit arithmetic)."""
    print("case: bellman-ford c1 max_sweeps=2^32-1", flush=True)
    gg = ctx.graph_from_arrays(c1.n, c1.deg, W=c1.W, S=c1.S, tile_heads=c1.rh, tile_tails=c1.rt,
                               tile_offsets=c1.roff, tile_order=1)
    try:
        d, run, _ = gg.bellman_ford(0, 0xFFFFFFFF)
    finally:
        gg.free()
    assert np.array_equal(d, O.bellman_ford_hops(c1.n, c1.rh, c1.rt, 0, 99)) and run < 99


@pytest.mark.parametrize("src", [0, 999])
def test_sssp_wide_weights(ctx, c3, src):
    # weights above 255: the sweep reads the 32-bit weights (no 1-byte copy)
    import dataclasses
    scale = np.uint32(4099)
    g = dataclasses.replace(c3, cw=c3.cw * scale, sw=c3.sw * scale, rw=c3.rw * scale)
    print(f"case: sssp c3 x{int(scale)} src={src}", flush=True)
    gg = upload(ctx, g, weighted=True)
    try:
        d, st = gg.sssp(src)
    finally:
        gg.free()
    ref = O.sssp_dijkstra(g.n, g.deg, g.cols, g.cw, src)
    assert int(g.cw.max()) > 255
    assert np.array_equal(d, ref), f"dist differ at {np.flatnonzero(d != ref)[:10]}"


def test_bfs_level_timing_off(ctx, c1):
    """Per-level event timing off: same levels and trace, only the total time.
    The loop then enqueues its levels back to back with no events between
    them, so it is also run from other sources, after a retile, and on a
    path deeper than one 8-level group, each against the oracle."""
    gg = upload(ctx, c1)
    try:
        lv_on, _, st_on = gg.bfs(0)
        ctx.set_level_timing(False)
        try:
            lv_off, _, st_off = gg.bfs(0)
            for src in (0, 5, 1000, 0):
                print(f"case: timing off, source {src}", flush=True)
                lv, pa, _ = gg.bfs(src)
                ref = O.bfs_levels(c1.n, c1.deg, c1.cols, src)
                assert np.array_equal(lv, ref), src
                assert O.check_parents(c1.n, c1.deg, c1.cols, src, lv, pa) == 0, src
            print("case: timing off after a retile to W=512 S=8", flush=True)
            gg.retile(512, 8)
            lv, _, st = gg.bfs(5)
            assert np.array_equal(lv, O.bfs_levels(c1.n, c1.deg, c1.cols, 5))
            assert st.total_ms > 0
        finally:
            ctx.set_level_timing(True)
        assert np.array_equal(lv_on, lv_off)
        assert st_on.steps == st_off.steps and np.array_equal(st_on.trace(), st_off.trace())
        assert st_off.edges_traversed == st_on.edges_traversed
        assert not any(st_off.sweep_times()) and not any(st_off.update_times())
        assert st_off.total_ms > 0 and any(st_on.sweep_times())
    finally:
        gg.free()
    n = 40  # a path 0 -> 1 -> ... -> 39: 39 levels, past the first 8
    path = O.tile_edgelist(n, np.arange(n - 1, dtype=np.uint32), np.arange(1, n, dtype=np.uint32), 4, 2)
    pg = upload(ctx, path)
    ctx.set_level_timing(False)
    try:
        for src in (0, 30):
            print(f"case: timing off, path of {n}, source {src}", flush=True)
            lv, _, st = pg.bfs(src)
            assert np.array_equal(lv, O.bfs_levels(n, path.deg, path.cols, src)), src
            assert st.steps == n - src
    finally:
        ctx.set_level_timing(True)
        pg.free()


def brute_force_step_bytes(g, lv, trace):
    """The byte model of each expanded frontier, from the levels alone:
    sparse step  4|F| + 8|F| + 4 D_F + 4 D_F + 12 |F'|
    dense step   8 E_act + 4 T_act + 3 ceil(N/8) + 8 |F'|
    plus ceil(N/8) on a switch to dense, ceil(N/8) + 4|F| on a switch to sparse.
    E_act / T_act: edges / count of the non-empty tiles whose head row holds a
    frontier vertex and whose tail column still has an unvisited vertex when
    the step begins -- counted here tile by tile over every edge."""
    W, n = g.W, g.n
    side = -(-n // W)
    heads = g.csr_heads.astype(np.int64)
    tails = g.cols.astype(np.int64)
    tile = (heads // W) * side + tails // W
    nbm = (n + 7) // 8
    out = []
    for k, (dense, f, dF) in enumerate(trace):
        fn = trace[k + 1][1] if k + 1 < len(trace) else 0
        if dense:
            row_hot = np.zeros(side, bool)
            row_hot[np.flatnonzero(lv == k) // W] = True
            col_open = np.zeros(side, bool)
            col_open[np.flatnonzero((lv > k) | (lv < 0)) // W] = True
            act = row_hot[heads // W] & col_open[tails // W]
            e_act = int(act.sum())
            t_act = len(np.unique(tile[act]))
            b = 8 * e_act + 4 * t_act + 3 * nbm + 8 * fn
        else:
            b = 4 * f + 8 * f + 4 * dF + 4 * dF + 12 * fn
        if k > 0 and dense != trace[k - 1][0]:
            b += nbm if dense else nbm + 4 * f
        out.append(b)
    return np.array(out, np.float64)


@pytest.mark.parametrize("W,S,src", [(1024, 16, 0), (1024, 16, 17), (256, 3, 0)])
def test_alg_bytes_equal_brute_force(ctx, W, S, src):
    """gphi_bfs_alg_bytes (E_act/T_act counted on the device) equals a
    tile-by-tile count over every edge; this is the numerator of the roofline."""
    print(f"case: alg bytes W={W} S={S} src={src}", flush=True)
    n, h, t = O.config_edges("c1")
    g = O.build_from_edgelist(n, h, t, W, S)
    gg = upload(ctx, g)
    try:
        lv, _, st = gg.bfs(src, parents=False)
        got = gg.alg_bytes(lv, st)
    finally:
        gg.free()
    tr = st.trace()
    assert any(r[0] == 1 for r in tr)
    want = brute_force_step_bytes(g, lv, tr.tolist())
    assert np.array_equal(got, want), (got, want)
