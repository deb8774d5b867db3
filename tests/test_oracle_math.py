"""Pins of the oracle that do not depend on the reference: closed forms,
brute force on tiny inputs and invariants, chosen so that a plausible
restatement mistake (a dropped term, an off-by-one level, a transposed edge,
a wrong tie rule) fails one of them."""
import ctypes

import numpy as np
import pytest

import oracle as O


def random_graph(n, m, seed, loops=False):
    rng = np.random.default_rng(seed)
    h = rng.integers(0, n, m).astype(np.uint32)
    t = rng.integers(0, n, m).astype(np.uint32)
    if not loops:
        keep = h != t
        h, t = h[keep], t[keep]
    k = np.unique((h.astype(np.uint64) << 32) | t)
    return (k >> 32).astype(np.uint32), (k & 0xFFFFFFFF).astype(np.uint32)


def dense_levels(n, h, t, src):
    """BFS levels by repeated boolean matrix-vector products."""
    A = np.zeros((n, n), bool)
    A[h, t] = True
    lv = np.full(n, -1, np.int32)
    lv[src] = 0
    front = np.zeros(n, bool)
    front[src] = True
    k = 0
    while front.any():
        k += 1
        nxt = A[front].any(axis=0) & (lv < 0)
        lv[nxt] = k
        front = nxt
    return lv


def test_glibc_rand_restatement_matches_libc():
    libc = ctypes.CDLL(None)
    libc.rand.restype = ctypes.c_int
    libc.srand(1)
    want = np.array([libc.rand() for _ in range(5000)], np.int32)
    assert np.array_equal(O.glibc_rand(1, 5000), want)
    libc.srand(12345)
    want = np.array([libc.rand() for _ in range(500)], np.int32)
    assert np.array_equal(O.glibc_rand(12345, 500), want)


@pytest.mark.parametrize("seed", range(6))
def test_bfs_levels_brute_force(seed):
    n = 60 + seed * 7
    h, t = random_graph(n, 3 * n, seed)
    g = O.tile_edgelist(n, h, t, 8, 2)
    for src in (0, 5, n - 1):
        ref = dense_levels(n, h, t, src)
        assert np.array_equal(O.bfs_levels(n, g.deg, g.cols, src), ref)
        lv, pa, tr = O.bfs_hybrid(g, src, t_ratio=20)
        assert np.array_equal(lv, ref)
        assert O.check_parents(n, g.deg, g.cols, src, lv, pa) == 0


def test_hybrid_direction_trace_closed_form():
    """dir(0) = sparse; dir(k) = dense iff |F_k| + sum outdeg(F_k) > M/20 (uint32),
    with F_k the vertices at level k (apps/bfs_simd/bfs.cpp:807-834)."""
    g = O.config_graph("c1")
    for src in (0, 17, 3001):
        lv, _, tr = O.bfs_hybrid(g, src)
        levels = lv.max() + 1
        want = []
        for k in range(levels):
            F = lv == k
            f, d = int(F.sum()), int(g.deg[F].astype(np.uint64).sum())
            dense = 0 if k == 0 else int(f + d > g.m // 20)
            want.append((dense, f, d))
        assert [tuple(map(int, r)) for r in tr] == want


def test_parent_checker_rejects_planted_mistakes():
    g = O.config_graph("c1")
    lv, pa, _ = O.bfs_hybrid(g, 0)
    assert O.check_parents(g.n, g.deg, g.cols, 0, lv, pa) == 0
    v = int(np.flatnonzero(lv == 2)[0])
    bad = pa.copy()
    bad[v] = int(np.flatnonzero(lv == 2)[1])           # same-level vertex: not a parent
    assert O.check_parents(g.n, g.deg, g.cols, 0, lv, bad) == v + 1
    bad = pa.copy()
    bad[v] = int(np.flatnonzero(lv == 1)[-1]) if not np.isin(v, g.cols) else bad[v]
    u = int(np.flatnonzero(lv == 1)[0])
    if v not in set(g.cols[int(O.row_offsets(g.deg)[u]):int(O.row_offsets(g.deg)[u + 1])]):
        bad[v] = u                                     # level-1 vertex without the edge
        assert O.check_parents(g.n, g.deg, g.cols, 0, lv, bad) == v + 1
    badl = lv.copy()
    badl[v] += 1                                       # off-by-one level
    assert O.check_parents(g.n, g.deg, g.cols, 0, badl, pa) != 0


@pytest.mark.parametrize("seed", range(4))
def test_dijkstra_brute_force_and_unit_weights(seed):
    n = 50
    h, t = random_graph(n, 200, 100 + seed)
    w = np.random.default_rng(seed).integers(1, 256, len(h)).astype(np.uint32)
    g = O.tile_edgelist(n, h, t, 8, 2, w)
    # Bellman-Ford by dense relaxation
    D = np.full(n, np.iinfo(np.int64).max // 4, np.int64)
    D[0] = 0
    for _ in range(n):
        cand = D[g.eh] + g.ew
        np.minimum.at(D, g.et, cand)
    ref = np.where(D >= np.iinfo(np.int64).max // 8, O.INT_MAX, D).astype(np.uint32)
    assert np.array_equal(O.sssp_dijkstra(n, g.deg, g.cols, g.cw, 0), ref)
    ones = np.ones(len(h), np.uint32)
    lv = O.bfs_levels(n, g.deg, g.cols, 0)
    d1 = O.sssp_dijkstra(n, g.deg, g.cols, ones, 0)
    assert np.array_equal(d1, np.where(lv >= 0, lv, O.INT_MAX).astype(np.uint32))


def test_pagerank_closed_forms():
    n = 37
    h = np.arange(n, dtype=np.uint32)
    t = ((h + 1) % n).astype(np.uint32)            # directed cycle: rank 1/N is a fixpoint
    deg = np.ones(n, np.uint32)
    r = O.pagerank_f64(n, h, t, deg, 30, 0.85, 1.0 / n)
    assert np.allclose(r, 1.0 / n, rtol=0, atol=1e-15)
    # general graph: dense power iteration of r' = (1-d)/N + d * A^T (r/deg)
    hh, tt = random_graph(40, 160, 7)
    deg = np.bincount(hh, minlength=40).astype(np.uint32)
    P = np.zeros((40, 40))
    np.add.at(P, (tt, hh), 1.0)
    x = np.full(40, 0.3)
    for _ in range(12):
        x = 0.15 / 40 + 0.85 * (P @ np.where(deg > 0, x / np.maximum(deg, 1), 0.0))
    assert np.allclose(O.pagerank_f64(40, hh, tt, deg, 12, 0.85, 0.3), x, rtol=1e-13, atol=0)


def test_bellman_ford_equals_levels():
    g = O.config_graph("c1")
    lv = O.bfs_levels(g.n, g.deg, g.cols, 0)
    bf = O.bellman_ford_hops(g.n, g.rh, g.rt, 0, 99)
    assert np.array_equal(bf, np.where(lv >= 0, lv, O.INT_MAX - 2).astype(np.uint32))


def test_degree_rank_tie_rule():
    """Descending (key, id); key = (unsigned)(deg + rand()/RAND_MAX): equal
    degrees go to the higher original id first unless the float sum rounds up."""
    deg = np.array([5, 3, 5, 7, 3, 3, 0, 5], np.uint32)
    r = O.degree_rank(deg)
    rv = O.glibc_rand(1, len(deg))
    key = (deg.astype(np.float32) + rv.astype(np.float32) / np.float32(2**31)).astype(np.uint32)
    order = sorted(range(len(deg)), key=lambda v: (int(key[v]), v), reverse=True)
    assert [int(r[v]) for v in order] == list(range(len(deg)))
    assert r[3] == 0 and r[6] == len(deg) - 1


def test_rmat_invariants():
    h, t = O.rmat_edges(12, 4096, 5, 80000)
    k = (h.astype(np.uint64) << 32) | t
    assert np.all(np.diff(k) > 0)                  # sorted, no duplicates
    assert np.all(h != t) and h.max() < 4096 and t.max() < 4096
    # at every bit level, the quadrant frequencies of raw samples follow
    # (a, b, c, d), and levels are independent: (a+b)^2 of the samples take
    # head bit 0 at both the top level and the next one
    hs, ts = np.empty(20000, np.int64), np.empty(20000, np.int64)
    for i in range(20000):
        hs[i], ts[i] = O.rmat_sample(5, i, 12)
    for level in range(12):
        hb, tb = (hs >> level) & 1, (ts >> level) & 1
        q = np.bincount(hb * 2 + tb, minlength=4) / len(hs)
        assert np.allclose(q, [0.57, 0.19, 0.19, 0.05], atol=0.015), (level, q)
    both = np.mean(((hs >> 11) & 1 == 0) & ((hs >> 10) & 1 == 0))
    assert abs(both - 0.76 ** 2) < 0.015


def test_hybrid_avx512_and_one_lane_sweeps_agree():
    """The oracle's dense sweep restated with AVX-512 (apps/bfs_simd/bfs.cpp:
    181-245) and one lane at a time give the same levels and direction trace
    (both equal the textbook BFS) and valid parents."""
    g = O.config_graph("c1")
    width = O.lib().orc_simd_width()
    try:
        for src in (0, 9, 4097):
            ref = O.bfs_levels(g.n, g.deg, g.cols, src)
            got = []
            for simd in (1, 0):
                O.lib().orc_set_simd(simd)
                lv, pa, tr = O.bfs_hybrid(g, src, nthreads=4)
                assert np.array_equal(lv, ref)
                assert O.check_parents(g.n, g.deg, g.cols, src, lv, pa) == 0
                got.append(tr)
            assert np.array_equal(got[0], got[1])
            assert any(r[0] == 1 for r in got[0])      # the sweep ran
    finally:
        O.lib().orc_set_simd(1 if width == 512 else 0)
