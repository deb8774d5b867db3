"""gphi_bfs_noalloc (the BFS entry point that allocates nothing) through the
C ABI: its device-resident levels and parents against gphi_bfs on the same
graph and source, on two calls in a row, the second on a smaller graph; and
the size limit of its fixed scratch array.  Output buffers are allocated by
the test itself (hipMalloc of the HIP runtime libgphi is linked against)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import bfs_simd_benchmarks_amd as G

pytestmark = pytest.mark.gpu

HIP_D2H = 2


def hip_runtime():
    """The HIP runtime libgphi.so is linked against (same loaded object)."""
    G.load_library()
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


class DevBuf:
    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(4 * n, 4)) == 0
        assert hip.hipMemset(self.p, 0x5A, max(4 * n, 4)) == 0   # no stale zeros

    def get(self):
        out = np.empty(self.n, np.int32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, 4 * self.n, HIP_D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def upload(ctx, g):
    return ctx.graph_from_arrays(g.n, g.deg, W=g.W, S=g.S, csr_edges=g.cols, tile_heads=g.sh,
                                 tile_tails=g.st, tile_offsets=g.soff)


def noalloc(gg, src, lv, pa):
    st = G.Stats()
    gg.ctx.check(gg.ctx.lib.gphi_bfs_noalloc(gg.h, src, lv.p if lv else None, pa.p if pa else None,
                                             C.byref(st)))
    return st


def check_against_bfs(ctx, hip, g, src):
    gg = upload(ctx, g)
    lv, pa = DevBuf(hip, g.n), DevBuf(hip, g.n)
    try:
        st = noalloc(gg, src, lv, pa)
        got_lv, got_pa = lv.get(), pa.get()
        want_lv, _, want_st = gg.bfs(src)
    finally:
        lv.free()
        pa.free()
        gg.free()
    assert np.array_equal(got_lv, want_lv), f"levels differ at {np.flatnonzero(got_lv != want_lv)[:10]}"
    assert np.array_equal(got_lv, O.bfs_levels(g.n, g.deg, g.cols, src))
    assert O.check_parents(g.n, g.deg, g.cols, src, got_lv, got_pa) == 0
    assert np.array_equal(st.trace(), want_st.trace())
    assert st.reached == want_st.reached and st.edges_traversed == want_st.edges_traversed


def test_noalloc_matches_bfs_twice_second_smaller(ctx, c1):
    hip = hip_runtime()
    print("case: noalloc c1 src=0", flush=True)
    check_against_bfs(ctx, hip, c1, 0)
    n, h, t = O.config_edges("c1")
    keep = (h < 5000) & (t < 5000)
    small = O.tile_edgelist(5000, h[keep], t[keep], 256, 4)
    print("case: noalloc 5000-vertex subgraph src=1", flush=True)
    check_against_bfs(ctx, hip, small, 1)


def test_noalloc_null_outputs_and_repeat(ctx, c1):
    hip = hip_runtime()
    gg = upload(ctx, c1)
    lv = DevBuf(hip, c1.n)
    try:
        print("case: noalloc c1 no outputs, then levels only", flush=True)
        st0 = noalloc(gg, 7, None, None)
        noalloc(gg, 7, lv, None)
        got = lv.get()
    finally:
        lv.free()
        gg.free()
    assert np.array_equal(got, O.bfs_levels(c1.n, c1.deg, c1.cols, 7))
    assert st0.reached == int((got >= 0).sum())


def test_noalloc_rejects_graph_beyond_scratch(ctx):
    print("case: noalloc scale-23 graph (too large)", flush=True)
    gg = ctx.generate(scale=23, nverts=1 << 23, seed=9, samples=1 << 22, W=4096, S=16)
    try:
        st = G.Stats()
        rc = ctx.lib.gphi_bfs_noalloc(gg.h, 0, None, None, C.byref(st))
        assert rc == -3, rc
        assert b"scratch" in ctx.lib.gphi_last_error(ctx.h)
        lv, _, _ = gg.bfs(0, parents=False)     # the allocating entry point still works
        assert lv[0] == 0
    finally:
        gg.free()
