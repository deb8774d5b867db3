"""Python binding of the oracle (CPU restatement of the reference path).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the cpu_baseline leg of ``bench.py`` -- as the checker, never as the thing
measured or shipped.  See oracle.h for what pins each function.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"
REF_DIR = ORACLE_DIR / "_ref"

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "lib"], check=True)
        L = C.CDLL(str(LIB_PATH))
        P, u32, u64, i32, f64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
        sig = {
            "orc_rmat_thresholds": (None, [f64, f64, f64, P]),
            "orc_rmat_sample": (None, [u64, u64, u32, P, P, P]),
            "orc_rmat_edges": (u64, [u32, u32, u64, u64, u64, f64, f64, f64, P, P, u64]),
            "orc_random_perm": (None, [u32, u64, P]),
            "orc_weights": (None, [u64, u64, u32, u32, P]),
            "orc_relabel": (None, [u64, P, P, P]),
            "orc_glibc_rand": (None, [u32, u64, P]),
            "orc_degree_rank": (None, [u32, P, P]),
            "orc_bfs_order_rank": (None, [u32, P, P, P]),
            "orc_untiled": (None, [u32, u64, P, P, P, P, P, P, P]),
            "orc_side": (u32, [u32, u32]),
            "orc_coo_tile": (None, [u32, u64, u32, P, P, P, P, P, P, P]),
            "orc_col_major": (None, [u32, u64, u32, u32, P, P, P, P, P, P, P, P]),
            "orc_write_parts": (i32, [C.c_char_p, u32, u64, P, P, P, u32]),
            "orc_write_u32_lines": (i32, [C.c_char_p, P, u64]),
            "orc_write_edgelist": (i32, [C.c_char_p, u32, u64, P, P, P]),
            "orc_bfs_levels": (None, [u32, P, P, u32, P]),
            "orc_bfs_hybrid": (i32, [u32, u64, u32, u32, P, P, P, P, P, u32, u32, i32, P, P, P, u32]),
            "orc_check_parents": (C.c_int64, [u32, P, P, u32, P, P]),
            "orc_simd_width": (i32, []),
            "orc_set_simd": (None, [i32]),
            "orc_sssp_dijkstra": (None, [u32, P, P, P, u32, P]),
            "orc_pagerank_f64": (None, [u32, u64, P, P, P, u32, f64, f64, P]),
            "orc_bellman_ford_hops": (None, [u32, u64, P, P, u32, u32, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


INT_MAX = 2147483647
RMAT_ABC = (0.57, 0.19, 0.19)


def rmat_edges(scale, nverts, seed, nsamples, target_unique=0, abc=RMAT_ABC):
    cap = target_unique if target_unique else nsamples
    h = np.empty(max(cap, 1), np.uint32)
    t = np.empty(max(cap, 1), np.uint32)
    m = lib().orc_rmat_edges(scale, nverts, seed, nsamples, target_unique, *abc, _p(h), _p(t), cap)
    return h[:m].copy(), t[:m].copy()


def rmat_sample(seed, i, scale, abc=RMAT_ABC):
    thr = np.zeros(3, np.uint32)
    lib().orc_rmat_thresholds(*abc, _p(thr))
    h, t = C.c_uint32(), C.c_uint32()
    lib().orc_rmat_sample(seed, i, scale, _p(thr), C.byref(h), C.byref(t))
    return h.value, t.value


def random_perm(n, seed):
    p = np.empty(n, np.uint32)
    lib().orc_random_perm(n, seed, _p(p))
    return p


def weights(m, seed, lo=1, hi=255):
    w = np.empty(max(m, 1), np.uint32)
    lib().orc_weights(m, seed, lo, hi, _p(w))
    return w[:m]


def glibc_rand(seed, n):
    out = np.empty(n, np.int32)
    lib().orc_glibc_rand(seed, n, _p(out))
    return out


def degree_rank(deg):
    deg = _u32(deg)
    r = np.empty(len(deg), np.uint32)
    lib().orc_degree_rank(len(deg), _p(deg), _p(r))
    return r


def row_offsets(deg):
    off = np.zeros(len(deg) + 1, np.uint64)
    np.cumsum(np.asarray(deg, np.uint64), out=off[1:])
    return off


def bfs_order_rank(n, deg, cols):
    off = row_offsets(deg)
    cols = _u32(cols)
    r = np.empty(n, np.uint32)
    lib().orc_bfs_order_rank(n, _p(off), _p(cols), _p(r))
    return r


def untiled(n, h, t, w=None):
    h, t = _u32(h), _u32(t)
    m = len(h)
    oh, ot = np.empty(m, np.uint32), np.empty(m, np.uint32)
    ow = np.empty(m, np.uint32) if w is not None else None
    w = _u32(w) if w is not None else None
    deg = np.empty(n, np.uint32)
    lib().orc_untiled(n, m, _p(h), _p(t), _p(w), _p(oh), _p(ot), _p(ow), _p(deg))
    return oh, ot, ow, deg


def side(n, W):
    return n // W + (1 if n % W else 0)


def coo_tile(n, W, h, t, w=None):
    h, t = _u32(h), _u32(t)
    m = len(h)
    s = side(n, W)
    oh, ot = np.empty(m, np.uint32), np.empty(m, np.uint32)
    ow = np.empty(m, np.uint32) if w is not None else None
    w = _u32(w) if w is not None else None
    off = np.empty(s * s, np.uint32)
    lib().orc_coo_tile(n, m, W, _p(h), _p(t), _p(w), _p(oh), _p(ot), _p(ow), _p(off))
    return oh, ot, ow, off


def col_major(n, W, S, h, t, w, off):
    h, t, off = _u32(h), _u32(t), _u32(off)
    m = len(h)
    oh, ot = np.empty(m, np.uint32), np.empty(m, np.uint32)
    ow = np.empty(m, np.uint32) if w is not None else None
    w = _u32(w) if w is not None else None
    noff = np.empty(len(off), np.uint32)
    lib().orc_col_major(n, m, W, S, _p(h), _p(t), _p(w), _p(off), _p(oh), _p(ot), _p(ow), _p(noff))
    return oh, ot, ow, noff


def write_parts(prefix, n, h, t, w=None, parts=64):
    h, t = _u32(h), _u32(t)
    w = _u32(w) if w is not None else None
    rc = lib().orc_write_parts(str(prefix).encode(), n, len(h), _p(h), _p(t), _p(w), parts)
    assert rc == 0, prefix


def write_u32_lines(path, a):
    a = _u32(a)
    assert lib().orc_write_u32_lines(str(path).encode(), _p(a), len(a)) == 0, path


def write_edgelist(path, n, h, t, w=None):
    h, t = _u32(h), _u32(t)
    w = _u32(w) if w is not None else None
    assert lib().orc_write_edgelist(str(path).encode(), n, len(h), _p(h), _p(t), _p(w)) == 0


def bfs_levels(n, deg, cols, src):
    off = row_offsets(deg)
    cols = _u32(cols)
    lv = np.empty(n, np.int32)
    lib().orc_bfs_levels(n, _p(off), _p(cols), src, _p(lv))
    return lv


def bfs_hybrid(g: "Graph", src, t_ratio=20, nthreads=0):
    lv = np.empty(g.n, np.int32)
    pa = np.empty(g.n, np.int32)
    cap = 4096
    tr = np.zeros(3 * cap, np.uint32)
    k = lib().orc_bfs_hybrid(g.n, g.m, g.W, g.S, _p(g.deg), _p(g.cols), _p(g.sh), _p(g.st),
                             _p(g.soff), src, t_ratio, nthreads, _p(lv), _p(pa), _p(tr), cap)
    assert k >= 0
    return lv, pa, tr[: 3 * min(k, cap)].reshape(-1, 3)


def check_parents(n, deg, cols, src, levels, parents):
    off = row_offsets(deg)
    cols = _u32(cols)
    lv = np.ascontiguousarray(levels, np.int32)
    pa = np.ascontiguousarray(parents, np.int32)
    return int(lib().orc_check_parents(n, _p(off), _p(cols), src, _p(lv), _p(pa)))


def sssp_dijkstra(n, deg, cols, wts, src):
    off = row_offsets(deg)
    cols, wts = _u32(cols), _u32(wts)
    d = np.empty(n, np.uint32)
    lib().orc_sssp_dijkstra(n, _p(off), _p(cols), _p(wts), src, _p(d))
    return d


def pagerank_f64(n, heads, tails, deg, iters, damping=0.85, init=1.0):
    heads, tails, deg = _u32(heads), _u32(tails), _u32(deg)
    r = np.empty(n, np.float64)
    lib().orc_pagerank_f64(n, len(heads), _p(heads), _p(tails), _p(deg), iters, damping, init, _p(r))
    return r


def bellman_ford_hops(n, heads, tails, src=0, sweeps=99):
    heads, tails = _u32(heads), _u32(tails)
    d = np.empty(n, np.uint32)
    lib().orc_bellman_ford_hops(n, len(heads), _p(heads), _p(tails), src, sweeps, _p(d))
    return d


# ---------------------------------------------------------------------------
# the tool pipeline, in memory: edge list -> F1 -> F4 -> F1 -> F2 -> F3
# (tools/tiling_process/process.sh:10-40 with the degree reorder of
# tools/vertex_id_reorder_to_degrees in place of vertex_id_remap)
# ---------------------------------------------------------------------------
@dataclass
class Graph:
    n: int
    m: int
    W: int
    S: int
    eh: np.ndarray              # `_reorder` edge list (1-based in files), as written
    et: np.ndarray
    ew: np.ndarray | None       # weights of that edge list
    deg: np.ndarray             # out-degrees (-nneibor)
    cols: np.ndarray            # CSR tails (_untiled order)
    cw: np.ndarray | None       # CSR weights
    rh: np.ndarray              # row-major tiles (_coo-tiled-W)
    rt: np.ndarray
    rw: np.ndarray | None
    roff: np.ndarray
    sh: np.ndarray              # column-stripe tiles (_col-S-coo-tiled-W)
    st: np.ndarray
    sw: np.ndarray | None
    soff: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def csr_heads(self):
        return np.repeat(np.arange(self.n, dtype=np.uint32), self.deg)


def build_from_edgelist(n, h, t, W, S, w=None, reorder="degree"):
    """Apply the reference tools to an edge list (0-based arrays)."""
    uh, ut, uw, udeg = untiled(n, h, t, w)
    if reorder == "degree":
        rank = degree_rank(udeg)
    elif reorder == "bfs":
        rank = bfs_order_rank(n, udeg, ut)
    else:
        rank = np.arange(n, dtype=np.uint32)
    eh, et = rank[uh], rank[ut]
    return tile_edgelist(n, eh, et, W, S, uw)


def tile_edgelist(n, eh, et, W, S, ew=None):
    ch, ct, cw, deg = untiled(n, eh, et, ew)
    rh, rt, rw, roff = coo_tile(n, W, ch, ct, cw)
    sh, st, sw, soff = col_major(n, W, S, rh, rt, rw, roff)
    return Graph(n, len(eh), W, S, eh, et, ew, deg, ct, cw, rh, rt, rw, roff, sh, st, sw, soff)


# Workloads (project spec for the synthetic inputs; the reference ships none)
CONFIGS = {
    "c1": dict(scale=16, nverts=1 << 16, seed=1, samples=16 << 16, target=0, perm=False, W=1024, S=16),
    "c2": dict(scale=21, nverts=1_632_803, seed=2, samples=1 << 27, target=30_622_564, perm=True,
               W=1024, S=16),
    "c4": dict(scale=25, nverts=1 << 25, seed=4, samples=36 << 25, target=0, perm=False, W=2048, S=32),
}


def config_edges(name):
    c = CONFIGS[name]
    h, t = rmat_edges(c["scale"], c["nverts"], c["seed"], c["samples"], c["target"])
    if c["perm"]:
        p = random_perm(c["nverts"], c["seed"])
        h, t = p[h], p[t]
    return c["nverts"], h, t


def config_graph(name, W=None, S=None, weight_seed=None):
    c = CONFIGS[name]
    n, h, t = config_edges(name)
    g = build_from_edgelist(n, h, t, W or c["W"], S or c["S"])
    if weight_seed is not None:
        w = weights(g.m, weight_seed)
        g = tile_edgelist(n, g.eh, g.et, g.W, g.S, w)
    return g
