"""Python host binding of libgphi, the MI355X-native tiled graph-traversal library.

The product is the C ABI in ``include/gphi.h`` (HIP kernels for gfx950 behind
``extern "C"`` entry points).  This module is a thin ctypes layer over it that
mirrors the reference apps' roles:

=====================  ==========================================================
reference              here
=====================  ==========================================================
graph_input            :meth:`Context.load_text` (apps/bfs_simd/bfs.cpp:968)
graph_prepare          :meth:`Graph.bfs` (apps/bfs_simd/bfs.cpp:753)
sssp_weighted          :meth:`Graph.sssp` (apps/sssp_simd/sssp.cpp:969)
page_rank              :meth:`Graph.pagerank` (apps/pageRank_simd/page_rank.cpp:362)
bellman-ford page_rank :meth:`Graph.bellman_ford` (apps/bellman-ford_simd/bellman-ford.cpp:291)
=====================  ==========================================================

There is no CPU fallback: if ``libgphi.so`` is missing or no GPU is visible,
every call raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libgphi.so"
MAX_STEPS = 512
TILES_COL_STRIPE = 0
TILES_ROW_MAJOR = 1

_ERRORS = {0: "OK", -1: "E_IO", -2: "E_FORMAT", -3: "E_RANGE", -4: "E_OOM", -5: "E_HIP",
           -6: "E_ARG", -7: "E_STATE"}


class GphiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.msg = msg


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("t_ratio", C.c_uint32)]


class GraphDesc(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64), ("tile_width", C.c_uint32),
                ("stripe_length", C.c_uint32), ("tile_order", C.c_int),
                ("degrees", C.c_void_p), ("csr_edges", C.c_void_p), ("csr_weights", C.c_void_p),
                ("tile_heads", C.c_void_p), ("tile_tails", C.c_void_p),
                ("tile_weights", C.c_void_p), ("tile_offsets", C.c_void_p)]


class RmatSpec(C.Structure):
    _fields_ = [("scale", C.c_uint32), ("nverts", C.c_uint32), ("seed", C.c_uint64),
                ("samples", C.c_uint64), ("target_unique", C.c_uint64), ("a", C.c_double),
                ("b", C.c_double), ("c", C.c_double), ("perm_seed", C.c_uint64)]


class WeightSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("lo", C.c_uint32), ("hi", C.c_uint32)]


class ServerOpts(C.Structure):
    _fields_ = [("max_batch", C.c_uint32), ("hbm_budget", C.c_uint64), ("sssp_lanes", C.c_uint32)]


class QueryInfo(C.Structure):
    _fields_ = [("latency_us", C.c_double), ("queue_us", C.c_double), ("run_us", C.c_double),
                ("batch", C.c_uint32), ("joined_level", C.c_uint32), ("launch", C.c_uint64)]


QUERY_BFS = 0
QUERY_SSSP = 1

REORDER_NONE = 0
REORDER_DEGREE = 1
REORDER_BFS = 2


class Stats(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("dir", C.c_uint32 * MAX_STEPS),
                ("fsize", C.c_uint32 * MAX_STEPS), ("fdeg", C.c_uint32 * MAX_STEPS),
                ("sweep_ms", C.c_float * MAX_STEPS), ("select_ms", C.c_float * MAX_STEPS),
                ("update_ms", C.c_float * MAX_STEPS),
                ("chunks", C.c_uint32 * MAX_STEPS),
                ("init_ms", C.c_float), ("total_ms", C.c_float), ("reached", C.c_uint64),
                ("edges_traversed", C.c_uint64)]

    def trace(self) -> np.ndarray:
        """(steps, 3) array of (dir, |F|, sum outdeg F) per expanded frontier."""
        k = self.steps
        return np.stack([np.ctypeslib.as_array(self.dir)[:k], np.ctypeslib.as_array(self.fsize)[:k],
                         np.ctypeslib.as_array(self.fdeg)[:k]], axis=1).astype(np.uint32)

    def sweep_times(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.sweep_ms)[: self.steps].copy()

    def select_times(self) -> np.ndarray:
        """Chunk selection time of each dense step (0 for sparse steps)."""
        return np.ctypeslib.as_array(self.select_ms)[: self.steps].copy()

    def update_times(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.update_ms)[: self.steps].copy()

    def swept_chunks(self) -> np.ndarray:
        """2048-edge chunks each dense step swept (0 for sparse steps)."""
        return np.ctypeslib.as_array(self.chunks)[: self.steps].copy()


_lib = None
_P = C.c_void_p


def _declare(lib):
    u32, u64, i32 = C.c_uint32, C.c_uint64, C.c_int
    sig = {
        "gphi_version": (C.c_char_p, []),
        "gphi_open": (i32, [C.POINTER(Opts), C.POINTER(_P)]),
        "gphi_close": (None, [_P]),
        "gphi_set_level_timing": (i32, [_P, i32]),
        "gphi_last_error": (C.c_char_p, [_P]),
        "gphi_graph_from_desc": (i32, [_P, C.POINTER(GraphDesc), C.POINTER(_P)]),
        "gphi_graph_load_text": (i32, [_P, C.c_char_p, u32, u32, i32, C.POINTER(_P)]),
        "gphi_graph_load_text_rowmajor": (i32, [_P, C.c_char_p, u32, u32, C.POINTER(_P)]),
        "gphi_graph_info": (i32, [_P, _P, _P, _P, _P, _P]),
        "gphi_graph_device_bytes": (u64, [_P]),
        "gphi_graph_free": (None, [_P]),
        "gphi_bfs": (i32, [_P, u32, _P, _P, C.POINTER(Stats)]),
        "gphi_sssp": (i32, [_P, u32, _P, C.POINTER(Stats)]),
        "gphi_bfs_batch": (i32, [_P, _P, u32, _P, C.POINTER(Stats)]),
        "gphi_sssp_batch": (i32, [_P, _P, u32, _P, C.POINTER(Stats)]),
        "gphi_pagerank": (i32, [_P, u32, C.c_float, C.c_float, _P, _P]),
        "gphi_bellman_ford": (i32, [_P, u32, u32, _P, _P, _P]),
        "gphi_bfs_alg_bytes": (i32, [_P, _P, C.POINTER(Stats), _P]),
        "gphi_graph_generate": (i32, [_P, C.POINTER(RmatSpec), u32, u32, i32, C.POINTER(WeightSpec),
                                      i32, C.POINTER(_P)]),
        "gphi_graph_download": (i32, [_P, _P, _P, _P, _P, _P, _P]),
        "gphi_graph_tile_offsets": (i32, [_P, _P]),
        "gphi_graph_retile": (i32, [_P, u32, u32, i32]),
        "gphi_graph_view": (i32, [_P, _P, C.POINTER(_P)]),
        "gphi_graph_save": (i32, [_P, C.c_char_p]),
        "gphi_graph_load": (i32, [_P, C.c_char_p, C.POINTER(_P)]),
        "gphi_host_read_text": (i32, [C.c_char_p, u32, u32, i32, i32, C.POINTER(_P)]),
        "gphi_host_from_device": (i32, [_P, C.POINTER(_P)]),
        "gphi_host_write_text": (i32, [_P, C.c_char_p]),
        "gphi_host_desc": (i32, [_P, C.POINTER(GraphDesc)]),
        "gphi_host_free": (None, [_P]),
        "gphi_server_create": (i32, [C.POINTER(ServerOpts), C.POINTER(_P)]),
        "gphi_server_add_graph": (i32, [_P, _P, C.POINTER(C.c_uint32)]),
        "gphi_server_submit": (i32, [_P, u32, u32, u32, _P, C.POINTER(C.c_uint64)]),
        "gphi_server_wait": (i32, [_P, C.c_uint64, C.POINTER(QueryInfo)]),
        "gphi_server_stats": (i32, [_P, u32, _P, _P, _P]),
        "gphi_server_destroy": (None, [_P]),
        "gphi_debug_fail_alloc": (None, [i32]),
        "gphi_bfs_noalloc": (i32, [_P, u32, _P, _P, C.POINTER(Stats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load_library(path: Path | str | None = None):
    """Load libgphi.so (built in-tree by ``python -m bfs_simd_benchmarks_amd.build``)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise GphiError(-7, f"{p} is not built; run python -m bfs_simd_benchmarks_amd.build")
    lib = _declare(C.CDLL(str(p)))
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _host_check(lib, rc):
    if rc:
        raise GphiError(rc, lib.gphi_last_error(None).decode())


class HostGraph:
    """graph_input's arrays in host memory (the reference's text formats, no device)."""

    def __init__(self, h):
        self.lib = load_library()
        self.h = h
        self.desc = GraphDesc()
        _host_check(self.lib, self.lib.gphi_host_desc(h, C.byref(self.desc)))

    @classmethod
    def read(cls, prefix, W: int, S: int, weighted: bool = False, row_major: bool = False):
        lib = load_library()
        h = _P()
        _host_check(lib, lib.gphi_host_read_text(str(prefix).encode(), W, S, int(weighted),
                                                  int(row_major), C.byref(h)))
        return cls(h)

    def write(self, prefix):
        _host_check(self.lib, self.lib.gphi_host_write_text(self.h, str(prefix).encode()))

    def arrays(self) -> dict:
        d = self.desc
        side = d.n // d.tile_width + (1 if d.n % d.tile_width else 0)

        def arr(ptr, count):
            if not ptr or count == 0:
                return None
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint32)), shape=(count,)).copy()
        return {"n": d.n, "m": d.m, "W": d.tile_width, "S": d.stripe_length, "order": d.tile_order,
                "deg": arr(d.degrees, d.n), "cols": arr(d.csr_edges, d.m), "cw": arr(d.csr_weights, d.m),
                "heads": arr(d.tile_heads, d.m), "tails": arr(d.tile_tails, d.m),
                "tw": arr(d.tile_weights, d.m), "offsets": arr(d.tile_offsets, side * side)}

    def free(self):
        if self.h:
            self.lib.gphi_host_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Resident graphs with continuous batching (gphi_server_*).  Each graph
    added must live in its own Context; the server takes both over."""

    def __init__(self, max_batch: int = 64, hbm_budget: int = 0, sssp_lanes: int = 0):
        self.lib = load_library()
        h = _P()
        _host_check(self.lib, self.lib.gphi_server_create(
            C.byref(ServerOpts(max_batch, hbm_budget, sssp_lanes)), C.byref(h)))
        self.h = h
        self.graphs = []
        self._n = {}          # graph id -> vertex count
        self._outs = {}       # ticket -> output array the worker thread writes into

    def add_graph(self, g: "Graph") -> int:
        gid = C.c_uint32()
        _host_check(self.lib, self.lib.gphi_server_add_graph(self.h, g.h, C.byref(gid)))
        g.h = None                      # owned by the server now
        g.ctx.h = None
        self.graphs.append(g)
        self._n[gid.value] = g.n
        return gid.value

    def submit(self, gid: int, kind: int, src: int, out: np.ndarray | None = None) -> int:
        """Queue a query.  `out` (optional) receives the result from a worker
        thread: a C-contiguous int32 array (BFS levels) or uint32 array (SSSP
        distances) of at least the graph's vertex count, kept alive here until
        wait() returns for the ticket."""
        if out is not None:
            want = np.int32 if kind == QUERY_BFS else np.uint32
            n = self._n.get(gid)
            if n is None:
                raise GphiError(-3, f"no graph {gid}")
            if (not isinstance(out, np.ndarray) or out.dtype != want or not out.flags.c_contiguous
                    or not out.flags.writeable or out.size < n):
                raise GphiError(-6, f"out must be a writeable C-contiguous {np.dtype(want).name} array "
                                    f"of at least {n} elements")
        t = C.c_uint64()
        _host_check(self.lib, self.lib.gphi_server_submit(self.h, gid, kind, src, _ptr(out), C.byref(t)))
        if out is not None:
            self._outs[t.value] = out
        return t.value

    def wait(self, ticket: int) -> QueryInfo:
        info = QueryInfo()
        try:
            _host_check(self.lib, self.lib.gphi_server_wait(self.h, ticket, C.byref(info)))
        finally:
            self._outs.pop(ticket, None)
        return info

    def stats(self, gid: int):
        b, s_, m = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _host_check(self.lib, self.lib.gphi_server_stats(self.h, gid, C.byref(b), C.byref(s_), C.byref(m)))
        return {"launches": b.value, "served": s_.value, "hbm_bytes": m.value}

    def close(self):
        if self.h:
            self.lib.gphi_server_destroy(self.h)
            self.h = None


class Context:
    """One HIP device + stream (gphi_open)."""

    def __init__(self, device: int = 0, t_ratio: int = 20):
        self.lib = load_library()
        h = _P()
        rc = self.lib.gphi_open(C.byref(Opts(device, t_ratio)), C.byref(h))
        if rc:
            raise GphiError(rc, self.lib.gphi_last_error(None).decode())
        self.h = h
        self.device = device

    def check(self, rc: int):
        if rc:
            raise GphiError(rc, self.lib.gphi_last_error(self.h).decode())

    def set_level_timing(self, on: bool):
        """Per-level HIP-event timing of traversals (Stats sweep/select/update
        times); on by default.  Off, only total_ms is recorded and the levels'
        kernels run back to back."""
        self.check(self.lib.gphi_set_level_timing(self.h, 1 if on else 0))

    def close(self):
        if self.h:
            self.lib.gphi_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def graph_from_arrays(self, n: int, degrees, *, W: int, S: int, csr_edges=None,
                          csr_weights=None, tile_heads=None, tile_tails=None, tile_weights=None,
                          tile_offsets=None, tile_order: int = TILES_COL_STRIPE) -> "Graph":
        """Upload graph_input's arrays (apps/bfs_simd/bfs.cpp:968-1118)."""
        keep = {k: (None if v is None else _u32(v)) for k, v in dict(
            degrees=degrees, csr_edges=csr_edges, csr_weights=csr_weights, tile_heads=tile_heads,
            tile_tails=tile_tails, tile_weights=tile_weights, tile_offsets=tile_offsets).items()}
        m = len(keep["csr_edges"]) if keep["csr_edges"] is not None else len(keep["tile_heads"])
        d = GraphDesc(n, m, W, S, tile_order, *(_ptr(keep[k]) for k in (
            "degrees", "csr_edges", "csr_weights", "tile_heads", "tile_tails", "tile_weights",
            "tile_offsets")))
        h = _P()
        self.check(self.lib.gphi_graph_from_desc(self.h, C.byref(d), C.byref(h)))
        return Graph(self, h)

    def generate(self, *, scale: int, nverts: int, seed: int, samples: int, target_unique: int = 0,
                 abc=(0.57, 0.19, 0.19), perm_seed: int = 0, W: int, S: int,
                 reorder: int = REORDER_DEGREE, weight_seed: int = 0, wlo: int = 1, whi: int = 255,
                 tile_order: int = TILES_COL_STRIPE) -> "Graph":
        """Generate an RMAT edge list on the device and build the graph in HBM, as the
        reference tools would lay it out (gphi_graph_generate)."""
        spec = RmatSpec(scale, nverts, seed, samples, target_unique, *abc, perm_seed)
        ws = WeightSpec(weight_seed, wlo, whi)
        h = _P()
        self.check(self.lib.gphi_graph_generate(self.h, C.byref(spec), W, S, reorder, C.byref(ws),
                                                tile_order, C.byref(h)))
        return Graph(self, h)

    def load_text(self, prefix: str, W: int, S: int, weighted: bool = False) -> "Graph":
        """prefix_col-S-coo-tiled-W-* + prefix_untiled-* (apps/bfs_simd/bfs.cpp:983-1113)."""
        h = _P()
        self.check(self.lib.gphi_graph_load_text(self.h, str(prefix).encode(), W, S, int(weighted),
                                                 C.byref(h)))
        return Graph(self, h)

    def view(self, graph: "Graph") -> "Graph":
        """A handle on `graph` (resident on this context's device) with its own
        traversal state, running on this context's stream (gphi_graph_view).
        Free it before the graph."""
        h = _P()
        self.check(self.lib.gphi_graph_view(self.h, graph.h, C.byref(h)))
        return Graph(self, h)

    def load_binary(self, path) -> "Graph":
        """A graph cache file written by Graph.save (gphi_graph_load)."""
        h = _P()
        self.check(self.lib.gphi_graph_load(self.h, str(path).encode(), C.byref(h)))
        return Graph(self, h)

    def load_text_rowmajor(self, prefix: str, W: int, S: int) -> "Graph":
        """prefix_coo-tiled-W-* (apps/pageRank_simd/page_rank.cpp:468-568)."""
        h = _P()
        self.check(self.lib.gphi_graph_load_text_rowmajor(self.h, str(prefix).encode(), W, S,
                                                          C.byref(h)))
        return Graph(self, h)


class Graph:
    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h
        n, m, W, S, wt = C.c_uint32(), C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_int()
        ctx.check(ctx.lib.gphi_graph_info(h, C.byref(n), C.byref(m), C.byref(W), C.byref(S),
                                          C.byref(wt)))
        self.n, self.m, self.W, self.S, self.weighted = n.value, m.value, W.value, S.value, bool(wt.value)

    @property
    def device_bytes(self) -> int:
        return int(self.ctx.lib.gphi_graph_device_bytes(self.h))

    def free(self):
        if self.h:
            self.ctx.lib.gphi_graph_free(self.h)
            self.h = None

    def bfs(self, src: int, parents: bool = True, levels: bool = True):
        """Hybrid BFS (graph_prepare).  Returns (levels int32|None, parents int32|None, Stats).
        With levels=parents=False the results stay in HBM (no device-to-host copy)."""
        lv = np.empty(self.n, np.int32) if levels else None
        pa = np.empty(self.n, np.int32) if parents else None
        st = Stats()
        self.ctx.check(self.ctx.lib.gphi_bfs(self.h, src, _ptr(lv), _ptr(pa), C.byref(st)))
        return lv, pa, st

    def bfs_batch(self, srcs, levels: bool = True):
        """Multi-source BFS of up to 256 sources in one traversal.  Returns
        (levels uint8[k, n] or None, Stats); 255 = unreached."""
        srcs = _u32(srcs)
        k = len(srcs)
        lv = np.empty((k, self.n), np.uint8) if levels else None
        st = Stats()
        self.ctx.check(self.ctx.lib.gphi_bfs_batch(self.h, _ptr(srcs), k, _ptr(lv), C.byref(st)))
        return lv, st

    def sssp(self, src: int):
        d = np.empty(self.n, np.uint32)
        st = Stats()
        self.ctx.check(self.ctx.lib.gphi_sssp(self.h, src, _ptr(d), C.byref(st)))
        return d, st

    def sssp_batch(self, srcs, dist: bool = True):
        """Weighted SSSP from up to 16 sources in one traversal.  Returns
        (dist uint32[k, n] or None, Stats); 2147483647 = unreached."""
        srcs = _u32(srcs)
        k = len(srcs)
        d = np.empty((k, self.n), np.uint32) if dist else None
        st = Stats()
        self.ctx.check(self.ctx.lib.gphi_sssp_batch(self.h, _ptr(srcs), k, _ptr(d), C.byref(st)))
        return d, st

    def pagerank(self, iters: int = 1, damping: float = 0.85, init: float = 1.0):
        r = np.empty(self.n, np.float32)
        ms = C.c_float()
        self.ctx.check(self.ctx.lib.gphi_pagerank(self.h, iters, damping, init, _ptr(r), C.byref(ms)))
        return r, ms.value

    def bellman_ford(self, src: int = 0, max_sweeps: int = 99):
        d = np.empty(self.n, np.uint32)
        run, ms = C.c_uint32(), C.c_float()
        self.ctx.check(self.ctx.lib.gphi_bellman_ford(self.h, src, max_sweeps, _ptr(d), C.byref(run),
                                                      C.byref(ms)))
        return d, run.value, ms.value

    def download(self, csr: bool = True, tiles: bool = True) -> dict:
        """Host copies of the graph arrays (degrees, CSR, tile order; weights if any)."""
        out = {"deg": np.empty(self.n, np.uint32)}
        if csr:
            out["cols"] = np.empty(self.m, np.uint32)
            out["cw"] = np.empty(self.m, np.uint32) if self.weighted else None
        if tiles:
            out["heads"] = np.empty(self.m, np.uint32)
            out["tails"] = np.empty(self.m, np.uint32)
            out["tw"] = np.empty(self.m, np.uint32) if self.weighted else None
        g = out.get
        self.ctx.check(self.ctx.lib.gphi_graph_download(
            self.h, _ptr(g("deg")), _ptr(g("cols")), _ptr(g("cw")), _ptr(g("heads")), _ptr(g("tails")),
            _ptr(g("tw"))))
        return out

    def retile(self, W: int, S: int, tile_order: int = TILES_COL_STRIPE):
        """Rebuild the tile arrays from the CSR with another (W, S) or order."""
        self.ctx.check(self.ctx.lib.gphi_graph_retile(self.h, W, S, tile_order))
        self.W, self.S = W, S

    def save(self, path):
        """Write the binary cache file of this graph (gphi_graph_save)."""
        self.ctx.check(self.ctx.lib.gphi_graph_save(self.h, str(path).encode()))

    def to_host(self) -> HostGraph:
        """Download into a HostGraph (writable as the reference's text files)."""
        h = _P()
        self.ctx.check(self.ctx.lib.gphi_host_from_device(self.h, C.byref(h)))
        return HostGraph(h)

    @property
    def side(self) -> int:
        return self.n // self.W + (1 if self.n % self.W else 0)

    def tile_offsets(self) -> np.ndarray:
        """The -offsets array of the stored tile order (SIDE^2 entries)."""
        out = np.empty(self.side * self.side, np.uint32)
        self.ctx.check(self.ctx.lib.gphi_graph_tile_offsets(self.h, _ptr(out)))
        return out

    def alg_bytes(self, levels: np.ndarray, st: Stats) -> np.ndarray:
        out = np.zeros(max(st.steps, 1), np.float64)
        lv = np.ascontiguousarray(levels, dtype=np.int32)
        self.ctx.check(self.ctx.lib.gphi_bfs_alg_bytes(self.h, _ptr(lv), C.byref(st), _ptr(out)))
        return out[: st.steps]


__all__ = ["Context", "Graph", "HostGraph", "Server", "QUERY_BFS", "QUERY_SSSP", "GphiError", "Stats", "load_library", "TILES_COL_STRIPE",
           "TILES_ROW_MAJOR", "LIB_PATH"]
