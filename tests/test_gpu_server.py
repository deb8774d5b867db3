"""Serving loop (gphi_server_*): queries on several resident graphs submitted
together are answered with batched launches that overlap across graphs, and
each result equals the query's result alone (oracle / single-source run)."""
import threading
import time

import numpy as np
import pytest

import oracle as O
import bfs_simd_benchmarks_amd as G
from bfs_simd_benchmarks_amd.workloads import pick_sources

pytestmark = pytest.mark.gpu


def lane(g, weighted=False):
    ctx = G.Context(0)
    return ctx.graph_from_arrays(g.n, g.deg, W=g.W, S=g.S, csr_edges=g.cols,
                                 csr_weights=g.cw if weighted else None, tile_heads=g.sh,
                                 tile_tails=g.st, tile_weights=g.sw if weighted else None,
                                 tile_offsets=g.soff)


@pytest.mark.parametrize("max_batch", [64, 1])
def test_mixed_queries_equal_alone(c1, c3, max_batch):
    print(f"case: server max_batch={max_batch}", flush=True)
    n2, h2, t2 = O.config_edges("c1")
    g2 = O.build_from_edgelist(n2, h2[::3], t2[::3], 512, 8)     # a third graph, other tiling
    srv = G.Server(max_batch=max_batch)
    try:
        ids = [srv.add_graph(lane(c1)), srv.add_graph(lane(c3, weighted=True)), srv.add_graph(lane(g2))]
        graphs = [c1, c3, g2]
        rng = np.random.default_rng(5)
        jobs = []
        for q in range(90):
            gi = int(rng.integers(0, 3))
            kind = G.QUERY_SSSP if (gi == 1 and q % 3 == 0) else G.QUERY_BFS
            src = int(pick_sources(graphs[gi].deg, 1 + q % 7, q)[-1]) if q % 5 else int(rng.integers(0, graphs[gi].n))
            out = np.empty(graphs[gi].n, np.int32 if kind == G.QUERY_BFS else np.uint32)
            jobs.append((gi, kind, src, out))
        tickets = []
        lock = threading.Lock()

        def submit(part):
            for gi, kind, src, out in part:
                t = srv.submit(ids[gi], kind, src, out)
                with lock:
                    tickets.append(t)
        ths = [threading.Thread(target=submit, args=(jobs[i::3],)) for i in range(3)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        infos = [srv.wait(t) for t in tickets]
        for gi, kind, src, out in jobs:
            g = graphs[gi]
            if kind == G.QUERY_BFS:
                assert np.array_equal(out, O.bfs_levels(g.n, g.deg, g.cols, src)), (gi, src)
            else:
                assert np.array_equal(out, O.sssp_dijkstra(g.n, g.deg, g.cols, g.cw, src)), (gi, src)
        launches = sum(srv.stats(i)["launches"] for i in ids)
        served = sum(srv.stats(i)["served"] for i in ids)
        assert served == len(jobs)
        if max_batch == 1:
            assert launches == len(jobs)
            assert max(i.batch for i in infos) == 1
        else:
            assert launches < len(jobs) and max(i.batch for i in infos) > 1
    finally:
        srv.close()


def test_admission_by_hbm_budget(c1):
    srv = G.Server(max_batch=64, hbm_budget=1 << 20)       # 1 MiB: nothing fits
    gg = lane(c1)
    try:
        with pytest.raises(G.GphiError) as ei:
            srv.add_graph(gg)
        assert ei.value.code == -4
    finally:
        gg.free()
        gg.ctx.close()
        srv.close()


def test_view_runs_beside_its_graph(c3):
    """Two handles on one resident graph (gphi_graph_view), each on its own
    stream, traverse at the same time and both match the oracle."""
    import threading
    base_ctx, other = G.Context(0), G.Context(0)
    g = base_ctx.graph_from_arrays(c3.n, c3.deg, W=c3.W, S=c3.S, csr_edges=c3.cols, csr_weights=c3.cw,
                                   tile_heads=c3.sh, tile_tails=c3.st, tile_weights=c3.sw,
                                   tile_offsets=c3.soff)
    v = other.view(g)
    try:
        srcs = [int(s) for s in pick_sources(c3.deg, 6, 21)]
        got = {}

        def run_bfs():
            got["bfs"] = [g.bfs(s)[0] for s in srcs]

        def run_sssp():
            got["sssp"] = [v.sssp(s)[0] for s in srcs]
        ths = [threading.Thread(target=run_bfs), threading.Thread(target=run_sssp)]
        print("case: view sssp beside base bfs", flush=True)
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        for s, lv, d in zip(srcs, got["bfs"], got["sssp"]):
            assert np.array_equal(lv, O.bfs_levels(c3.n, c3.deg, c3.cols, s))
            assert np.array_equal(d, O.sssp_dijkstra(c3.n, c3.deg, c3.cols, c3.cw, s))
        with pytest.raises(G.GphiError) as ei:
            v.retile(256, 4)
        assert ei.value.code == -7
    finally:
        v.free()
        g.free()
        other.close()
        base_ctx.close()


def test_query_joins_running_traversal(c1):
    """Continuous batching: BFS queries submitted while a multi-source
    traversal runs join it at a level boundary (same launch, joined_level > 0)
    and still get exactly their own BFS levels.

    Graph: the rmat16 graph plus a chain of 200 vertices hanging off vertex 0
    (so a traversal that reaches vertex 0 lasts 200 more levels) and a
    separate chain of 20,000 vertices whose first vertex, as a lone plain BFS,
    keeps the lane busy while the first wave of queries queues up."""
    base, chain_b, chain_a = c1.n, 200, 20_000
    n = base + chain_b + chain_a
    h = [c1.csr_heads, np.array([0], np.uint32), np.arange(base, base + chain_b - 1, dtype=np.uint32),
         np.arange(base + chain_b, n - 1, dtype=np.uint32)]
    t = [c1.cols, np.array([base], np.uint32), np.arange(base + 1, base + chain_b, dtype=np.uint32),
         np.arange(base + chain_b + 1, n, dtype=np.uint32)]
    g = O.tile_edgelist(n, np.concatenate(h), np.concatenate(t), 1024, 16)
    wave = [int(s) for s in pick_sources(c1.deg, 16, 77)]
    srv = G.Server(max_batch=64)
    try:
        gid = srv.add_graph(lane(g))
        outs = {}

        def submit(src):
            outs[src] = np.empty(n, np.int32)
            return srv.submit(gid, G.QUERY_BFS, src, outs[src])
        print("case: server join, blocker then two waves", flush=True)
        blocker = submit(base + chain_b)
        time.sleep(0.05)
        first = [submit(s) for s in wave[:8]]
        srv.wait(blocker)
        second = [submit(s) for s in wave[8:]]
        info1 = [srv.wait(x) for x in first]
        info2 = [srv.wait(x) for x in second]
    finally:
        srv.close()
    for s in [base + chain_b] + wave:
        assert np.array_equal(outs[s], O.bfs_levels(n, g.deg, g.cols, s)), s
    launch = info1[0].launch
    assert all(i.launch == launch for i in info1), [i.launch for i in info1]
    joined = [i for i in info2 if i.launch == launch and i.joined_level > 0]
    print(f"joined the running launch: {len(joined)} of {len(info2)}, at levels "
          f"{sorted(i.joined_level for i in joined)}", flush=True)
    assert joined, [(i.launch, i.joined_level) for i in info2]


def test_sssp_queries_share_a_traversal(c3):
    """SSSP queries that pile up on a lane run as one multi-source SSSP
    traversal; every answer equals Dijkstra's."""
    print("case: server SSSP batches", flush=True)
    srv = G.Server(max_batch=64, sssp_lanes=1)
    try:
        gid = srv.add_graph(lane(c3, weighted=True))
        srcs = [int(s) for s in pick_sources(c3.deg, 40, 91)]
        outs = [np.empty(c3.n, np.uint32) for _ in srcs]
        tickets = [srv.submit(gid, G.QUERY_SSSP, s, o) for s, o in zip(srcs, outs)]
        infos = [srv.wait(t) for t in tickets]
        for s, o in zip(srcs, outs):
            assert np.array_equal(o, O.sssp_dijkstra(c3.n, c3.deg, c3.cols, c3.cw, s)), s
        assert max(i.batch for i in infos) > 1
        st = srv.stats(gid)
        assert st["served"] == len(srcs) and st["launches"] < len(srcs)
    finally:
        srv.close()
