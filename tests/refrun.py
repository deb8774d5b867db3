"""Drive the reference programs compiled from source into oracle/_ref (by
oracle/Makefile) and the oracle's restatement of the same steps, on the same
inputs, in a scratch directory.  Used by tests/test_oracle_vs_ref.py (when
oracle/_ref exists, i.e. in the build container) and by
tests/golden/make_golden.py, which freezes the reference outputs as fixtures
for machines without the reference tree."""
from __future__ import annotations

import hashlib
import subprocess
from pathlib import Path

import numpy as np

import oracle as O

REF = O.REF_DIR
PARTS = 64


def have_ref(*names) -> bool:
    return all((REF / n).exists() for n in names)


def run_ref(name: str, *args, cwd: Path, timeout: int = 600) -> str:
    r = subprocess.run([str(REF / name), *map(str, args)], cwd=cwd, capture_output=True,
                       text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{name} {args} failed: {r.stdout[-500:]} {r.stderr[-500:]}")
    return r.stdout


def files_of(prefix: Path, with_offsets: bool = False) -> list[Path]:
    out = [Path(f"{prefix}-{k}") for k in range(PARTS)] + [Path(f"{prefix}-nneibor")]
    if with_offsets:
        out.append(Path(f"{prefix}-offsets"))
    return out


def digest(paths) -> str:
    h = hashlib.sha256()
    for p in paths:
        name = Path(p).name  # hash the format suffix, not the graph's base name
        h.update((name.split("-")[-1] if "-" in name else name.rsplit("_", 1)[-1]).encode())
        h.update(Path(p).read_bytes())
    return h.hexdigest()


def same_bytes(a_paths, b_paths) -> list[str]:
    bad = []
    for a, b in zip(a_paths, b_paths):
        if Path(a).read_bytes() != Path(b).read_bytes():
            bad.append(Path(a).name)
    return bad


# ---- the oracle writing the same files the tools write ----------------------
def oracle_untiled_files(prefix: Path, n, h, t, w=None):
    uh, ut, uw, deg = O.untiled(n, h, t, w)
    O.write_parts(f"{prefix}_untiled", n, uh, ut, uw)
    O.write_u32_lines(f"{prefix}_untiled-nneibor", deg)
    return uh, ut, uw, deg


def oracle_tiled_files(prefix: Path, n, h, t, W, S, w=None):
    """prefix is an edge list in _untiled order already bucketed by head"""
    uh, ut, uw, deg = O.untiled(n, h, t, w)
    rh, rt, rw, roff = O.coo_tile(n, W, uh, ut, uw)
    O.write_parts(f"{prefix}_coo-tiled-{W}", n, rh, rt, rw)
    O.write_u32_lines(f"{prefix}_coo-tiled-{W}-offsets", roff)
    O.write_u32_lines(f"{prefix}_coo-tiled-{W}-nneibor", deg)
    sh, st, sw, soff = O.col_major(n, W, S, rh, rt, rw, roff)
    O.write_parts(f"{prefix}_col-{S}-coo-tiled-{W}", n, sh, st, sw)
    O.write_u32_lines(f"{prefix}_col-{S}-coo-tiled-{W}-offsets", soff)
    O.write_u32_lines(f"{prefix}_col-{S}-coo-tiled-{W}-nneibor", deg)


def read_levels_txt(path: Path) -> np.ndarray:
    return np.loadtxt(path, dtype=np.int64).astype(np.int32)


def pipeline(work: Path, n: int, h, t, W: int, S: int, w=None, tag="g"):
    """Run tools + oracle on one edge list; returns a dict of comparisons and
    the reference outputs.  Edge lists are 0-based numpy arrays."""
    ref, orc = work / "ref", work / "orc"
    ref.mkdir(parents=True, exist_ok=True)
    orc.mkdir(parents=True, exist_ok=True)
    res = {"mismatch": {}, "digest": {}}
    # raw edge list -> _untiled (csr_tiling)
    O.write_edgelist(ref / tag, n, h, t)
    run_ref("csr_tiling", ref / tag, cwd=ref)
    uh, ut, _, deg = oracle_untiled_files(orc / tag, n, h, t)
    res["mismatch"]["raw untiled"] = same_bytes(files_of(ref / f"{tag}_untiled"),
                                                files_of(orc / f"{tag}_untiled"))
    res["digest"]["raw untiled"] = digest(files_of(ref / f"{tag}_untiled"))
    # degree relabel (vertex_id_reorder_to_degrees)
    run_ref("degree_reorder", ref / tag, cwd=ref)
    rank = O.degree_rank(deg)
    O.write_edgelist(orc / f"{tag}_degreeReordered", n, rank[uh], rank[ut])
    res["mismatch"]["degreeReordered"] = same_bytes([ref / f"{tag}_degreeReordered"],
                                                    [orc / f"{tag}_degreeReordered"])
    res["digest"]["degreeReordered"] = digest([ref / f"{tag}_degreeReordered"])
    # BFS-order relabel (vertex_id_remap)
    run_ref("bfs_remap", ref / tag, cwd=ref)
    brank = O.bfs_order_rank(n, deg, ut)
    O.write_edgelist(orc / f"{tag}_reorder", n, brank[uh], brank[ut])
    res["mismatch"]["bfs_reorder"] = same_bytes([ref / f"{tag}_reorder"], [orc / f"{tag}_reorder"])
    res["digest"]["bfs_reorder"] = digest([ref / f"{tag}_reorder"])
    # the degree-relabelled list becomes the base "R" of the tiled formats
    eh, et = rank[uh], rank[ut]
    base = f"{tag}R"
    (ref / base).write_bytes((ref / f"{tag}_degreeReordered").read_bytes())
    run_ref("csr_tiling", ref / base, cwd=ref)
    run_ref("coo_tiling", ref / base, W, W, cwd=ref)
    run_ref("column_major_tile", ref / base, W, S, S, cwd=ref)
    oracle_untiled_files(orc / base, n, eh, et)
    oracle_tiled_files(orc / base, n, eh, et, W, S)
    for fmt, off in (("untiled", False), (f"coo-tiled-{W}", True), (f"col-{S}-coo-tiled-{W}", True)):
        res["mismatch"][fmt] = same_bytes(files_of(ref / f"{base}_{fmt}", off),
                                          files_of(orc / f"{base}_{fmt}", off))
        res["digest"][fmt] = digest(files_of(ref / f"{base}_{fmt}", off))
    # reference known answers on the degree-relabelled graph
    run_ref("sssp_naive_dbg", ref / base, cwd=ref)
    res["ref_levels0"] = read_levels_txt(ref / "distances.txt")
    run_ref("pagerank_naive_dbg", ref / base, cwd=ref)
    res["ref_rank1"] = np.loadtxt(ref / "ranks.txt", dtype=np.float64)
    out = run_ref("bellman_ford", ref / base, W, S, cwd=ref)
    lines = [ln.split() for ln in out.strip().splitlines()[-20:]]
    res["ref_bf20"] = np.array([int(x[1]) for x in lines], np.int64)
    assert [int(x[0]) for x in lines] == list(range(20))
    res["graph"] = O.tile_edgelist(n, eh, et, W, S)
    if w is not None:
        wt = f"{tag}w"
        O.write_edgelist(ref / wt, n, eh, et, w)
        run_ref("csr_tiling_w", ref / wt, cwd=ref)
        run_ref("coo_tiling_w", ref / wt, W, W, cwd=ref)
        run_ref("column_major_tile_w", ref / wt, W, S, S, cwd=ref)
        uh2, ut2, uw2, _ = oracle_untiled_files(orc / wt, n, eh, et, w)
        oracle_tiled_files(orc / wt, n, eh, et, W, S, w)
        for fmt, off in (("untiled", False), (f"coo-tiled-{W}", True),
                         (f"col-{S}-coo-tiled-{W}", True)):
            res["mismatch"][f"weighted {fmt}"] = same_bytes(files_of(ref / f"{wt}_{fmt}", off),
                                                            files_of(orc / f"{wt}_{fmt}", off))
            res["digest"][f"weighted {fmt}"] = digest(files_of(ref / f"{wt}_{fmt}", off))
    return res


TINY = dict(scale=10, nverts=1024, seed=7, samples=8192, W=64, S=4, wseed=9)
TINY_ODD = dict(scale=10, nverts=1000, seed=11, samples=6000, W=100, S=3, wseed=13)


def tiny_edges(spec):
    return O.rmat_edges(spec["scale"], spec["nverts"], spec["seed"], spec["samples"])
