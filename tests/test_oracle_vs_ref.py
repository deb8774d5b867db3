"""The oracle against the reference programs themselves, live: runs only where
oracle/_ref was compiled from the reference sources (the build container)."""
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import refrun as R

pytestmark = pytest.mark.skipif(
    not R.have_ref("csr_tiling", "csr_tiling_w", "coo_tiling", "coo_tiling_w", "column_major_tile",
                   "column_major_tile_w", "degree_reorder", "bfs_remap", "sssp_naive_dbg",
                   "pagerank_naive_dbg", "bellman_ford"),
    reason="oracle/_ref not built (needs the reference tree)")


@pytest.mark.parametrize("spec", [R.TINY, R.TINY_ODD], ids=["tiny", "tiny_odd"])
def test_tools_and_apps_agree_with_oracle(spec):
    h, t = R.tiny_edges(spec)
    n = spec["nverts"]
    w = O.weights(len(h), spec["wseed"])
    with tempfile.TemporaryDirectory() as d:
        res = R.pipeline(Path(d), n, h, t, spec["W"], spec["S"], w=w)
    bad = {k: v for k, v in res["mismatch"].items() if v}
    assert not bad, bad
    g = res["graph"]
    assert np.array_equal(O.bfs_levels(n, g.deg, g.cols, 0), res["ref_levels0"])
    assert np.array_equal(O.bellman_ford_hops(n, g.rh, g.rt, 0, 99)[:20], res["ref_bf20"])
