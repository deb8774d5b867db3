"""Write the golden fixtures of this directory from the reference programs.

Runs in the build container only (needs oracle/_ref, i.e. the reference
sources compiled by oracle/Makefile).  For each input graph it runs the
reference tools and apps on files written from the same seeded edge list and
stores, as data:
  reference_outputs.json : sha256 of every reference-written file set
                           (_untiled, _degreeReordered, _reorder,
                           _coo-tiled-W, _col-S-coo-tiled-W, weighted forms)
                           and bellman-ford_simd's printed dist[0..19]
  <graph>_ref_levels0.npz: hop levels from vertex 0 written by apps/sssp_naive
                           (-DONEDEBUG distances.txt)
  <graph>_ref_rank1.npz  : ranks after one iteration from rank=1 written by
                           apps/pageRank_naive (-DONEDEBUG ranks.txt, "%lf")
Usage: python tests/golden/make_golden.py
"""
import json
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))

import oracle as O  # noqa: E402
import refrun as R  # noqa: E402


def graphs():
    for name, spec in (("tiny", R.TINY), ("tiny_odd", R.TINY_ODD)):
        h, t = R.tiny_edges(spec)
        yield name, spec, spec["nverts"], h, t, O.weights(len(h), spec["wseed"])
    c = O.CONFIGS["c1"]
    n, h, t = O.config_edges("c1")
    spec = dict(config="c1", W=c["W"], S=c["S"], wseed=3)
    yield "c1", spec, n, h, t, O.weights(len(h), 3)


def main():
    assert R.have_ref("csr_tiling", "sssp_naive_dbg"), "build oracle/_ref first (make -C oracle ref)"
    out = {}
    for name, spec, n, h, t, w in graphs():
        with tempfile.TemporaryDirectory() as d:
            res = R.pipeline(Path(d), n, h, t, spec["W"], spec["S"], w=w, tag=name)
        bad = {k: v for k, v in res["mismatch"].items() if v}
        assert not bad, f"{name}: oracle writers differ from the reference tools: {bad}"
        out[name] = {"spec": {k: v for k, v in spec.items()}, "n": int(n), "m": int(len(h)),
                     "digests": res["digest"], "bf_dist_0_19": res["ref_bf20"].tolist()}
        np.savez_compressed(HERE / f"{name}_ref_levels0.npz", levels=res["ref_levels0"])
        np.savez_compressed(HERE / f"{name}_ref_rank1.npz", rank=res["ref_rank1"])
        print(name, "ok", out[name]["m"], "edges")
    (HERE / "reference_outputs.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
