"""The autotuner's W bounds restate tools/tile-size-computer/tile-size-computer.cpp:21-56:
checked against that tool compiled from its source (when built) and against
values worked out by hand from its loops."""
import subprocess
from pathlib import Path

import pytest

from bfs_simd_benchmarks_amd.autotune import candidate_widths, tile_size_bounds

TOOL = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "tile_size_computer"

CASES = [(33554432, 1175673973), (65536, 1000000), (1632803, 30622564), (1000, 5000),
         (100, 100), (4096, 40000000), (3000000, 10), (2 ** 31, 4000000000), (77777, 777777)]


def test_bounds_by_hand():
    # rmat25: SIDE^2 <= 2^32-1 first at W=1024; M/SIDE^2 >= 8 first at 4096 (17);
    # SIDE <= 64 first at 524288, so max = 262144
    assert tile_size_bounds(33554432, 1175673973) == (4096, 262144)
    # rmat16: W=64 gives 1024^2 tiles; avg 0 <8 until W=256 (16); SIDE<=64 at 1024 -> 512
    assert tile_size_bounds(65536, 1000000) == (256, 512)


def test_candidates_fit_32bit_tile_keys():
    for n, m in CASES:
        ws = candidate_widths(n, m)
        assert ws == sorted(set(ws))
        for w in ws:
            side = n // w + (n % w > 0)
            assert side * side <= 1 << 32
    assert candidate_widths(33554432, 1175673973) == [512, 1024, 2048, 4096, 8192, 16384]


@pytest.mark.skipif(not TOOL.exists(), reason="reference tool not built (no reference tree)")
@pytest.mark.parametrize("n,m", CASES)
def test_bounds_match_reference_tool(n, m):
    out = subprocess.run([str(TOOL), str(n), str(m)], capture_output=True, text=True, timeout=30).stdout
    lo = int(out.split("min_tile_size:")[1].split(",")[0])
    hi = int(out.split("max_tile_size:")[1])
    assert tile_size_bounds(n, m) == (lo, hi)
