"""CPU checks of the product boundary: the C-ABI library loads and exports every
entry point include/gphi.h declares; without a GPU it fails loudly."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "gphi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gphi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("gphi_open", "gphi_bfs", "gphi_sssp", "gphi_pagerank", "gphi_bellman_ford",
                 "gphi_graph_load_text", "gphi_graph_from_desc"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import bfs_simd_benchmarks_amd as G
    assert G.LIB_PATH.exists(), "libgphi.so not built"
    lib = ctypes.CDLL(str(G.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    import bfs_simd_benchmarks_amd as G
    assert b"gfx950" in G.load_library().gphi_version()


def test_open_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import bfs_simd_benchmarks_amd as G
    with pytest.raises(G.GphiError) as ei:
        G.Context(0)
    assert ei.value.code == -5
