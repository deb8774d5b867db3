"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built libgphi.so;
everything else runs on the CPU (oracle, host logic, library symbols)."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libgphi.so")


@pytest.fixture(scope="session")
def c1():
    import oracle as O
    return O.config_graph("c1")


@pytest.fixture(scope="session")
def c3():
    import oracle as O
    return O.config_graph("c1", weight_seed=3)


@pytest.fixture(scope="session")
def ctx():
    import bfs_simd_benchmarks_amd as G
    with G.Context(0) as c:
        yield c
