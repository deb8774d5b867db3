"""bench.py's host logic: the --gpus contract and the traffic figure it reads
from the committed rocprofv3 summary."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stdout


def test_traffic_comes_from_the_committed_profile():
    import bench
    b, src, mean_ms = bench.measured_traffic("c4", 2048, 32)
    assert src == "profiles/bfs_dense_kernel_rmat25-ef36.md"
    # FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, in bytes: a few GB per launch
    assert 1e9 < b < 2e10
    assert 0.1 < mean_ms < 10
    assert bench.measured_traffic("c4", 4096, 16) == (None, None, None)   # another tiling: no figure
