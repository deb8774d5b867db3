"""Build libgphi.so (HIP kernels for gfx950 + host C++) and the drop-in CLIs in-tree.

Run as ``python -m bfs_simd_benchmarks_amd.build`` or through
``__graft_entry__.build()``.  Objects go to ``bfs_simd_benchmarks_amd/_build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
CLI = PKG / "cli"
BUILD = PKG / "_build"
LIB = PKG / "libgphi.so"
BIN = PKG / "bin"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # keep hipcc from emitting packed FP32 ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wextra",
              "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}"]
LINK_FLAGS = [f"-L{ROCM / 'lib'}", "-lamdhip64", "-fopenmp", "-pthread", f"-Wl,-rpath,{ROCM / 'lib'}"]

HIP_SOURCES = ["device.hip", "gen.hip", "batch.hip", "noalloc.hip", "sssp_batch.hip"]
HOST_SOURCES = ["api.cpp", "formats.cpp", "server.cpp", "cache.cpp"]
CLIS = {"bfs": "bfs_main.cpp", "sssp": "sssp_main.cpp", "page_rank": "pagerank_main.cpp",
        "bellman-ford": "bellman_ford_main.cpp"}


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd[:3])} ...")
    # hipcc's host pass warns that it does not know the packed-fp32 feature
    noise = [ln for ln in (r.stdout + r.stderr).splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "warning generated" not in ln]
    if verbose and noise:
        print("\n".join(noise))


def build(verbose: bool = False, jobs: int = 8) -> Path:
    """Compile every HIP/C++ source of the package; returns the library path."""
    BUILD.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.h")) + [PKG.parent / "include" / "gphi.h"]
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        procs.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        procs.append(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    del headers
    # compile in parallel
    running = []
    for cmd in procs:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                              text=True)))
    failed = []
    for cmd, p in running:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((cmd, out))
        elif verbose:
            noise = [ln for ln in out.splitlines() if ln.strip() and "packed-fp32-ops" not in ln
                     and "warning generated" not in ln]
            if noise:
                print("\n".join(noise))
    if failed:
        for cmd, out in failed:
            sys.stderr.write(" ".join(cmd) + "\n" + out + "\n")
        raise RuntimeError("libgphi compile failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-o", str(tmp), *map(str, objs), *LINK_FLAGS], verbose)
    os.replace(tmp, LIB)
    # drop-in CLIs
    for name, src in CLIS.items():
        path = CLI / src
        if not path.exists():
            continue
        _run(["g++", "-O2", "-std=c++17", "-fopenmp", f"-I{PKG.parent / 'include'}", f"-I{CLI}", str(path),
              "-o", str(BIN / name), f"-L{PKG}", "-lgphi", f"-Wl,-rpath,$ORIGIN/..", *LINK_FLAGS],
             verbose)
    return LIB


def build_oracle(verbose: bool = False) -> None:
    """Build the checker: oracle/liboracle.so, and oracle/_ref when the reference is present."""
    root = PKG.parent / "oracle"
    _run(["make", "-s", "-C", str(root), "lib"], verbose)
    if Path(os.environ.get("GPHI_REFERENCE", "/root/reference")).is_dir() and shutil.which("g++"):
        _run(["make", "-s", "-j8", "-C", str(root), "ref"], verbose)


if __name__ == "__main__":
    build(verbose=True)
    build_oracle(verbose=True)
    print(f"built {LIB}")
