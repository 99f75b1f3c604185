"""Builds the MI355X smoke library: every csrc/*.hip compiled by hipcc for
gfx950 (in parallel), linked into lib/libvisual-simulation-of-smoke.so, and the
two C-ABI callers: the HIP counterpart of the reference benchmark CLI
(cli/benchmark.hip) and the reference visualizer's scene driver
(cli/scene_plume.hip)."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libvisual-simulation-of-smoke.so"
BENCH_SRC = PKG / "cli" / "benchmark.hip"
BENCH_BIN = PKG / "lib" / "visual-simulation-of-smoke-benchmark"
SCENE_SRC = PKG / "cli" / "scene_plume.hip"
SCENE_BIN = PKG / "lib" / "visual-simulation-of-smoke-scene"

ARCH = os.environ.get("SMOKE_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-fvisibility=hidden",
    # no packed FP32 math on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", f"-I{REPO / 'include'}", f"-I{CSRC}",
]


def _compile(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for src, out in ((BENCH_SRC, BENCH_BIN), (SCENE_SRC, SCENE_BIN)):
        if not src.exists():
            continue
        cmd = [HIPCC, *FLAGS, str(src), "-o", str(out), f"-L{LIB.parent}", "-lvisual-simulation-of-smoke", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
