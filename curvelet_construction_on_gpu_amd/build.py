"""Build the native parts in-tree.

    python -m curvelet_construction_on_gpu_amd.build        # everything
    python curvelet_construction_on_gpu_amd/build.py --only lib

Products (git-ignored, shipped with the tree):
    curvelet_construction_on_gpu_amd/lib/libcurvelet_gpu.so   HIP pipeline + C ABI
    curvelet_construction_on_gpu_amd/bin/CURVELET             drop-in command line
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib" / "libcurvelet_gpu.so"
CLI = PKG / "bin" / "CURVELET"
ARCH = os.environ.get("CVG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

# Device code: no FP contraction (the reference's x86-64 build has no FMA, and
# bit-exact decisions need the same rounding points), no fast-math, packed
# FP32 disabled.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> Path:
    LIB.parent.mkdir(parents=True, exist_ok=True)
    objs = []
    objdir = PKG / "lib" / "_obj"
    objdir.mkdir(parents=True, exist_ok=True)
    for src in ["cvg_pipeline.hip"]:
        obj = objdir / (Path(src).stem + ".o")
        run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in ["cvg_io.cpp", "cvg_synth.cpp", "cvg_server.cpp"]:
        obj = objdir / (Path(src).stem + ".o")
        run([CXX, "-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    run([HIPCC, "-shared", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> Path:
    CLI.parent.mkdir(parents=True, exist_ok=True)
    run([CXX, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(CSRC / "cvg_cli.cpp"),
         f"-L{LIB.parent}", "-lcurvelet_gpu", "-Wl,-rpath,$ORIGIN/../lib", "-o", str(CLI)])
    return CLI


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["lib", "cli"], default=None)
    args = ap.parse_args(argv)
    if args.only in (None, "lib"):
        build_lib()
    if args.only in (None, "cli"):
        build_cli()
    return 0


if __name__ == "__main__":
    sys.exit(main())
