"""Build the native libraries of sparkalog_amd in-tree.

Produces, under sparkalog_amd/_lib/:
  libsparkalog_hip.so       gfx950 kernels + C ABI + device executor (hipcc)
  libsparkalog_workload.so  seeded workload generators (g++, host only)

Run as `python -m sparkalog_amd.build`. Every call compiles from source.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_lib"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["ops.hip", "segsort.hip", "xindex.hip", "compact.hip", "index.hip", "ingest.hip", "capi.hip", "executor.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed-fp32 instruction selection stays off for every HIP translation unit
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip() -> Path:
    OUT.mkdir(exist_ok=True)
    objs = []
    obj_dir = OUT / "obj"
    obj_dir.mkdir(exist_ok=True)
    for src in HIP_SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    target = OUT / "libsparkalog_hip.so"
    tmp = OUT / "libsparkalog_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    tmp.replace(target)
    return target


def build_workload() -> Path:
    OUT.mkdir(exist_ok=True)
    target = OUT / "libsparkalog_workload.so"
    tmp = OUT / "libsparkalog_workload.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-shared", "-fPIC", f"-I{INCLUDE}",
          str(CSRC / "workload.cpp"), "-o", str(tmp), "-lpthread"])
    tmp.replace(target)
    return target


def build_all() -> None:
    build_workload()
    build_hip()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
