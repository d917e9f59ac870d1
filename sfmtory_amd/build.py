"""Build the native libraries in-tree.

  sfmtory_amd/lib/libsfmba.so     HIP engine for gfx950 (hipcc)
  sfmtory_amd/lib/libsfmscene.so  synthetic scene generator (g++)

Run as ``python -m sfmtory_amd.build``.  No CUDA, no cmake: one hipcc line
per translation unit, then a shared link.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = PKG / "_build"

HIP_SOURCES = ["engine_host.hip", "engine_kernels.hip", "dense_kernels.hip", "intrinsics_kernels.hip",
               "window_kernels.hip", "window_kernels_512.hip", "window_kernels_256.hip", "pcg_kernels.hip",
               "noalloc.hip"]

# host-only translation units (no device code): g++, linked into libsfmba
HOST_SOURCES = ["colmap_io.cpp", "scope.cpp"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
              "-I/opt/rocm/include"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 math stays off: the FP32 kernels use scalar forms only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{proc.stdout}")
    # surface warnings other than the host pass's unknown-feature notice
    lines = [l for l in proc.stdout.splitlines() if l.strip() and "packed-fp32-ops" not in l
             and "not a recognized feature" not in l]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def build_engine() -> Path:
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    hipcc = _hipcc()
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = OBJ / (Path(src).stem + ".o")
        objs.append(obj)
        jobs.append([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = OBJ / (Path(src).stem + ".o")
        objs.append(obj)
        jobs.append(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, min(8, os.cpu_count() or 1, len(jobs)))) as ex:
        list(ex.map(_run, jobs))
    out = LIB / "libsfmba.so"
    tmp = LIB / "libsfmba.so.tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_scenegen() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libsfmscene.so"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(out), str(CSRC / "scenegen.cpp")])
    return out


def build_all() -> None:
    build_scenegen()
    build_engine()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
