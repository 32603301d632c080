"""Builds libnarval_gpu.so (HIP kernels for gfx950 + C++ host runtime) in-tree.

Every source is compiled by hipcc with:
  * --offload-arch=gfx950 (CDNA4 only),
  * -ffp-contract=off: each multiply and add rounds separately, so device
    arithmetic matches a plain C++ evaluation of the same formula,
  * -Xclang -target-feature -Xclang -packed-fp32-ops: no packed-FP32 VALU ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libnarval_gpu.so"
ROOT = PKG.parent
FACADE_SRC = ROOT / "tests" / "cpp" / "facade_check.cpp"
FACADE_BIN = ROOT / "tests" / "cpp" / "facade_check"
SOURCES = ["render.hip", "wavefront.hip", "batch.hip", "guide.hip", "api.cpp", "bvh_build.cpp", "generators.cpp",
           "scene_io.cpp", "image_io.cpp"]
HEADERS = ["ne_math.h", "ne_scene.h", "ne_trace.h", "ne_launch.h", "ne_vmf.h", "ne_guide.h", "bvh_build.h"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build the HIP library)")


COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str) -> Path:
    obj = BUILD / (src.replace(".", "_") + ".o")
    cmd = [hipcc(), "-c", str(CSRC / src), "-o", str(obj), *COMMON]
    cmd += ["-x", "hip", "--offload-arch=gfx950"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"compilation failed: {src}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-o", str(tmp), *map(str, objs), "--offload-arch=gfx950", "-pthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    build_facade_check()
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_facade_check() -> Path:
    """g++ build of the C++ facade check program (include/ne_gpu.hpp over the
    C ABI), linked against the in-tree library."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(FACADE_SRC), "-o", str(FACADE_BIN),
           f"-L{PKG}", "-lnarval_gpu", "-Wl,-rpath,$ORIGIN/../../narvalengine_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("facade check build failed")
    return FACADE_BIN


if __name__ == "__main__":
    build(verbose=True)
