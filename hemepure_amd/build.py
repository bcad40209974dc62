"""Builds the native library hemepure_amd/libhemepure_amd.so with hipcc for gfx950.

Every source is compiled with -ffp-contract=off: the double-precision path
reproduces the reference's rounding operation by operation, so no multiply-add
may be fused.  Packed FP32 instructions are disabled for device code.
Run as ``python -m hemepure_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libhemepure_amd.so"

SOURCES = ["capi.hip", "batch.hip", "server.hip", "lbm_fp64.hip", "lbm_fp32.hip", "lattice.cpp", "geometry_gen.cpp", "gmy_reader.cpp",
           "extraction.cpp"]
# Host-only sources built with the system C++ compiler: their complex
# arithmetic and transcendental calls must round as the reference's build of
# the same expressions does.
GXX_SOURCES = ["iolet_math.cpp"]
ARCH = os.environ.get("HLB_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build hemepure_amd")


def common_flags() -> list[str]:
    return [
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-command-line-argument",
        f"-I{PKG.parent / 'include'}",
    ]


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [hipcc(), *common_flags(), "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def _compile_gxx(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", f"-I{PKG.parent / 'include'}",
           "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{cxx} failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES)) + list(ex.map(_compile_gxx, GXX_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-lz", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
