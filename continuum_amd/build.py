"""Builds libcnq.so (HIP for gfx950 + the C++ host builder) in-tree, and the
oracle's C++ library that the tests check against.

Run as ``python -m continuum_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "continuum_amd"
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcnq.so"
ORACLE = ROOT / "oracle"
ORACLE_LIB = ORACLE / "liboracle.so"

ARCH = os.environ.get("CNQ_OFFLOAD_ARCH", "gfx950")

# No fused multiply-adds anywhere (results must match the oracle bit for bit);
# packed FP32 instructions off on the device.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]

HIP_SOURCES = ["cnq_api.hip", "cnq_cone.hip", "cnq_primary.hip", "cnq_refit.hip", "cnq_cache.hip", "cnq_server.hip",
               "cnq_vcache.hip", "cnq_sky.hip"]
CXX_SOURCES = ["bvh_build.cpp", "cache_host.cpp", "vcache_host.cpp"]
ORACLE_SOURCES = ["cone_oracle.cpp", "bvh_oracle.cpp", "cone3_oracle.cpp", "vcache_oracle.cpp"]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_library(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        jobs.append(["g++", *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["g++", *CXX_FLAGS, "-shared", "-pthread", "-o", str(tmp), *[str(ORACLE / s) for s in ORACLE_SOURCES]])
    os.replace(tmp, ORACLE_LIB)
    if verbose:
        print(f"built {ORACLE_LIB}")
    return ORACLE_LIB


def build_all(verbose: bool = False) -> None:
    build_oracle(verbose)
    build_library(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
