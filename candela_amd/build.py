"""Builds libcandela_rt.so in-tree (HIP for gfx950 + C++17 host code).

    python -m candela_amd.build            # library only
    python -m candela_amd.build --oracle   # library + the test oracle (oracle/)

Host translation units are compiled by g++ (SSE2 scalar FP32, no FMA
contraction, like the reference constructor's own build); the HIP kernels by
hipcc for gfx950 with FMA contraction off and packed-FP32 instructions
disabled.  Objects go to candela_amd/_build/, the library next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcandela_rt.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("CRT_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["bvh_build.cpp", "runtime.cpp", "wide_build.cpp", "scene_cache.cpp"]
HIP_SOURCES = ["trace.hip", "persistent.hip", "stack.hip", "spec.hip", "batched.hip", "shade.hip", "wide.hip", "attr.hip", "path.hip",
               "raygen.hip", "gpu_build.hip", "wide_gpu.hip", "noalloc.hip", "spec_bvh2.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
HOST_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wextra",
    "-D__HIP_PLATFORM_AMD__",
    f"-I{ROCM}/include",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _sources(names: list[str]) -> list[str]:
    return [n for n in names if os.path.exists(os.path.join(CSRC, n))]


def build_library(verbose: bool = True) -> str:
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    cxx = shutil.which("g++") or "g++"
    os.makedirs(BUILD, exist_ok=True)
    inc = [f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
    jobs = []
    for src in _sources(HOST_SOURCES):
        jobs.append([cxx, *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, src + ".o")])
    for src in _sources(HIP_SOURCES):
        jobs.append([hipcc, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, src + ".o")])
    objs = [j[-1] for j in jobs]
    # translation units are independent: compile them side by side
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1), 8))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, j) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-j4"])


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true", help="also build oracle/ (test checker)")
    args = ap.parse_args(argv)
    build_library()
    if args.oracle:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
