"""Builds libcsstitch.so (HIP kernels for gfx950 + host runtime + C ABI).

    python crossstage-stitch_amd/build.py [--jobs N] [--verbose]

Every .hip and .cpp file under csrc/ is compiled with hipcc into _build/ and
linked into cs_stitch/libcsstitch.so, next to the Python binding that loads it.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "cs_stitch" / "libcsstitch.so"
BUILD = PKG.parent / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CS_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    # Every multiply and add is rounded separately, as in the CPU semantics the
    # stages must reproduce bit for bit.
    "-ffp-contract=off",
    "-I", str(CSRC), "-I", str(PKG.parent / "include"),
]
HIP_ONLY = [
    f"--offload-arch={ARCH}",
    # Packed-FP32 instruction forms are kept out of the generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc",
]


def sources():
    return sorted(list(CSRC.rglob("*.hip")) + list(CSRC.rglob("*.cpp")))


def compile_one(src: Path, verbose: bool) -> Path:
    rel = src.relative_to(CSRC)
    obj = BUILD / (str(rel).replace("/", "__") + ".o")
    cmd = [HIPCC, *COMMON]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *HIP_ONLY]
    else:
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include"]
    cmd += ["-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip() and verbose:
        print(r.stderr, file=sys.stderr)
    return obj


def build(jobs: int = 0, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_harness(verbose)
    return OUT


HARNESS_SRC = PKG / "harness" / "stitch_harness.cpp"
HARNESS = PKG / "bin" / "stitch_harness"
STRESS_SRC = PKG / "harness" / "stitch_stress.cpp"
STRESS = PKG / "bin" / "stitch_stress"


def build_harness(verbose: bool = False) -> Path:
    """The stitch_harness and stitch_stress command-line drivers, linked
    against libcsstitch.so (found next to them through the run path)."""
    HARNESS.parent.mkdir(exist_ok=True)
    for src, exe in ((HARNESS_SRC, HARNESS), (STRESS_SRC, STRESS)):
        cmd = [HIPCC, *COMMON, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", str(src),
               "-o", str(exe), f"-L{OUT.parent}", "-lcsstitch", "-Wl,-rpath,$ORIGIN/../cs_stitch",
               "-L/opt/rocm/lib", "-lamdhip64", "-lpthread"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{exe.name} build failed\n{r.stdout}\n{r.stderr}")
    return HARNESS


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
