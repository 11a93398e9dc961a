"""Compile libcmg_hip.so (every HIP/C++ source under csrc/) for gfx950 with hipcc.

Usage: python -m cybermeshgenerator_amd.build   (or cybermeshgenerator_amd.build.build())

Flags:
  --offload-arch=gfx950            MI355X only
  -ffp-contract=off                no FMA contraction: the naive FP64 determinants, plane
                                   heights, cell centres and FP64 distances must round
                                   exactly as the reference's x86-64 build; FMAs that are
                                   wanted are written explicitly (__builtin_fma[f])
  -target-feature -packed-fp32-ops packed FP32 instructions are kept out of every kernel
  -munsafe-fp-atomics              hardware float atomics for the SDF adjoint scatter
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcmg_hip.so")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip", *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link libcmg_hip.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) +
                  glob.glob(os.path.join(CSRC, "*.cpp")))
    workers = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_native_tests()
    if verbose:
        print(f"built {LIB}")
    return LIB


def _build_native_tests() -> None:
    """The C-ABI smoke (C) and the C++ API test, linked against libcmg_hip.so."""
    tdir = os.path.join(ROOT, "tests", "native")
    if not os.path.isdir(tdir):
        return
    inc = ["-I" + os.path.join(ROOT, "include")]
    link = ["-L" + HERE, "-lcmg_hip", "-Wl,-rpath,$ORIGIN/../../cybermeshgenerator_amd"]
    jobs = [["gcc", "-std=c11", "-O1", *inc, os.path.join(tdir, "smoke_c_abi.c"), "-o",
             os.path.join(tdir, "smoke_c_abi"), *link],
            ["g++", "-std=c++20", "-O1", *inc, os.path.join(tdir, "test_cpp_api.cpp"), "-o",
             os.path.join(tdir, "test_cpp_api"), *link]]
    # host build of the register-only exact stages (device-and-host source)
    jobs.append([HIPCC, "-x", "hip", "-O2", "-std=c++20", "-ffp-contract=off",
                 "--offload-arch=gfx950", "-I" + CSRC, *inc,
                 os.path.join(tdir, "register_stages.cpp"), "-o",
                 os.path.join(tdir, "register_stages")])
    for cmd in jobs:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"native test build failed:\n{' '.join(cmd)}\n{r.stderr}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
