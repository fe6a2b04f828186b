"""Build the gfx950 shared library of the transform path, in tree.

    python accelerate-fourier_amd/build.py

compiles csrc/*.hip and csrc/*.cpp with hipcc for gfx950 (objects in
csrc/_build/), links accelerate-fourier_amd/libaccelerate_fourier_amd.so and
then builds tests/c_binding/replay (the C replay of the reference-side
binding) against it.  Every call recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libaccelerate_fourier_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["fft_kernels.hip", "fast_pass.hip", "fast_ext.hip", "decomp.hip", "karatsuba.hip", "abi.cpp", "planner.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # no packed FP32 instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        # keep warnings visible but do not fail (the host pass warns about the fp32 feature)
        msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if msg.strip():
            sys.stderr.write(msg + "\n")
    return obj


REPLAY_SRC = os.path.join(REPO, "tests", "c_binding", "replay.c")
REPLAY_BIN = os.path.join(REPO, "tests", "c_binding", "replay")


def build_replay(verbose: bool = True) -> str:
    """The C program that replays the reference-side binding's call sequence
    against the library (a GPU test runs it): plain C on the HIP runtime's C
    API, built with gcc, linked to the library by relative rpath."""
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
           "-I", os.path.join(REPO, "include"), REPLAY_SRC, "-o", REPLAY_BIN + ".tmp", "-L", HERE,
           "-laccelerate_fourier_amd", "-L", "/opt/rocm/lib", "-lamdhip64", "-lm",
           "-Wl,-rpath,$ORIGIN/../../accelerate-fourier_amd", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed on replay.c:\n{r.stdout}\n{r.stderr}")
    os.replace(REPLAY_BIN + ".tmp", REPLAY_BIN)
    if verbose:
        print(f"built {REPLAY_BIN}")
    return REPLAY_BIN


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_replay(verbose)
    return LIB


if __name__ == "__main__":
    build()
