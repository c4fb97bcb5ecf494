"""Build libmeinv.so in-tree for gfx950 with hipcc (no JIT cache, no torch)."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = CSRC / "libmeinv.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions stay off: hipcc is told the target lacks them
# FP32 division and sqrt use the hardware reciprocal/rsq (about 1 ulp) instead
# of the correctly rounded sequences, and FP32 denormals flush to zero, so a
# reciprocal is one instruction instead of a range-scaled sequence of five;
# FP64 arithmetic is unaffected
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero",
         "-Wno-unused-result"]


HOST = HERE / "host"
HOST_BIN = HOST / "meinv"


def build_host(verbose: bool = False) -> Path:
    """The MEINV-style host program (plain C++ over the C ABI), linked
    against the in-tree libmeinv.so."""
    cxx = os.environ.get("CXX", "g++")
    srcs = [str(HOST / f) for f in ("meinv_main.cpp", "fits.cpp", "control.cpp")]
    tmp = HOST_BIN.with_suffix(".tmp")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-o", str(tmp), *srcs, f"-L{CSRC}", "-lmeinv",
           "-Wl,-rpath,$ORIGIN/../csrc", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the meinv host program failed")
    os.replace(tmp, HOST_BIN)
    return HOST_BIN


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    lib = build_lib(verbose, extra)
    build_host(verbose)
    return lib


def build_lib(verbose: bool = False, extra: list[str] | None = None) -> Path:
    srcs = [str(CSRC / "me_api.hip")]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *(extra or []), "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmeinv.so")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    extra = sys.argv[1:]
    print(build(verbose=True, extra=extra))
