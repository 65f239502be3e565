"""Builds the native library admm_legalization_amd/_lib/libskyline_amd.so.

HIP sources are compiled for gfx950 only.  Floating-point flags: no FMA
contraction (the legalizer's bit-parity contract) and no packed-FP32
instruction selection.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libskyline_amd.so")
CMD = os.path.join(LIBDIR, "skyline_cmd")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
             "-Wno-unused-function"]

# (source, compiler) pairs; header-only .inc.hip files are pulled in by includes
UNITS = [
    ("lg_device.hip", "hip"),
    ("lg_grid.hip", "hip"),
    ("lg_host.cpp", "cxx"),
    ("dens_device.hip", "hip"),
    ("dens_host.cpp", "cxx"),
    ("wl_device.hip", "hip"),
    ("gp_device.hip", "hip"),
    ("gp_host.cpp", "cxx"),
    ("lefdef.cpp", "cxx"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src, kind in UNITS:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        if kind == "hip":
            _run([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj], verbose)
        else:
            _run(["g++", *CXX_FLAGS, "-c", path, "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp, *objs,
          f"-Wl,-rpath,{ROCM}/lib", f"-Wl,-rpath,{ROCM}/lib/llvm/lib"], verbose)
    os.replace(tmp, LIB)
    # the command driver (skyline_cmd: read_design / legalize / write_result)
    _run(["g++", "-O2", "-std=c++17", os.path.join(CSRC, "skyline_cmd.cpp"), "-L", LIBDIR, "-lskyline_amd",
          f"-Wl,-rpath,{LIBDIR}", "-o", CMD + ".tmp"], verbose)
    os.replace(CMD + ".tmp", CMD)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
