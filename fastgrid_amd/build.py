"""Builds libfgmaps.so (host front end + gfx950 HIP kernels + C ABI) and the
fastgrid4 / fgsynth executables, in-tree under fastgrid_amd/.

    python -m fastgrid_amd.build          # or: python fastgrid_amd/build.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("FG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HOST_SOURCES = ["params.cpp", "loader.cpp", "writers.cpp", "synth.cpp", "mapset.cpp", "capi_host.cpp"]
DEVICE_SOURCES = ["capi_device.hip"]
HEADERS = ["host.hpp", "capi_common.hpp", "kernels.hpp", "kernels.hip", "kernels_plain.hip", "kernels_interp.hip"]

CXXFLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-I" + INCLUDE]
HIPFLAGS = [
    "-std=c++17", "-O3", "-fPIC", "--offload-arch=" + ARCH,
    # packed FP32 (v_pk_*_f32) stays off: every FP32 op is a plain VALU op
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-fno-fast-math", "-Wno-unused-result", "-I" + INCLUDE,
]

LIB = os.path.join(HERE, "libfgmaps.so")
BIN = os.path.join(HERE, "bin")


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + p.stdout)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    # hipcc's host pass warns about the packed-fp32 target feature; drop it
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out.strip():
        sys.stderr.write(out + "\n")


def build(verbose: bool = False) -> str:
    """Compiles everything; returns the path of libfgmaps.so."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        objs.append(obj)
        jobs.append([CXX, *CXXFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    rpath = "-Wl,-rpath,$ORIGIN/.."
    _run([CXX, "-std=c++17", "-O2", "-I" + INCLUDE, "-I" + CSRC, os.path.join(CSRC, "fastgrid4_main.cpp"),
          "-o", os.path.join(BIN, "fastgrid4"), "-L" + HERE, "-lfgmaps", rpath, "-lpthread"])
    _run([CXX, "-std=c++17", "-O2", "-I" + INCLUDE, os.path.join(CSRC, "fgsynth_main.cpp"),
          "-o", os.path.join(BIN, "fgsynth"), "-L" + HERE, "-lfgmaps", rpath])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
