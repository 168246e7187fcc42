"""Builds the in-tree native library libmcold.so for gfx950.

Every source is compiled by hipcc (host .cpp files too, so the shared
physics header sees one compiler), with FMA contraction off so GPU histories
are bit-identical to the CPU restatement, and packed-FP32 disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmcold.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST = os.path.join(HERE, "host")
CLI = os.path.join(HERE, "mc")
SOURCES = ["csrc/mc_xml.cpp", "csrc/mc_problem.cpp", "csrc/mc_api.cpp", "csrc/mc_report.cpp",
           "csrc/mc_engine.hip", "csrc/mc_kernels.hip", "host/simulator.cpp"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-I" + CSRC]
DEVICE = ["--offload-arch=gfx950", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    path = os.path.join(HERE, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp, *objs, "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the `mc <io_dir>` command-line drop-in, linked against the library
    cmd = ["g++", "-O2", "-std=c++17", "-o", CLI + ".tmp", os.path.join(HOST, "main.cpp"),
           os.path.join(HOST, "simulator.cpp"), "-L" + HERE, "-lmcold", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(CLI + ".tmp", CLI)
    if verbose:
        print("built", LIB, "and", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
